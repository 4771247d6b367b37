"""CPU oracle for the modified Born series (MBS) path.

TEST INFRASTRUCTURE ONLY.  Nothing in the ``wavesim_amd`` package imports this
package; only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline``
leg of ``bench.py`` use it, and only as the checker / the timed CPU path.

``oracle.mbs`` restates the reference MATLAB implementation line by line in
numpy (complex64 and complex128), ``oracle.analytic`` restates the reference's
closed-form homogeneous-medium solution and adds the textbook 2-D/3-D Green's
functions used as physics checks.  Each function cites the reference file and
line it follows and names the check that pins it (see ``tests/test_oracle*.py``).
"""
