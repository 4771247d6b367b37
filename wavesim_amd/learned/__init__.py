"""Learned warm start for the MBS solver (no counterpart in the reference).

``prior``   background-medium field of a plane-wave source (closed form)
``fno``     factorised 3-D Fourier neural operator, forward/backward on the
            project's own HIP kernels (BF16 MFMA GEMMs, FP32 accumulate)
"""
