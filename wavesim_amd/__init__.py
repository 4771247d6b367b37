"""MI355X-native modified Born series (MBS) solver.

The MBS loop of the reference wavesim (WaveSimBase.run_algorithm /
RunWaveSimAlgorithmMex) runs on one GPU as hand-written gfx950 HIP kernels in
``libwavesim_mbs.so`` (C ABI: include/wavesim_mbs.h).  This package is the
host side: ``WaveSim`` / ``WaveSimVector`` / ``Source`` mirror the reference
classes, ``MBSSolver`` wraps the handle API, ``PSTD`` mirrors the reference's
time-domain comparison solver (PSTD.m) on the same FFT kernels.
"""
from .simulation import (Medium, SimulationGrid, Source, SimState, WaveSim, WaveSimBase, exec_many,
                         WaveSimVector, efficient_size, symrange)
from .solver import MBSSolver, RhsResult, SolverGroup
from .pstd import PSTD

__all__ = ["Medium", "SimulationGrid", "Source", "SimState", "WaveSim", "WaveSimBase", "exec_many", "SolverGroup",
           "WaveSimVector", "efficient_size", "symrange", "MBSSolver", "RhsResult", "PSTD"]
