"""MI355X (gfx950) FDTD engine for Aether's Yee solver.

The compute path is the HIP library ``libaether_fdtd.so`` behind the C ABI of
``include/aether_fdtd.h``; this package only loads it (``capi.engine()``) and
builds synthetic inputs (``workloads``).  There is no CPU fallback.
"""
from .capi import (EX, EY, EZ, FP16, FP32, FP64, HX, HY, HZ, MODE_CUSTOM, MODE_NORM, Backend, FDTDError, Grid,
                   Material, ModeSpec, SensorReq, engine, run_normal_incidence, run_single_particle)

__all__ = ["EX", "EY", "EZ", "HX", "HY", "HZ", "FP16", "FP32", "FP64", "MODE_CUSTOM", "MODE_NORM", "Backend",
           "FDTDError", "Grid", "Material", "ModeSpec", "SensorReq", "engine", "run_normal_incidence",
           "run_single_particle"]
