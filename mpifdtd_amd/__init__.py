"""mpifdtd_amd -- MI355X-native 2-D FDTD solvers (TM / TE) with NTFF.

A drop-in for rennone/mpiFDTD's solvers on one GPU (the UPML ones, with a
fused FP32 path, and the split-field PML and non-standard FDTD ones in FP64):
the C host keeps the reference's field / models / simulator / solver entry
points and calls hand-written gfx950 HIP kernels through a thin C ABI
(include/*.h).
"""
from ._native import (  # noqa: F401
    D_X, D_XY, D_Y, LAYER, MIE_CYLINDER, MORPHO_SCALE, MPI_TE_UPML_2D, MPI_TM_UPML_2D, NO_MODEL,
    NS_TE_2D, NS_TM_2D, TE_2D, TE_UPML_2D, TM_2D, TM_UPML_2D, ZIGZAG, FieldInfo, ModelParams,
    NativeLibraryMissing, LIB_PATH, lib,
)
from .simulator import Simulator, field_info, model_params  # noqa: F401
