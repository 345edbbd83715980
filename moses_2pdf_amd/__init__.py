"""MI355X-native engine for MoSES_2PDF's adNOC_2Pb time-step loop (landslide and inflow
modes): HIP kernels for gfx950 behind the C ABI of include/moses2pdf.h."""
from .engine import (M2P_FP32, M2P_FP64, M2P_INFLOW, M2P_LANDSLIDE, Batch,  # noqa: F401
                     M2PError, RunKernels, Simulation, default_params, footprint, lib)
