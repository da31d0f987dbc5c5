"""MI355X (gfx950) Numerical Flow Iteration rho path.

The hot path of NuFI (paulwilhelmvlasov/NumericalFlowIteration): for every
(x, v) quadrature node, trace the characteristic backward through the stored
electric-field spline history (Stormer-Verlet), evaluate f0 at the foot point
and reduce over v into rho; then the periodic Poisson solve and spline fit
append the next history slice — all on the device, behind the C ABI of
include/nufi_hip.h (libnufi_hip.so).
"""
from .config import (Config, F0, FP32C, FP64, config_1d, config_3d, landau_1d, landau_3d,
                     multimode_1d, twostream_1d, sweep_runs, sweep_config, adjoint_modes)
from ._capi import NufiError, NufiIOError, NufiRangeError, device_count, library_path
from .history import read_history, read_info, write_history
from .scheduler import BatchScheduler, GpuScheduler

__all__ = [
    "Config", "F0", "FP64", "FP32C", "config_1d", "config_3d", "landau_1d", "landau_3d",
    "multimode_1d", "twostream_1d", "sweep_runs", "sweep_config", "adjoint_modes",
    "NufiError", "NufiIOError", "NufiRangeError", "device_count", "library_path", "GpuScheduler",
    "BatchScheduler", "read_history", "read_info", "write_history",
]
