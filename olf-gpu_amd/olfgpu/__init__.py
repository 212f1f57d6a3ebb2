"""olfgpu — MI355X-native AC Newton-Raphson and batched N-1 behind the reference's AcSolver / provider surface.

The numerics live in libolfhip.so (HIP, gfx950) behind the C ABI of include/olfhip.h; this package is the host-side
mirror of the reference's Java layer (parameters, the GPU_NEWTON_RAPHSON solver, the batched N-1 solver) plus the
seeded synthetic workloads.
"""
from . import gridgen  # noqa: F401
from ._native import (BatchContext, OlfHipError, Structure, autotune, device_count,  # noqa: F401
                      STATUS_NAMES, LU_FP64, LU_FP32_REFINED)
from .solver import (GpuBatchedSecurityAnalysisSolver, GpuLoadFlowParameters, GpuNewtonRaphson,  # noqa: F401
                     NewtonRaphsonParameters, NAME, PROVIDER_NAME)
