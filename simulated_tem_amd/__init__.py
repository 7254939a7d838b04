"""simulated_tem_amd: MI355X-native dynamical two-beam imaging (d2bi hot path).

The compute path is libd2bi_gpu.so (hand-written gfx950 HIP kernels behind
the C-ABI of include/d2bi_gpu.h).  This package is the Python host mirror of
the reference's imaging interface (Lib_DynamicalTwoBeamImaging /
Lib_LinkCell3D / gaussianBlurImage); every compute call goes to the HIP
library and raises if it is not built or no device is present.
"""
from .abi import (D2biRecord, D2biPlan, D2biStats, D2BI_MODE_FP64, D2BI_MODE_FAST,  # noqa: F401
                  D2biError)

__all__ = ["D2biRecord", "D2biPlan", "D2biStats", "D2BI_MODE_FP64", "D2BI_MODE_FAST",
           "D2biError"]
