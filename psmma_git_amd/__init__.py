"""MI355X-native PSMMA: structured x dense products on gfx950 MFMA.

The compute lives in libpsmma_hip.so (HIP kernels + C ABI, include/psmma.h);
``api`` mirrors the reference's entry points over it.
"""
from . import api  # noqa: F401
from .api import (  # noqa: F401
    PREC_FP64, PREC_SPLIT_BF16, PREC_SPLIT_I8, PsmmaError, psmma_cauchy, psmma_cauchyeig,
    psmma_hankel, psmma_toeplitz, pzmdftl, pzmdftl1,
)

__version__ = "0.1.0"
