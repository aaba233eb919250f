"""MI355X-native AMM power method for the largest eigenvalue of A A^T.

Hot path: powerMethod_by_exact / _previous / _proposed / _proposed_v2 of
Takuya-Okuyama/eigenvalueComputation_by_AMM, rebuilt as hand-written HIP
kernels for gfx950 behind the C ABI in include/amm.h (libamm.so).
"""
from .api import (  # noqa: F401
    AmmBatch,
    AmmContext,
    BatchResult,
    AmmError,
    RunParams,
    RunResult,
    powerMethod_by_exact,
    powerMethod_by_previous,
    powerMethod_by_proposed,
    powerMethod_by_proposed_v2,
)
