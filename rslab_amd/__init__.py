"""rslab_amd: MI355X-native numeric factor and solve for rslab's supernodal
left-looking LDL^T (Bunch-Kaufman), behind rslab's LdltSymbolic / LdltSolver /
MixedLdltSolver surface. The numeric work runs in librslab_amd.so (HIP, gfx950).
"""
from .solver import (  # noqa: F401
    CscMatrix,
    DeviceError,
    DimensionMismatch,
    Inertia,
    InvalidInput,
    IoError,
    LdltFactors,
    LdltSolver,
    LdltSymbolic,
    MixedInfo,
    MixedLdltSolver,
    NumericallyRankDeficient,
    OrderingMethod,
    OutOfMemory,
    Precision,
    RslabError,
    SolverSettings,
    ZeroPivotAction,
    amd_order,
    nd_order,
)

__all__ = [
    "CscMatrix",
    "DeviceError",
    "DimensionMismatch",
    "Inertia",
    "InvalidInput",
    "IoError",
    "LdltFactors",
    "LdltSolver",
    "LdltSymbolic",
    "MixedInfo",
    "MixedLdltSolver",
    "NumericallyRankDeficient",
    "OrderingMethod",
    "OutOfMemory",
    "Precision",
    "RslabError",
    "SolverSettings",
    "ZeroPivotAction",
    "amd_order",
    "nd_order",
]
