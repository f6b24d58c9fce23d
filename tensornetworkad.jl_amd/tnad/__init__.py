"""tnad: MI355X (gfx950) backend for TensorNetworkAD.jl's CTMRG / TRG hot path.

The compute lives in libtnad.so (HIP kernels behind the C ABI of
include/tnad.h); this package is the host-side mirror of the reference's
Julia API over that ABI.
"""
from .api import (  # noqa: F401
    MIXED_BF16,
    MIXED_FP32,
    Context,
    CtmrgTape,
    DeviceArray,
    Heisenberg,
    Ising,
    OptimResult,
    SquareCTMRGRuntime,
    SquareIPEPS,
    TFIsing,
    ctmrg,
    contract,
    ctmrgstep,
    default_context,
    diaglocalhamiltonian,
    dmodel_tensor,
    double_layer,
    energy,
    energy_and_grad,
    expectationvalue,
    getchi,
    getd,
    gets,
    hamiltonian,
    indexperm_symmetrize,
    mag_tensor,
    magnetisation,
    magofbeta,
    make_opts,
    model_tensor,
    num_grad,
    optimiseipeps,
    rrule_energy,
    rrule_trg,
    time_contract,
    trg,
    trg_and_grad,
    trg_svd,
    trg_svd_back,
)
from ._lib import LIB_PATH, DimensionMismatch, TnadError, load  # noqa: F401
