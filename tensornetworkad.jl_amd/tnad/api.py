"""Host-side mirror of TensorNetworkAD.jl's API over the HIP C ABI.

Names, argument meaning and error behaviour follow the reference
(src/TensorNetworkAD.jl:6-10 exports): ``ctmrg``, ``ctmrgstep``, ``trg``,
``trg_svd``, ``energy``, ``expectationvalue``, ``optimiseipeps``,
``indexperm_symmetrize``, ``SquareIPEPS``, ``SquareCTMRGRuntime``,
``hamiltonian``, ``model_tensor``, ``num_grad``.  Arrays are NumPy arrays with
Julia's (column-major) index order; they are handed to the library as
Fortran-ordered buffers.  Device-resident inputs (``DeviceArray``) are passed
as device pointers.  All numerical work of the hot path runs in libtnad on the
GPU; the small model builders (Hamiltonians, the Ising tensor) are inputs,
built here exactly as the reference builds them on the host.

Differentiation: the reference attaches Zygote through ``@adjoint``; here
``energy_and_grad`` / ``trg_and_grad`` return the value and gradient of one
fused library call, and ``rrule_energy`` / ``rrule_trg`` give the
ChainRulesCore-style (value, pullback) pair.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import DimensionMismatch, Opts, Stats, TnadError


# ---------------------------------------------------------------------------
# context
# ---------------------------------------------------------------------------
class Context:
    """One GPU, one HIP stream (tnad_create / tnad_destroy)."""

    def __init__(self, device: int = 0):
        self.lib = _lib.load()
        h = C.c_void_p()
        rc = self.lib.tnad_create(device, C.byref(h))
        if rc != 0:
            raise TnadError(rc, f"tnad_create(device={device}) failed (no HIP device?)")
        self.h = h
        self.device = device

    def check(self, rc):
        if rc != 0:
            msg = self.lib.tnad_last_error(self.h).decode()
            if rc == -2:
                raise DimensionMismatch(rc, msg)
            raise TnadError(rc, msg)

    def stats(self) -> dict:
        s = Stats()
        self.check(self.lib.tnad_get_stats(self.h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in Stats._fields_}

    def close(self):
        if getattr(self, "h", None):
            self.lib.tnad_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx: Context | None = None


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


class DeviceArray:
    """A device buffer holding a column-major FP64 array (e.g. from torch: pass
    ``tensor.data_ptr()`` of a contiguous float64 CUDA tensor and the Julia shape)."""

    def __init__(self, ptr: int, shape):
        self.ptr = int(ptr)
        self.shape = tuple(shape)


def _host(x):
    """column-major host copy: ComplexF64 for complex input (interleaved re, im as Julia stores it),
    Float64 otherwise"""
    if np.iscomplexobj(x):
        return np.asfortranarray(np.asarray(x, dtype=np.complex128))
    return np.asfortranarray(np.asarray(x, dtype=np.float64))


def _in(x):
    if isinstance(x, DeviceArray):
        return C.c_void_p(x.ptr), x
    a = _host(x)
    return C.c_void_p(a.ctypes.data), a


def _real_only(x, what):
    if np.iscomplexobj(x):
        raise TypeError(f"{what}: this entry point is Float64 only; complex bulks go through energy / "
                        "energy_and_grad (tnad_energy_z)")


def _out(shape, dtype=np.float64):
    a = np.zeros(shape, dtype=dtype, order="F")
    return C.c_void_p(a.ctypes.data), a


def make_opts(precision="fp64", switch_tol=0.0, bwd_precision="fp64", eig_block=0, eig_max_iters=200,
              eig_warm=False, eig_warm_iters=200, eig_tol=1e-13, bwd_last=0) -> Opts:
    """Run options (tnad_opts).  Defaults give the reference semantics."""
    o = Opts()
    _lib.load().tnad_opts_default(C.byref(o))
    o.precision = _lib.PREC[precision]
    o.switch_tol = switch_tol
    o.bwd_precision = _lib.PREC[bwd_precision]
    o.eig_block = eig_block
    o.eig_max_iters = eig_max_iters
    o.eig_warm = int(bool(eig_warm))
    o.eig_warm_iters = eig_warm_iters
    o.eig_tol = eig_tol
    o.bwd_last = bwd_last
    return o


MIXED_FP32 = dict(precision="fp32", switch_tol=1e-5, eig_warm=True)
MIXED_BF16 = dict(precision="bf16", switch_tol=1e-2, eig_warm=True)


# ---------------------------------------------------------------------------
# models (src/hamiltonianmodels.jl, src/exampletensors.jl, src/variationalipeps.jl:4-20)
# ---------------------------------------------------------------------------
SX = np.array([[0.0, 1.0], [1.0, 0.0]])
SY = np.array([[0.0, -1.0j], [1.0j, 0.0]])
SZ = np.array([[1.0, 0.0], [0.0, -1.0]])
ID2 = np.eye(2)


class HamiltonianModel:
    pass


class Ising(HamiltonianModel):
    pass


@dataclass
class TFIsing(HamiltonianModel):
    hx: float


@dataclass
class Heisenberg(HamiltonianModel):
    Jz: float = 1.0
    Jx: float = 1.0
    Jy: float = 1.0


def _outer(a, b):
    return np.einsum("ij,kl->ijkl", a, b)


def hamiltonian(model: HamiltonianModel) -> np.ndarray:
    """src/hamiltonianmodels.jl:28-33 (TFIsing) and :53-59 (Heisenberg)."""
    if isinstance(model, Heisenberg):
        h = model.Jz * _outer(SZ, SZ) - model.Jx * _outer(SX, SX) - model.Jy * _outer(SY, SY)
        h = np.einsum("ijcd,kc,ld->ijkl", h, SX, SX.conj().T)
        return np.asfortranarray(np.real(h / 2))
    if isinstance(model, TFIsing):
        hx = model.hx
        return np.asfortranarray(np.real(-2 * _outer(SZ, SZ) - hx / 2 * _outer(SX, ID2) - hx / 2 * _outer(ID2, SX)))
    raise TypeError(f"no hamiltonian for {model!r}")


def diaglocalhamiltonian(diag) -> np.ndarray:
    """src/variationalipeps.jl:15-20."""
    diag = np.asarray(diag, dtype=float)
    n = len(diag)
    h = np.diag(diag)
    idm = np.eye(n)
    r = lambda x, s: np.reshape(x, s, order="F")
    return np.asfortranarray(r(h, (n, n, 1, 1)) * r(idm, (1, 1, n, n)) + r(h, (1, 1, n, n)) * r(idm, (n, n, 1, 1)))


def _ising_q(beta):
    cb, sb = math.sqrt(math.cosh(beta)), math.sqrt(math.sinh(beta))
    return 1 / math.sqrt(2) * np.array([[cb + sb, cb - sb], [cb - sb, cb + sb]])


def _ising_dq(beta):
    cb, sb = math.sqrt(math.cosh(beta)), math.sqrt(math.sinh(beta))
    dcb, dsb = math.sinh(beta) / (2 * cb), math.cosh(beta) / (2 * sb)
    return 1 / math.sqrt(2) * np.array([[dcb + dsb, dcb - dsb], [dcb - dsb, dcb + dsb]])


def _delta4(last=1.0):
    a = np.zeros((2, 2, 2, 2))
    a[0, 0, 0, 0] = 1.0
    a[1, 1, 1, 1] = last
    return a


def model_tensor(model: HamiltonianModel, beta: float) -> np.ndarray:
    """src/exampletensors.jl:30-35."""
    if not isinstance(model, Ising):
        raise TypeError("model_tensor is defined for Ising()")
    q = _ising_q(beta)
    return np.asfortranarray(np.einsum("abcd,ai,bj,ck,dl->ijkl", _delta4(), q, q, q, q))


def dmodel_tensor(model: HamiltonianModel, beta: float) -> np.ndarray:
    """d model_tensor / d beta (product rule over the four q factors)."""
    q, dq, a = _ising_q(beta), _ising_dq(beta), _delta4()
    e = lambda *m: np.einsum("abcd,ai,bj,ck,dl->ijkl", a, *m)
    return np.asfortranarray(e(dq, q, q, q) + e(q, dq, q, q) + e(q, q, dq, q) + e(q, q, q, dq))


def mag_tensor(model: HamiltonianModel, beta: float) -> np.ndarray:
    """src/exampletensors.jl:43-48."""
    q = _ising_q(beta)
    return np.asfortranarray(np.einsum("abcd,ai,bj,ck,dl->ijkl", _delta4(-1.0), q, q, q, q))


ISING_BETA_C = math.log(1 + math.sqrt(2)) / 2


def magofbeta(model: HamiltonianModel, beta: float) -> float:
    """src/exampletensors.jl:77 (Yang's closed form)."""
    return (1 - math.sinh(2 * beta) ** -4) ** 0.125 if beta > ISING_BETA_C else 0.0


# ---------------------------------------------------------------------------
# iPEPS and CTMRG runtime (src/ipeps.jl, src/ctmrg.jl)
# ---------------------------------------------------------------------------
@dataclass
class SquareIPEPS:
    """src/ipeps.jl:16-22: bulk of shape (d, d, d, d, s); DimensionMismatch otherwise."""
    bulk: np.ndarray

    def __post_init__(self):
        s = np.shape(self.bulk)
        if len(s) != 5 or not (s[0] == s[1] == s[2] == s[3]):
            raise DimensionMismatch(-2, f"size of tensor error, should be `(d, d, d, d, s)`, got {s}.")
        self.bulk = _host(self.bulk)

    @property
    def d(self):
        return self.bulk.shape[0]

    @property
    def s(self):
        return self.bulk.shape[4]


def getd(ipeps: SquareIPEPS) -> int:
    return ipeps.d


def gets(ipeps: SquareIPEPS) -> int:
    return ipeps.s


def indexperm_symmetrize(ipeps: SquareIPEPS, ctx: Context | None = None) -> SquareIPEPS:
    """src/ipeps.jl:32-39 on the GPU (Float64 bulks; the complex path symmetrises inside energy)."""
    ctx = ctx or default_context()
    _real_only(ipeps.bulk, "indexperm_symmetrize")
    pin, _keep = _in(ipeps.bulk)
    pout, out = _out(ipeps.bulk.shape)
    ctx.check(ctx.lib.tnad_indexperm_symmetrize(ctx.h, pin, ipeps.d, ipeps.s, pout))
    return SquareIPEPS(out)


def double_layer(ipeps: SquareIPEPS, ctx: Context | None = None, want_ap=True):
    """src/variationalipeps.jl:30-34: returns (ap, a) with D = d^2."""
    ctx = ctx or default_context()
    _real_only(ipeps.bulk, "double_layer")
    d, s = ipeps.d, ipeps.s
    D = d * d
    pin, _keep = _in(ipeps.bulk)
    pap, ap = _out((D, D, D, D, s, s)) if want_ap else (C.c_void_p(0), None)
    pa, a = _out((D, D, D, D))
    ctx.check(ctx.lib.tnad_double_layer(ctx.h, pin, d, s, pap, pa))
    return ap, a


@dataclass
class SquareCTMRGRuntime:
    """src/ctmrg.jl:23-34: bulk (D,D,D,D), corner (chi,chi), edge (chi,D,chi)."""
    bulk: np.ndarray
    corner: np.ndarray
    edge: np.ndarray

    @staticmethod
    def new(bulk, env: str, chi: int, ctx: Context | None = None, rng=None) -> "SquareCTMRGRuntime":
        """SquareCTMRGRuntime(bulk, Val(env), chi) (src/ctmrg.jl:62-86)."""
        bulk = _host(bulk)
        D = bulk.shape[0]
        if env == "raw":
            ctx = ctx or default_context()
            pa, _k = _in(bulk)
            pC, Cm = _out((chi, chi))
            pT, Tm = _out((chi, D, chi))
            ctx.check(ctx.lib.tnad_init_raw(ctx.h, pa, D, chi, pC, pT))
            return SquareCTMRGRuntime(bulk, Cm, Tm)
        if env == "random":
            # src/ctmrg.jl:66-72 (draw from NumPy's generator; Julia's randn stream is not reproducible here)
            rng = rng or np.random.default_rng()
            Cm = rng.standard_normal((chi, chi))
            Tm = rng.standard_normal((chi, D, chi))
            Cm = Cm + Cm.T
            Tm = Tm + np.transpose(Tm, (2, 1, 0))
            return SquareCTMRGRuntime(bulk, _host(Cm), _host(Tm))
        raise ValueError(f"unknown environment initialisation {env!r}")

    @property
    def chi(self):
        return self.corner.shape[0]

    @property
    def D(self):
        return self.bulk.shape[0]


def getchi(rt: SquareCTMRGRuntime) -> int:
    return rt.chi


def ctmrgstep(rt: SquareCTMRGRuntime, vals=None, ctx: Context | None = None, **opts):
    """src/ctmrg.jl:126-153: returns (new runtime, vals)."""
    ctx = ctx or default_context()
    chi, D = rt.chi, rt.D
    o = make_opts(**opts)
    pa, _a = _in(rt.bulk)
    pC, _c = _in(rt.corner)
    pT, _t = _in(rt.edge)
    pCo, Co = _out((chi, chi))
    pTo, To = _out((chi, D, chi))
    pv, v = _out((chi * D,))
    ctx.check(ctx.lib.tnad_ctmrgstep(ctx.h, pa, D, chi, pC, pT, C.byref(o), pCo, pTo, pv))
    return SquareCTMRGRuntime(rt.bulk, Co, To), v


@dataclass
class CtmrgTape:
    """Device checkpoint of a ctmrg run (tnad_tape); ``back`` runs the reverse sweep."""
    ctx: Context
    handle: C.c_void_p
    D: int
    chi: int

    def back(self, Cbar, Tbar) -> np.ndarray:
        pC, _c = _in(Cbar)
        pT, _t = _in(Tbar)
        pa, abar = _out((self.D,) * 4)
        self.ctx.check(self.ctx.lib.tnad_ctmrg_back(self.ctx.h, self.handle, pC, pT, pa))
        return abar

    def free(self):
        if self.handle:
            self.ctx.lib.tnad_tape_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def ctmrg(rt: SquareCTMRGRuntime, tol: float, maxit: int, ctx: Context | None = None, keep_tape=False,
          return_info=False, **opts):
    """src/ctmrg.jl:110-117 (fixedpoint + StopFunction semantics)."""
    ctx = ctx or default_context()
    chi, D = rt.chi, rt.D
    o = make_opts(**opts)
    pa, _a = _in(rt.bulk)
    pC0, _c = _in(rt.corner)
    pT0, _t = _in(rt.edge)
    pC, Cm = _out((chi, chi))
    pT, Tm = _out((chi, D, chi))
    pv, v = _out((chi * D,))
    it = C.c_int(0)
    tape = C.c_void_p()
    ctx.check(ctx.lib.tnad_ctmrg(ctx.h, pa, D, chi, pC0, pT0, float(tol), int(maxit), C.byref(o), pC, pT, pv,
                                 C.byref(it), C.byref(tape) if keep_tape else None))
    out = SquareCTMRGRuntime(rt.bulk, Cm, Tm)
    if not (keep_tape or return_info):
        return out
    info = {"iters": it.value, "vals": v, **ctx.stats()}
    if keep_tape:
        return out, info, CtmrgTape(ctx, tape, D, chi)
    return out, info


def expectationvalue(h, ap, rt: SquareCTMRGRuntime, ctx: Context | None = None) -> float:
    """src/variationalipeps.jl:49-56."""
    ctx = ctx or default_context()
    s = np.shape(h)[0]
    ph, _h = _in(h)
    pap, _ap = _in(ap)
    pC, _c = _in(rt.corner)
    pT, _t = _in(rt.edge)
    e = C.c_double(0.0)
    ctx.check(ctx.lib.tnad_expectationvalue(ctx.h, ph, s, pap, rt.D, rt.chi, pC, pT, C.byref(e)))
    return e.value


def energy_and_grad(h, ipeps: SquareIPEPS, chi: int, tol: float, maxit: int, ctx: Context | None = None,
                    want_grad=True, return_info=False, **opts):
    """energy (src/variationalipeps.jl:28-40) and d energy / d bulk in one library call."""
    ctx = ctx or default_context()
    if not isinstance(ipeps, SquareIPEPS):
        ipeps = SquareIPEPS(ipeps)
    if np.iscomplexobj(ipeps.bulk) or np.iscomplexobj(h):
        if opts:
            raise ValueError("the complex path runs in FP64 with the reference's semantics: no run options")
        return _energy_and_grad_z(h, ipeps, chi, tol, maxit, ctx, want_grad, return_info)
    o = make_opts(**opts)
    ph, _h = _in(h)
    pb, _b = _in(ipeps.bulk)
    pg, g = _out(ipeps.bulk.shape) if want_grad else (None, None)
    e = C.c_double(0.0)
    it = C.c_int(0)
    ctx.check(ctx.lib.tnad_energy(ctx.h, ph, pb, ipeps.d, ipeps.s, int(chi), float(tol), int(maxit), C.byref(o),
                                  C.byref(e), pg, C.byref(it)))
    if return_info:
        return e.value, g, {"iters": it.value, **ctx.stats()}
    return e.value, g


def _energy_and_grad_z(h, ipeps, chi, tol, maxit, ctx, want_grad, return_info):
    """ComplexF64 energy (complex e/n) and the gradient of Re(energy) (tnad_energy_z)."""
    hz = np.asfortranarray(np.asarray(h, dtype=np.complex128))
    bz = np.asfortranarray(np.asarray(ipeps.bulk, dtype=np.complex128))
    pg, g = _out(bz.shape, np.complex128) if want_grad else (None, None)
    e = np.zeros(2)
    it = C.c_int(0)
    ctx.check(ctx.lib.tnad_energy_z(ctx.h, C.c_void_p(hz.ctypes.data), C.c_void_p(bz.ctypes.data), ipeps.d,
                                    ipeps.s, int(chi), float(tol), int(maxit), C.c_void_p(e.ctypes.data), pg,
                                    C.byref(it)))
    ez = complex(e[0], e[1])
    if return_info:
        return ez, g, {"iters": it.value, **ctx.stats()}
    return ez, g


def svd(A, ctx: Context | None = None):
    """LinearAlgebra.svd (thin) of a complex m x p matrix on the GPU (m + p <= 96): (U, S, V) with
    A = U diag(S) V^H."""
    ctx = ctx or default_context()
    A = np.asfortranarray(np.asarray(A, dtype=np.complex128))
    m, p = A.shape
    K = min(m, p)
    pu, U = _out((m, K), np.complex128)
    pv, V = _out((p, K), np.complex128)
    ps, S = _out((K,))
    ctx.check(ctx.lib.tnad_svd_z(ctx.h, C.c_void_p(A.ctypes.data), m, p, pu, ps, pv))
    return U, S, V


def svd_back(A, dU=None, dS=None, dV=None, ctx: Context | None = None):
    """svd_back (src/trg.jl:72-106) through the GPU's complex SVD of A: cotangents of U, S, V (each
    optional) -> cotangent of A.  None for all three gives None, as the reference (:73)."""
    if dU is None and dS is None and dV is None:
        return None
    ctx = ctx or default_context()
    A = np.asfortranarray(np.asarray(A, dtype=np.complex128))
    m, p = A.shape
    K = min(m, p)
    keep = []

    def arg(x, shape, dtype):
        if x is None:
            return None
        x = np.asfortranarray(np.asarray(x, dtype=dtype))
        if x.shape != shape:
            raise ValueError(f"svd_back: cotangent of shape {x.shape}, expected {shape}")
        keep.append(x)
        return C.c_void_p(x.ctypes.data)

    pdu = arg(dU, (m, K), np.complex128)
    pds = arg(dS, (K,), np.float64)
    pdv = arg(dV, (p, K), np.complex128)
    pa, dA = _out((m, p), np.complex128)
    ctx.check(ctx.lib.tnad_svd_back_z(ctx.h, C.c_void_p(A.ctypes.data), m, p, pdu, pds, pdv, pa))
    return dA


def magnetisation(model: HamiltonianModel, beta: float, chi: int, rng=None, tol: float = 1e-6, maxit: int = 100,
                  ctx: Context | None = None) -> float:
    """src/exampletensors.jl:57-69: |<m>| of the model from CTMRG (:random start, tol 1e-6,
    maxit 100) and the four-site environment, every contraction on the GPU."""
    ctx = ctx or default_context()
    a = model_tensor(model, beta)
    m = mag_tensor(model, beta)
    rt = ctmrg(SquareCTMRGRuntime.new(a, "random", chi, rng=rng), tol, maxit, ctx=ctx)
    Cm, Tm = rt.corner, rt.edge
    # ctc = ein"ia,ajb,bk -> ijk"(corner, edge, corner)
    ct = contract("ia,ajb->ijb", Cm, Tm, ctx=ctx)
    ctc = contract("ijb,bk->ijk", ct, Cm, ctx=ctx)
    # env = ein"alc,ckd,bjd,bia -> ijkl"(ctc, edge, ctc, edge), pairwise
    x1 = contract("alc,ckd->alkd", ctc, Tm, ctx=ctx)
    x2 = contract("bjd,bia->jdia", ctc, Tm, ctx=ctx)
    env = contract("alkd,jdia->ijkl", x1, x2, ctx=ctx)
    mag = contract("ijkl,ijkl->", env, m, ctx=ctx)
    nrm = contract("ijkl,ijkl->", env, a, ctx=ctx)
    return abs(float(np.ravel(mag)[0]) / float(np.ravel(nrm)[0]))


def energy(h, ipeps: SquareIPEPS, chi: int, tol: float, maxit: int, ctx: Context | None = None, **opts) -> float:
    """src/variationalipeps.jl:28-40."""
    return energy_and_grad(h, ipeps, chi, tol, maxit, ctx=ctx, want_grad=False, **opts)[0]


def rrule_energy(h, ipeps: SquareIPEPS, chi: int, tol: float, maxit: int, ctx: Context | None = None, **opts):
    """ChainRulesCore-style rule: (energy, pullback) with pullback(ebar) -> bulk cotangent."""
    e, g = energy_and_grad(h, ipeps, chi, tol, maxit, ctx=ctx, **opts)
    return e, (lambda ebar: ebar * g)


def contract(spec: str, A, B, precision="fp64", ctx: Context | None = None, alpha: float = 1.0,
             beta: float = 0.0, Cin=None) -> np.ndarray:
    """One library contraction ``C = alpha ein"spec"(A, B) + beta Cin`` (column-major tensors,
    single-character index names, every output index must appear in A or B)."""
    ctx = ctx or default_context()
    lhs, ic = spec.replace(" ", "").split("->")
    ia, ib = lhs.split(",")
    A, B = _host(A), _host(B)
    ext = {}
    for idx, T in ((ia, A), (ib, B)):
        for ch, n in zip(idx, T.shape):
            ext[ch] = n
    dc = tuple(ext[ch] for ch in ic)
    arr = lambda t: (C.c_longlong * len(t))(*t)
    pa, _a = _in(A)
    pb, _b = _in(B)
    pc, out = _out(dc)
    if beta != 0.0:
        out[...] = _host(Cin)
    ctx.check(ctx.lib.tnad_contract(ctx.h, _lib.PREC[precision], ia.encode(), arr(A.shape), pa, ib.encode(),
                                    arr(B.shape), pb, ic.encode(), arr(dc), float(alpha), float(beta), pc))
    return out


def time_contract(spec: str, extents: dict, precision="fp64", reps: int = 20,
                  ctx: Context | None = None) -> float:
    """Average device milliseconds of one launch of ``ein"spec"`` at the given index extents
    (HIP events on the context's stream, operands resident on the device)."""
    ctx = ctx or default_context()
    lhs, ic = spec.replace(" ", "").split("->")
    ia, ib = lhs.split(",")
    arr = lambda idx: (C.c_longlong * len(idx))(*[extents[ch] for ch in idx])
    ms = C.c_double(0.0)
    ctx.check(ctx.lib.tnad_time_contract(ctx.h, _lib.PREC[precision], ia.encode(), arr(ia), ib.encode(), arr(ib),
                                         ic.encode(), arr(ic), int(reps), C.byref(ms)))
    return ms.value


def time_syev(m: int = 176, reps: int = 10, ctx: Context | None = None):
    """Average device ms of the divide-and-conquer eigensolver's stages on a random symmetric m x m matrix:
    (tridiagonalisation, divide and conquer, back-transformation + sort), HIP events on the context's stream."""
    ctx = ctx or default_context()
    ms = (C.c_double * 3)()
    ctx.check(ctx.lib.tnad_time_syev(ctx.h, int(m), int(reps), ms))
    return tuple(ms)


def time_jacobi(m: int = 192, sweeps: int = 3, reps: int = 5, ctx: Context | None = None):
    """(average device ms, rounds) of one launch of the Jacobi rotation kernel on a random symmetric m x m
    matrix, `sweeps` sweeps (HIP events on the context's stream)."""
    ctx = ctx or default_context()
    ms, rounds = C.c_double(0.0), C.c_int(0)
    ctx.check(ctx.lib.tnad_time_jacobi(ctx.h, int(m), int(sweeps), int(reps), C.byref(ms), C.byref(rounds)))
    return ms.value, rounds.value


# ---------------------------------------------------------------------------
# TRG (src/trg.jl)
# ---------------------------------------------------------------------------
def trg_and_grad(a, chi: int, niter: int, tol: float = 1e-16, ctx: Context | None = None, want_grad=True):
    """trg (src/trg.jl:13-30) and d lnZ / d a in one library call.  A complex tensor gives a complex lnZ and
    the gradient of Re(lnZ) (tnad_trg_z)."""
    ctx = ctx or default_context()
    a = _host(a)
    if np.iscomplexobj(a):
        q = a.shape[0]
        pg, g = _out(a.shape, np.complex128) if want_grad else (None, None)
        lz = np.zeros(2)
        ctx.check(ctx.lib.tnad_trg_z(ctx.h, C.c_void_p(a.ctypes.data), q, int(chi), int(niter), float(tol),
                                     C.c_void_p(lz.ctypes.data), pg))
        return complex(lz[0], lz[1]), g
    q = a.shape[0]
    pa, _a = _in(a)
    pg, g = _out(a.shape) if want_grad else (None, None)
    lnz = C.c_double(0.0)
    ctx.check(ctx.lib.tnad_trg(ctx.h, pa, q, int(chi), int(niter), float(tol), C.byref(lnz), pg))
    return lnz.value, g


def trg(a, chi: int, niter: int, tol: float = 1e-16, ctx: Context | None = None) -> float:
    """src/trg.jl:13-30: ln Z per site."""
    return trg_and_grad(a, chi, niter, tol, ctx=ctx, want_grad=False)[0]


def rrule_trg(a, chi: int, niter: int, tol: float = 1e-16, ctx: Context | None = None):
    lnz, g = trg_and_grad(a, chi, niter, tol, ctx=ctx)
    return lnz, (lambda zbar: zbar * g)


def trg_svd(t, dmax: int, tol: float, ctx: Context | None = None):
    """src/trg.jl:33-44: returns (u, v)."""
    ctx = ctx or default_context()
    t = _host(t)
    d1, d2, d3, d4 = t.shape
    kmax = min(dmax, d1 * d2, d3 * d4)
    pt, _t = _in(t)
    pu, u = _out((d1, d2, kmax))
    pv, v = _out((kmax, d3, d4))
    k = C.c_int(0)
    ctx.check(ctx.lib.tnad_trg_svd(ctx.h, pt, d1, d2, d3, d4, int(dmax), float(tol), pu, pv, C.byref(k)))
    kk = k.value
    return np.asfortranarray(u[:, :, :kk]), np.asfortranarray(v[:kk, :, :])


def trg_svd_back(t, dmax: int, tol: float, ubar, vbar, ctx: Context | None = None):
    """Pullback of trg_svd (src/trg.jl:33-44 through svd_back, src/trg.jl:55-61,72-106): the cotangent of t
    from those of u (d1, d2, k) and v (k, d3, d4); k is the kept count of trg_svd(t, dmax, tol)."""
    ctx = ctx or default_context()
    t = _host(t)
    d1, d2, d3, d4 = t.shape
    ubar = np.asfortranarray(_host(ubar), dtype=np.float64)
    vbar = np.asfortranarray(_host(vbar), dtype=np.float64)
    k = ubar.shape[2] if ubar.ndim == 3 else -1
    if k < 1 or ubar.shape != (d1, d2, k) or vbar.shape != (k, d3, d4):
        raise ValueError("trg_svd_back: ubar must be (d1, d2, k) and vbar (k, d3, d4)")
    pt, _t = _in(t)
    pub, _ub = _in(ubar)
    pvb, _vb = _in(vbar)
    ptb, tbar = _out((d1, d2, d3, d4))
    ctx.check(ctx.lib.tnad_trg_svd_back(ctx.h, pt, d1, d2, d3, d4, int(dmax), float(tol), pub, pvb, k, ptb))
    return tbar


# ---------------------------------------------------------------------------
# finite differences (src/autodiff.jl:44,58-63) and the optimiser (src/variationalipeps.jl:67-75)
# ---------------------------------------------------------------------------
def num_grad(f, x, delta: float = 1e-5):
    """Central differences (f(x+d/2) - f(x-d/2)) / d, elementwise for arrays."""
    if np.isscalar(x):
        return (f(x + delta / 2) - f(x - delta / 2)) / delta
    x = _host(x)
    g = np.zeros_like(x)
    for i in np.ndindex(x.shape):
        xp, xm = x.copy(order="F"), x.copy(order="F")
        xp[i] += delta / 2
        xm[i] -= delta / 2
        g[i] = (f(xp) - f(xm)) / delta
    return g


@dataclass
class OptimResult:
    minimizer: np.ndarray
    minimum: float
    iterations: int
    f_calls: int
    trace: list = field(default_factory=list)


def optimiseipeps(ipeps: SquareIPEPS, h, chi: int, tol: float, maxit: int, f_tol: float = 1e-6,
                  iterations: int = 1000, m: int = 20, ctx: Context | None = None, store_trace=False, **opts):
    """src/variationalipeps.jl:67-75: minimise energy(h, ipeps) over the bulk with L-BFGS(m=20).

    The reference calls Optim.jl with separate f and g closures, so every
    evaluation runs the forward twice; here each L-BFGS evaluation is one
    fused energy + gradient library call.  The L-BFGS driver is SciPy's
    (host side, as Optim.jl is in the reference)."""
    from scipy.optimize import minimize

    ctx = ctx or default_context()
    shape = ipeps.bulk.shape
    trace = []
    # a complex bulk is optimised over (Re, Im) with the objective real(energy) (src/variationalipeps.jl:71);
    # its gradient dRe(e)/dRe x + i dRe(e)/dIm x splits the same way
    cplx = np.iscomplexobj(ipeps.bulk)
    nel = int(np.prod(shape))

    def unpack(x):
        if cplx:
            x = x[:nel] + 1j * x[nel:]
        return np.reshape(x, shape, order="F")

    def fg(x):
        e, g = energy_and_grad(h, SquareIPEPS(unpack(x)), chi, tol, maxit, ctx=ctx, **opts)
        e = float(np.real(e))
        if store_trace:
            trace.append(e)
        g = np.reshape(g, -1, order="F")
        return e, (np.concatenate([g.real, g.imag]) if cplx else g)

    x0 = np.reshape(ipeps.bulk, -1, order="F")
    if cplx:
        x0 = np.concatenate([x0.real, x0.imag])
    res = minimize(fg, x0, jac=True, method="L-BFGS-B",
                   options=dict(maxcor=m, ftol=f_tol, gtol=1e-12, maxiter=iterations))
    return OptimResult(unpack(res.x), float(res.fun), int(res.nit), int(res.nfev), trace)
