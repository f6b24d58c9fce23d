"""CPU FP64 oracle for the CTMRG / TRG hot path of TensorNetworkAD.jl.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this module, and only as the
checker (or as the timed CPU baseline).  The product path (the HIP library and
its ctypes host layer) never calls it.

This is a NumPy restatement of the reference algorithm.  Every function cites
the reference file:line it follows (paths relative to the reference repo).
Arrays follow Julia's semantics: all reshapes are column-major (order='F') and
every ``ein"..."`` string is restated as the same ``np.einsum`` string, which
operates on logical indices and is therefore layout independent.

Reverse mode: the reference differentiates with Zygote, which unrolls every
CTMRG/TRG iteration and uses the hand-written ``svd_back`` (src/trg.jl:72-106)
plus OMEinsum's einsum adjoints.  Zygote is not available here, so the oracle
carries its own hand-written reverse sweep of exactly that unrolled program
(``energy_and_grad``, ``trg_and_grad``).  Each adjoint is the textbook
einsum / norm / permutation adjoint; the whole sweep is pinned against central
finite differences (``num_grad``, src/autodiff.jl:58-63) and against the
reference's published gradients.

Parity pins (all exercised in tests/test_oracle.py):
  * trg: golden scalars test/trg.jl:18, README.md:60, README.md:70,
    docs/src/userguide.md:28; Onsager closed form; FD of d lnZ / d beta.
  * trg_svd: reconstruction at tol=0 (test/trg.jl:6-10).
  * svd_back: central FD on small random matrices (as test/svd.jl).
  * ctmrg/energy: product-state energies -1, +1, 0 (test/variationalipeps.jl:14-25);
    Yang magnetisation (test/ctmrg.jl:35-42); step invariants (C=C^T,
    T=T^{kji}, |C|=|T|=1, vals[0]=1); gauge invariance of the energy.
  * energy_and_grad: central FD on every bulk entry at small chi
    (test/variationalipeps.jl:121-134 pattern).
  * hamiltonian(Heisenberg()): the matrix printed at README.md:83-100.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

ETA_SVD = 1e-40          # src/trg.jl:72 (Lorentzian broadening in svd_back)
EPS0F0 = float(np.finfo(np.float32).smallest_subnormal)  # eps(0f0), src/autodiff.jl:26


def ein(spec, *ops):
    return np.einsum(spec, *ops, optimize=True)


def reshape_f(x, shape):
    """Julia ``reshape`` (column-major)."""
    return np.reshape(x, shape, order="F")


# --------------------------------------------------------------------------
# inputs: Hamiltonians and model tensors
# --------------------------------------------------------------------------
SX = np.array([[0.0, 1.0], [1.0, 0.0]])                 # src/variationalipeps.jl:4
SY = np.array([[0.0, -1.0j], [1.0j, 0.0]])              # src/variationalipeps.jl:5
SZ = np.array([[1.0, 0.0], [0.0, -1.0]])                # src/variationalipeps.jl:6
ID2 = np.eye(2)                                         # src/variationalipeps.jl:7


def hamiltonian_heisenberg(Jz=1.0, Jx=1.0, Jy=1.0):
    """src/hamiltonianmodels.jl:53-59: Jz szsz - Jx sxsx - Jy sysy, site 2 rotated by sx, real(h/2)."""
    h = (Jz * ein("ij,kl->ijkl", SZ, SZ) - Jx * ein("ij,kl->ijkl", SX, SX)
         - Jy * ein("ij,kl->ijkl", SY, SY))
    h = ein("ijcd,kc,ld->ijkl", h, SX, SX.conj().T)
    return np.real(h / 2).copy()


def hamiltonian_tfising(hx):
    """src/hamiltonianmodels.jl:28-33."""
    return (-2 * ein("ij,kl->ijkl", SZ, SZ) - hx / 2 * ein("ij,kl->ijkl", SX, ID2)
            - hx / 2 * ein("ij,kl->ijkl", ID2, SX))


def diaglocalhamiltonian(diag):
    """src/variationalipeps.jl:15-20 (test input builder)."""
    diag = np.asarray(diag, dtype=float)
    n = len(diag)
    h = np.diag(diag)
    idm = np.eye(n)
    return (reshape_f(h, (n, n, 1, 1)) * reshape_f(idm, (1, 1, n, n))
            + reshape_f(h, (1, 1, n, n)) * reshape_f(idm, (n, n, 1, 1)))


def _ising_q(beta):
    cb, sb = math.sqrt(math.cosh(beta)), math.sqrt(math.sinh(beta))
    return 1 / math.sqrt(2) * np.array([[cb + sb, cb - sb], [cb - sb, cb + sb]])


def _ising_dq(beta):
    """d q / d beta for _ising_q (chain rule through sqrt(cosh), sqrt(sinh))."""
    cb, sb = math.sqrt(math.cosh(beta)), math.sqrt(math.sinh(beta))
    dcb = math.sinh(beta) / (2 * cb)
    dsb = math.cosh(beta) / (2 * sb)
    return 1 / math.sqrt(2) * np.array([[dcb + dsb, dcb - dsb], [dcb - dsb, dcb + dsb]])


def _delta4(sign_last=1.0):
    a = np.zeros((2, 2, 2, 2))
    a[0, 0, 0, 0] = 1.0
    a[1, 1, 1, 1] = sign_last
    return a


def model_tensor_ising(beta):
    """src/exampletensors.jl:30-35."""
    q = _ising_q(beta)
    return ein("abcd,ai,bj,ck,dl->ijkl", _delta4(), q, q, q, q)


def dmodel_tensor_ising(beta):
    """d model_tensor / d beta (product rule over the four q factors)."""
    a, q, dq = _delta4(), _ising_q(beta), _ising_dq(beta)
    return (ein("abcd,ai,bj,ck,dl->ijkl", a, dq, q, q, q) + ein("abcd,ai,bj,ck,dl->ijkl", a, q, dq, q, q)
            + ein("abcd,ai,bj,ck,dl->ijkl", a, q, q, dq, q) + ein("abcd,ai,bj,ck,dl->ijkl", a, q, q, q, dq))


def mag_tensor_ising(beta):
    """src/exampletensors.jl:43-48."""
    q = _ising_q(beta)
    return ein("abcd,ai,bj,ck,dl->ijkl", _delta4(-1.0), q, q, q, q)


def tensorfromclassical(ham):
    """src/exampletensors.jl:17-21 (matrix sqrt of the elementwise Boltzmann weights)."""
    w = np.exp(ham)
    ev, evec = np.linalg.eigh(w)
    q = (evec * np.sqrt(ev.astype(complex))) @ evec.T
    q = np.real_if_close(q)
    return ein("ij,ik,il,im->jklm", q, q, q, q)


ISING_BETA_C = math.log(1 + math.sqrt(2)) / 2   # src/exampletensors.jl:2


def magofbeta(beta):
    """Yang's closed form, src/exampletensors.jl:77."""
    return (1 - math.sinh(2 * beta) ** -4) ** 0.125 if beta > ISING_BETA_C else 0.0


def onsager_lnz(beta, npts=4000):
    """Onsager's ln Z per site of the 2D Ising model (closed form, independent pin).

    ln Z = ln 2 + 1/(2 pi^2) int_0^pi int_0^pi ln[cosh^2 2b - sinh 2b (cos t1 + cos t2)] dt1 dt2.
    The double integral is done with Gauss-Legendre quadrature in both variables."""
    x, w = np.polynomial.legendre.leggauss(npts)
    t = (x + 1) * math.pi / 2
    wt = w * math.pi / 2
    c2 = math.cosh(2 * beta) ** 2
    s2 = math.sinh(2 * beta)
    ct = np.cos(t)
    f = np.log(c2 - s2 * (ct[:, None] + ct[None, :]))
    return math.log(2) + (wt @ f @ wt) / (2 * math.pi ** 2)


# --------------------------------------------------------------------------
# iPEPS front: symmetrisation and double layer
# --------------------------------------------------------------------------
_PERMS = [(0, 3, 2, 1, 4),   # left-right   src/ipeps.jl:34
          (2, 1, 0, 3, 4),   # up-down      src/ipeps.jl:35
          (1, 0, 3, 2, 4),   # diagonal     src/ipeps.jl:36
          (3, 2, 1, 0, 4)]   # rotation     src/ipeps.jl:37


def check_square_ipeps(bulk):
    """src/ipeps.jl:17-22: DimensionMismatch unless the four virtual dims agree."""
    s = bulk.shape
    if len(s) != 5 or not (s[0] == s[1] == s[2] == s[3]):
        raise ValueError(f"size of tensor error, should be `(d, d, d, d, s)`, got {s}.")


def indexperm_symmetrize(bulk):
    """src/ipeps.jl:32-39: x += permutedims(x, p) for the four permutations, then x / norm(x)."""
    check_square_ipeps(bulk)
    x = bulk
    for p in _PERMS:
        x = x + np.transpose(x, p)
    return x / np.linalg.norm(x)


def indexperm_symmetrize_back(bulk, xbar):
    """Adjoint of indexperm_symmetrize (norm adjoint of src/autodiff.jl:23-29 included)."""
    xs = [bulk]
    x = bulk
    for p in _PERMS:
        x = x + np.transpose(x, p)
        xs.append(x)
    nrm = np.linalg.norm(x)
    # y = x / n : xbar_x = ybar/n ; nbar = -<ybar, x>/n^2 ; norm adjoint: nbar * x / (n + eps(0f0))
    nbar = -np.vdot(x, xbar).real / nrm ** 2
    g = xbar / nrm + nbar * x / (nrm + EPS0F0)
    for p in reversed(_PERMS):
        g = g + np.transpose(g, np.argsort(p))
    return g


def double_layer(bulk):
    """src/variationalipeps.jl:30-34: ap = A (x) conj(A) with fused (ket, bra) legs; a = tr_phys ap."""
    d = bulk.shape[0]
    s = bulk.shape[4]
    D = d * d
    ap = ein("abcdx,ijkly->aibjckdlxy", bulk, bulk.conj())
    ap = reshape_f(ap, (D, D, D, D, s, s))
    a = ein("ijklaa->ijkl", ap)
    return ap, a


def double_layer_back(bulk, apbar, abar):
    """Adjoint of double_layer: cotangents of ap (6-index) and a -> cotangent of the bulk.

    Cotangent convention for complex arrays (the one of Zygote's gradients): xbar with
    dL = Re sum conj(xbar) dx, i.e. xbar = dL/dRe x + i dL/dIm x.  For ap = ein(A, conj(A)) the
    first operand receives ein(apbar, A) and the second, through the conj, conj(ein(apbar, conj(A)))."""
    d = bulk.shape[0]
    s = bulk.shape[4]
    g6 = apbar + ein("ijkl,xy->ijklxy", abar, np.eye(s))
    g10 = reshape_f(g6, (d, d, d, d, d, d, d, d, s, s))
    gb = ein("aibjckdlxy,ijkly->abcdx", g10, bulk)
    gb = gb + ein("aibjckdlxy,abcdx->ijkly", g10.conj(), bulk)
    return gb


# --------------------------------------------------------------------------
# CTMRG
# --------------------------------------------------------------------------
def init_raw(a, chi):
    """src/ctmrg.jl:74-86: corner = sum_kl a, edge = sum_l a, embedded in zeros (truncated)."""
    D = a.shape[0]
    C = np.zeros((chi, chi), dtype=a.dtype)
    T = np.zeros((chi, D, chi), dtype=a.dtype)
    cinit = ein("ijkl->ij", a)
    tinit = ein("ijkl->ijk", a)
    m = min(chi, D)
    C[:m, :m] = cinit[:m, :m]
    T[:m, :, :m] = tinit[:m, :, :m]
    return C, T


def init_random(a, chi, rng):
    """src/ctmrg.jl:66-72 (random draw with NumPy's generator instead of Julia's randn)."""
    D = a.shape[0]
    C = rng.standard_normal((chi, chi))
    T = rng.standard_normal((chi, D, chi))
    C = C + C.conj().T
    T = T + np.transpose(T.conj(), (2, 1, 0))
    return C, T


@dataclass
class StepCache:
    C: np.ndarray
    T: np.ndarray
    cp: np.ndarray
    tp: np.ndarray
    U: np.ndarray
    s: np.ndarray
    V: np.ndarray
    z: np.ndarray
    C2: np.ndarray
    T2: np.ndarray


def phase_weights(n):
    """Weights of the complex phase rule: w_r = 1 + frac(0.6180339887 r), r = 0..n-1."""
    r = np.arange(n, dtype=np.float64)
    return 1.0 + np.mod(0.6180339887 * r, 1.0)


def fix_phases(U):
    """Phase gauge of complex singular / eigenvectors: column j is multiplied by conj(q)/|q| with
    q = sum_r w_r U[r, j] (phase_weights), so that q becomes real positive; when |q| <= 1e-6 |U_j|
    the largest-modulus entry (first on ties) is made real positive instead.  Real columns stay real
    (the phase is a sign).  Returns (U', phases)."""
    n, k = U.shape
    w = phase_weights(n)
    ph = np.ones(k, dtype=U.dtype)
    for j in range(k):
        q = np.sum(w * U[:, j])
        if abs(q) <= 1e-6 * np.linalg.norm(U[:, j]):
            q = U[int(np.argmax(np.abs(U[:, j]))), j]
        if q != 0:
            ph[j] = np.conj(q) / abs(q)
    return U * ph[None, :], ph


def herm_svd(M):
    """SVD of a Hermitian matrix through its eigendecomposition: U = eigenvectors sorted by |lambda|
    descending (stable), s = |lambda|, V = U sign(lambda), phases fixed by fix_phases.  For a complex
    Hermitian matrix the reference's result depends on LAPACK's singular-vector phases (the energy
    contracts l with l, not conj(l), src/variationalipeps.jl:53-54, so it is invariant under real
    orthogonal but not under unitary changes of the environment's gauge): this rule is the oracle's
    fixed choice, which the GPU path follows."""
    lam, X = np.linalg.eigh(M)
    o = np.argsort(-np.abs(lam), kind="stable")
    lam, X = lam[o], X[:, o]
    X, _ = fix_phases(X)
    sg = np.where(lam < 0, -1.0, 1.0)
    return X, np.abs(lam), (X * sg[None, :]).conj().T


def ctmrgstep(a, C, T, keep=False):
    """src/ctmrg.jl:126-153.  Returns (C', T', vals) [and the cache for the reverse sweep].
    Real tensors: LAPACK's SVD.  Complex tensors: herm_svd (fixed phase rule)."""
    chi = C.shape[0]
    D = a.shape[0]
    n = chi * D
    cp = ein("ad,iba,dcl,jkcb->ijlk", C, T, T, a)            # :130
    tp = ein("iam,jkla->ijklm", T, a)                        # :131
    cpmat = reshape_f(cp, (n, n))                            # :134
    cpmat = cpmat + cpmat.conj().T                           # :135
    if np.iscomplexobj(cpmat):
        U, s, Vh = herm_svd(cpmat)                           # :136
    else:
        U, s, Vh = np.linalg.svd(cpmat)                      # :136 (LAPACK gesdd)
    z = reshape_f(U[:, :chi], (chi, D, chi))                 # :137
    C1 = ein("abcd,abi,cdj->ij", cp, z.conj(), z)            # :139
    T1 = ein("abjcd,abi,dck->ijk", tp, z.conj(), z)          # :140
    vals = s / s[0]                                          # :142
    C2 = C1 + C1.conj().T                                    # :145
    T2 = T1 + ein("ijk->kji", T1.conj())                     # :146
    C3 = C2 / np.linalg.norm(C2)                             # :149
    T3 = T2 / np.linalg.norm(T2)                             # :150
    if keep:
        return C3, T3, vals, StepCache(C, T, cp, tp, U, s, Vh.conj().T, z, C2, T2)
    return C3, T3, vals


def svd_back(U, S, V, dU, dS, dV, eta=ETA_SVD):
    """src/trg.jl:72-106 (restated term by term; complex diagonal term included)."""
    if dU is None and dS is None and dV is None:
        return None
    S2 = S ** 2
    Sinv = S / (S2 + eta)
    F = S2[None, :] - S2[:, None]
    F = F / (F ** 2 + eta)
    res = np.zeros((len(S), len(S)), dtype=np.result_type(U, V, S))
    if dU is not None:
        UdU = U.conj().T @ dU
        J = F * UdU
        res = res + (J + J.conj().T) * S[None, :]
        if np.iscomplexobj(UdU):
            res = res + np.diag(1j * np.imag(np.diag(UdU)) * Sinv)
    if dV is not None:
        VdV = V.conj().T @ dV
        K = F * VdV
        res = res + S[:, None] * (K + K.conj().T)
    if dS is not None:
        res = res + np.diag(dS)
    res = U @ res @ V.conj().T
    if dU is not None and U.shape[0] != U.shape[1]:
        res = res + (dU - U @ (U.conj().T @ dU)) * Sinv[None, :] @ V.conj().T
    if dV is not None and V.shape[0] != V.shape[1]:
        res = res + (U * Sinv[None, :]) @ (dV.conj().T - (dV.conj().T @ V) @ V.conj().T)
    return res


def _norm_back(x, ybar):
    """Adjoint of y = x / norm(x) with the norm adjoint of src/autodiff.jl:23-29."""
    nrm = np.linalg.norm(x)
    nbar = -np.vdot(x, ybar).real / nrm ** 2
    return ybar / nrm + nbar * x / (nrm + EPS0F0)


def ctmrgstep_back(a, cache: StepCache, C3bar, T3bar):
    """Reverse sweep of one ctmrgstep.  Returns (Cbar, Tbar, abar).

    Real or complex: every einsum adjoint contracts the cotangent with the complex conjugates of the
    other operands (the conj of a holomorphic map's derivative); an operand entering as conj(z)
    passes the conjugate of its cotangent on to z.  On real arrays every conj is the identity."""
    C, T, cp, tp, U, s, V, z = cache.C, cache.T, cache.cp, cache.tp, cache.U, cache.s, cache.V, cache.z
    chi = C.shape[0]
    D = a.shape[0]
    n = chi * D
    cj = np.conj
    C2bar = _norm_back(cache.C2, C3bar)
    T2bar = _norm_back(cache.T2, T3bar)
    C1bar = C2bar + C2bar.conj().T                               # C2 = C1 + C1'
    T1bar = T2bar + ein("ijk->kji", T2bar).conj()                # T2 = T1 + conj(T1^{kji})
    zc = z.conj()
    # T1 = ein("abjcd,abi,dck->ijk", tp, conj(z), z)
    tpbar = ein("ijk,abi,dck->abjcd", T1bar, cj(zc), cj(z))
    zcbar = ein("abjcd,dck,ijk->abi", cj(tp), cj(z), T1bar)
    zbar = ein("abjcd,abi,ijk->dck", cj(tp), cj(zc), T1bar)
    # C1 = ein("abcd,abi,cdj->ij", cp, conj(z), z)
    cpbar = ein("ij,abi,cdj->abcd", C1bar, cj(zc), cj(z))
    zcbar = zcbar + ein("abcd,cdj,ij->abi", cj(cp), cj(z), C1bar)
    zbar = zbar + ein("abcd,abi,ij->cdj", cj(cp), cj(zc), C1bar)
    zbar = zbar + zcbar.conj()
    dU = np.zeros((n, n), dtype=np.result_type(U, zbar))
    dU[:, :chi] = reshape_f(zbar, (n, chi))
    Mbar = svd_back(U, s, V, dU, None, None)
    cpmatbar = Mbar + Mbar.conj().T                              # cpmat += cpmat'
    cpbar = cpbar + reshape_f(cpmatbar, (chi, D, chi, D))
    # tp = ein("iam,jkla->ijklm", T, a)
    Tbar = ein("ijklm,jkla->iam", tpbar, cj(a))
    abar = ein("ijklm,iam->jkla", tpbar, cj(T))
    # cp = ein("ad,iba,dcl,jkcb->ijlk", C, T, T, a)
    Cbar = ein("ijlk,iba,dcl,jkcb->ad", cpbar, cj(T), cj(T), cj(a))
    Tbar = Tbar + ein("ijlk,ad,dcl,jkcb->iba", cpbar, cj(C), cj(T), cj(a))
    Tbar = Tbar + ein("ijlk,ad,iba,jkcb->dcl", cpbar, cj(C), cj(T), cj(a))
    abar = abar + ein("ijlk,ad,iba,dcl->jkcb", cpbar, cj(C), cj(T), cj(T))
    return Cbar, Tbar, abar


@dataclass
class CtmrgResult:
    C: np.ndarray
    T: np.ndarray
    iters: int
    vals_trace: list = field(default_factory=list)
    caches: list = field(default_factory=list)


def ctmrg(a, C0, T0, tol, maxit, keep=False):
    """src/ctmrg.jl:110-117 with fixedpoint/StopFunction (src/fixedpoint.jl:11-41).

    StopFunction is evaluated on every state including the initial one: the counter
    (constructed as -1, src/ctmrg.jl:114) is incremented, the run stops once
    counter > maxit, otherwise when norm(vals - oldvals) <= tol (L2); oldvals starts
    at Inf, so the first two evaluations (NaN, Inf) never stop.  Hence at most
    maxit + 1 steps, and maxit = 0 still takes one step."""
    chi, D = C0.shape[0], a.shape[0]
    oldvals = np.full(chi * D, np.inf)
    vals = oldvals
    C, T = C0, T0
    counter = -1
    res = CtmrgResult(C, T, 0)
    while True:
        counter += 1
        if counter > maxit:
            break
        with np.errstate(invalid="ignore"):
            diff = np.linalg.norm(vals - oldvals)
        if diff <= tol:
            break
        oldvals = vals
        out = ctmrgstep(a, C, T, keep=keep)
        if keep:
            C, T, vals, cache = out
            res.caches.append(cache)
        else:
            C, T, vals = out
        res.vals_trace.append(vals)
        res.iters += 1
    res.C, res.T = C, T
    return res


def _half_plaquette(C, T, apn):
    """l = ein"ab,ica,bde,cjfdlm,eg,gfk -> ijklm"(C, T, T, apn, C, T) (src/variationalipeps.jl:52),
    evaluated pairwise: E1 = T C, E2 = E1 T, E3 = C T, E4 = E2 apn, l = E4 E3."""
    E1 = ein("ica,ab->icb", T, C)
    E2 = ein("icb,bde->icde", E1, T)
    E3 = ein("eg,gfk->efk", C, T)
    E4 = ein("icde,cjfdlm->iejflm", E2, apn)
    l = ein("iejflm,efk->ijklm", E4, E3)
    return l, (E1, E2, E3, E4)


def expectationvalue(h, ap, C, T):
    """src/variationalipeps.jl:49-56."""
    apn = ap / np.linalg.norm(ap)
    l, _ = _half_plaquette(C, T, apn)
    e = ein("abcij,abckl,ijkl->", l, l, h)
    nn = ein("ijkaa,ijkbb->", l, l)
    return e / nn


def expectationvalue_back(h, ap, C, T, Ebar=1.0):
    """Reverse sweep of expectationvalue for the loss Re(Ebar * e/n) (Re(e/n) is what optimiseipeps
    differentiates, src/variationalipeps.jl:71; for real tensors it is e/n).  Returns (apbar, Cbar, Tbar)."""
    cj = np.conj
    nrm = np.linalg.norm(ap)
    apn = ap / nrm
    l, (E1, E2, E3, E4) = _half_plaquette(C, T, apn)
    e = ein("abcij,abckl,ijkl->", l, l, h)
    tr = ein("ijkaa->ijk", l)
    nn = np.sum(tr * tr)
    ebar = cj(Ebar / nn)                 # q = e/n, qbar = Ebar: ebar = conj(dq/de) qbar
    nnbar = cj(-Ebar * e / nn ** 2)
    lbar = ebar * cj(ein("abckl,ijkl->abcij", l, h) + ein("abcij,ijkl->abckl", l, h))
    s = l.shape[3]
    lbar = lbar + 2 * nnbar * ein("ijk,ab->ijkab", cj(tr), np.eye(s))
    # l = E4 E3
    E4bar = ein("ijklm,efk->iejflm", lbar, cj(E3))
    E3bar = ein("iejflm,ijklm->efk", cj(E4), lbar)
    # E4 = E2 apn
    E2bar = ein("iejflm,cjfdlm->icde", E4bar, cj(apn))
    apnbar = ein("icde,iejflm->cjfdlm", cj(E2), E4bar)
    # E3 = C T
    Cbar = ein("efk,gfk->eg", E3bar, cj(T))
    Tbar = ein("eg,efk->gfk", cj(C), E3bar)
    # E2 = E1 T
    E1bar = ein("icde,bde->icb", E2bar, cj(T))
    Tbar = Tbar + ein("icb,icde->bde", cj(E1), E2bar)
    # E1 = T C
    Tbar = Tbar + ein("icb,ab->ica", E1bar, cj(C))
    Cbar = Cbar + ein("ica,icb->ab", cj(T), E1bar)
    apbar = _norm_back(ap, apnbar)
    return apbar, Cbar, Tbar


def energy(h, bulk, chi, tol, maxit, return_state=False):
    """src/variationalipeps.jl:28-40."""
    b = indexperm_symmetrize(bulk)
    ap, a = double_layer(b)
    C0, T0 = init_raw(a, chi)
    res = ctmrg(a, C0, T0, tol, maxit)
    e = expectationvalue(h, ap, res.C, res.T)
    if return_state:
        return e, res
    return e


def energy_frozen_init(h, bulk, chi, tol, maxit, env0):
    """energy() with the initial environment held fixed at ``env0 = (C0, T0)``: the
    function whose derivative Zygote computes, since _initializect_square is
    @nograd (src/autodiff.jl:5).  Central FD of this function pins the reverse
    sweep at any iteration count."""
    b = indexperm_symmetrize(bulk)
    ap, a = double_layer(b)
    res = ctmrg(a, env0[0], env0[1], tol, maxit)
    return expectationvalue(h, ap, res.C, res.T)


def raw_env_of(bulk, chi):
    """The :raw initial environment of ``energy`` for this bulk."""
    return init_raw(double_layer(indexperm_symmetrize(bulk))[1], chi)


def energy_and_grad(h, bulk, chi, tol, maxit):
    """Energy and d energy / d bulk of the unrolled program (what Zygote computes for
    src/variationalipeps.jl:73): the :raw init is @nograd (src/autodiff.jl:5), vals
    only feed the @nograd StopFunction (src/autodiff.jl:4).  For a complex bulk the energy is
    complex (e/n) and the gradient is that of Re(energy) (src/variationalipeps.jl:71), in Zygote's
    convention dRe(e)/dRe x + i dRe(e)/dIm x."""
    b = indexperm_symmetrize(bulk)
    ap, a = double_layer(b)
    C0, T0 = init_raw(a, chi)
    res = ctmrg(a, C0, T0, tol, maxit, keep=True)
    e = expectationvalue(h, ap, res.C, res.T)
    apbar, Cbar, Tbar = expectationvalue_back(h, ap, res.C, res.T)
    abar = np.zeros_like(a)
    for cache in reversed(res.caches):
        Cbar, Tbar, ab = ctmrgstep_back(a, cache, Cbar, Tbar)
        abar = abar + ab
    bbar = double_layer_back(b, apbar, abar)
    g = indexperm_symmetrize_back(bulk, bbar)
    return e, g, res.iters


def magnetisation(beta, chi, rng, tol=1e-6, maxit=100):
    """src/exampletensors.jl:57-69 (:random init drawn from ``rng``)."""
    a = model_tensor_ising(beta)
    m = mag_tensor_ising(beta)
    C0, T0 = init_random(a, chi, rng)
    res = ctmrg(a, C0, T0, tol, maxit)
    C, T = res.C, res.T
    ctc = ein("ia,ajb,bk->ijk", C, T, C)
    env = ein("alc,ckd,bjd,bia->ijkl", ctc, T, ctc, T)
    mag = ein("ijkl,ijkl->", env, m)
    nrm = ein("ijkl,ijkl->", env, a)
    return abs(mag / nrm)


# --------------------------------------------------------------------------
# TRG
# --------------------------------------------------------------------------
def trg_svd(t, dmax, tol, keep=False):
    """src/trg.jl:33-44.  Keep count = min(#{s > tol} + 1, dmax, len(s)) (searchsortedfirst
    on the descending s finds the first index with s <= tol)."""
    d1, d2, d3, d4 = t.shape
    tmat = reshape_f(t, (d1 * d2, d3 * d4))
    U, s, Vh = np.linalg.svd(tmat, full_matrices=False)
    first_le = int(np.sum(s > tol)) + 1            # 1-based searchsortedfirst(s, tol, rev=true)
    k = min(first_le, dmax, len(s))
    sq = np.sqrt(s[:k])
    u = reshape_f(U[:, :k] * sq[None, :], (d1, d2, k))
    v = reshape_f(Vh[:k, :] * sq[:, None], (k, d3, d4))
    if keep:
        return u, v, (U, s, Vh.conj().T, k, t.shape)
    return u, v


def trg_svd_back(cache, ubar, vbar):
    """Reverse of trg_svd through svd_back (dU and dV both present).  Real or complex (cotangent
    convention of double_layer_back): u = U sqrt(s) and v = sqrt(s) V' give dU = ubar sqrt(s),
    dV' = vbar sqrt(s) and, for the real s, dS = Re(<U, ubar> + <V', vbar>) / (2 sqrt(s))."""
    U, s, V, k, shape = cache
    d1, d2, d3, d4 = shape
    sq = np.sqrt(s[:k])
    ub = reshape_f(ubar, (d1 * d2, k))
    vb = reshape_f(vbar, (k, d3 * d4))
    Vh = V.conj().T
    dU = np.zeros_like(U)
    dU[:, :k] = ub * sq[None, :]
    dS = np.zeros_like(s)
    sqbar = np.real(np.sum(ub.conj() * U[:, :k], axis=0) + np.sum(vb.conj() * Vh[:k, :], axis=1))
    dS[:k] = sqbar / (2 * sq)
    dVh = np.zeros_like(Vh)
    dVh[:k, :] = vb * sq[:, None]
    tbar = svd_back(U, s, V, dU, dS, dVh.conj().T)
    return reshape_f(tbar, shape)


def trg(a, chi, niter, tol=1e-16, with_grad_seed=None):
    """src/trg.jl:13-30: ln Z per site.  If ``with_grad_seed`` is an array da0 (the
    derivative of the input tensor along a scalar parameter), also return
    d lnZ along that direction via the reverse sweep."""
    lnz = 0.0
    tape = []
    for n in range(1, niter + 1):
        idx = int(np.argmax(np.abs(a)))
        maxval = float(np.abs(a).flat[idx])
        a_in = a
        a = a / maxval
        lnz += 2.0 ** (1 - n) * math.log(maxval)
        dr_ul = ein("urdl->drul", a)
        ld_ru = ein("urdl->ldru", a)
        dr, ul, c1 = trg_svd(dr_ul, chi, tol, keep=True)
        ld, ru, c2 = trg_svd(ld_ru, chi, tol, keep=True)
        a = ein("npu,por,dom,lmn->urdl", dr, ld, ul, ru)
        tape.append((a_in, idx, maxval, c1, c2, dr, ul, ld, ru))
    trace = ein("ijij->", a)
    cplx = np.iscomplexobj(a)
    lnz += (np.log(trace) if cplx else math.log(trace)) / 2.0 ** niter
    if with_grad_seed is None:
        return lnz
    # reverse sweep of Re(lnZ) (the loss of test/trg.jl:26-27 for complex tensors; lnZ itself for real)
    cj = np.conj
    abar = np.zeros_like(a)
    for i in range(a.shape[0]):
        for j in range(a.shape[1]):
            abar[i, j, i, j] = 1.0
    abar *= cj(1.0 / (2.0 ** niter * trace))
    for n in range(niter, 0, -1):
        a_in, idx, maxval, c1, c2, dr, ul, ld, ru = tape[n - 1]
        drbar = ein("urdl,por,dom,lmn->npu", abar, cj(ld), cj(ul), cj(ru))
        ldbar = ein("urdl,npu,dom,lmn->por", abar, cj(dr), cj(ul), cj(ru))
        ulbar = ein("urdl,npu,por,lmn->dom", abar, cj(dr), cj(ld), cj(ru))
        rubar = ein("urdl,npu,por,dom->lmn", abar, cj(dr), cj(ld), cj(ul))
        g1 = trg_svd_back(c1, drbar, ulbar)   # d/d dr_ul
        g2 = trg_svd_back(c2, ldbar, rubar)   # d/d ld_ru
        anbar = ein("drul->urdl", g1) + ein("ldru->urdl", g2)
        # a_n = a_in / maxval ; lnz += 2^(1-n) log(maxval); maxval = |a_in[idx]| (real)
        mbar = 2.0 ** (1 - n) / maxval - np.real(np.sum(anbar.conj() * a_in)) / maxval ** 2
        abar = anbar / maxval
        flat = abar.reshape(-1)
        z = a_in.flat[idx]
        flat[idx] += mbar * (z / abs(z))
    return lnz, float(np.real(np.sum(abar.conj() * with_grad_seed)))


def trg_and_grad(beta, chi, niter, tol=1e-16):
    """ln Z(beta) per site and d lnZ / d beta through model_tensor (README.md:67-70)."""
    return trg(model_tensor_ising(beta), chi, niter, tol, with_grad_seed=dmodel_tensor_ising(beta))


# --------------------------------------------------------------------------
# finite differences
# --------------------------------------------------------------------------
def num_grad(f, x, delta=1e-5):
    """src/autodiff.jl:44,58-63: central differences (f(x+d/2) - f(x-d/2)) / d, elementwise."""
    if np.isscalar(x):
        return (f(x + delta / 2) - f(x - delta / 2)) / delta
    g = np.zeros_like(x, dtype=float)
    for i in np.ndindex(x.shape):
        xp = x.copy()
        xm = x.copy()
        xp[i] += delta / 2
        xm[i] -= delta / 2
        g[i] = (f(xp) - f(xm)) / delta
    return g


def symmetrized_random_bulk(D, s, seed, dist="uniform"):
    """Seeded bulk tensor as used by the workloads: NumPy PCG64 draw, column-major
    (D,D,D,D,s).  Not symmetrised here (energy() symmetrises)."""
    rng = np.random.default_rng(seed)
    if dist == "uniform":
        x = rng.random(D ** 4 * s)
    else:
        x = rng.standard_normal(D ** 4 * s)
    return reshape_f(x, (D, D, D, D, s))
