"""GPU parity of the CTMRG / energy path: libtnad (through the C ABI) against the
CPU oracle on the same seeded inputs.  Gauge-invariant quantities only (vals,
energies, gradients with respect to the bulk, projectors), since eigenvector
signs are a free gauge (C -> S C S, T -> S T S)."""
import sys

import numpy as np
import pytest

from oracle import tnad_oracle as O

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _bulk(seed, d=2, s=2, dist="normal"):
    return O.symmetrized_random_bulk(d, s, seed, dist)


# --- front of energy(): symmetrisation, double layer, :raw init ----------------------
def test_symmetrize_double_layer_init_raw(ctx):
    import tnad
    _say("case: symmetrize/double_layer/init_raw d=3")
    b = _bulk(0, d=3)
    bs = tnad.indexperm_symmetrize(tnad.SquareIPEPS(b), ctx=ctx).bulk
    assert np.allclose(bs, O.indexperm_symmetrize(b), atol=1e-14)
    ap, a = tnad.double_layer(tnad.SquareIPEPS(bs), ctx=ctx)
    ap_o, a_o = O.double_layer(bs)
    assert np.allclose(ap, ap_o, atol=1e-15) and np.allclose(a, a_o, atol=1e-15)
    rt = tnad.SquareCTMRGRuntime.new(a, "raw", 12, ctx=ctx)
    C0, T0 = O.init_raw(a_o, 12)
    assert np.allclose(rt.corner, C0, atol=1e-14) and np.allclose(rt.edge, T0, atol=1e-14)


@pytest.mark.parametrize("D,chi", [(2, 4), (5, 3), (4, 4)])
def test_init_raw_non_symmetric_and_truncated(ctx, D, chi):
    # the reference's doctest input (src/ctmrg.jl:52-60: randn(2,2,2,2), chi = 4) and D > chi, where
    # src/ctmrg.jl:79-84 drops the entries outside the chi-sized arrays; non-symmetric inputs, so an
    # index-order mistake on either side shows
    import tnad
    _say(f"case: init_raw non-symmetric D={D} chi={chi}")
    a = np.random.default_rng(100 + D).standard_normal((D, D, D, D))
    rt = tnad.SquareCTMRGRuntime.new(a, "raw", chi, ctx=ctx)
    C0, T0 = O.init_raw(a, chi)
    m = min(D, chi)
    assert np.allclose(rt.corner[:m, :m], a.sum(axis=(2, 3))[:m, :m], atol=1e-13)
    assert np.allclose(rt.edge[:m, :, :m], a.sum(axis=3)[:m, :, :m], atol=1e-13)
    assert np.array_equal(rt.corner, C0) or np.allclose(rt.corner, C0, atol=1e-14)
    assert np.allclose(rt.edge, T0, atol=1e-14)


def test_unsupported_sizes_are_rejected_before_launch(ctx):
    # host-side limits (csrc/tnad_limits.h): the kept block of the eigensolver is at most 192 wide above
    # n = 192; these calls must fail with TNAD_ERR_UNSUPPORTED before anything is launched
    import tnad
    _say("case: size limits through the C ABI")
    n = 200
    A = np.random.default_rng(0).standard_normal((n, n))
    A = A + A.T
    with pytest.raises(tnad.TnadError) as ei:
        _symeig(ctx, A, n, 193)
    assert ei.value.code == tnad.ERR_UNSUPPORTED
    with pytest.raises(tnad.TnadError) as ei:
        tnad.trg_svd(np.ones((12, 12, 12, 12)), 96, 1e-16, ctx=ctx)
    assert ei.value.code == tnad.ERR_UNSUPPORTED


def test_dimension_mismatch_is_an_error(ctx):
    import tnad
    with pytest.raises(tnad.DimensionMismatch):
        tnad.SquareIPEPS(np.zeros((3, 3, 4, 3, 2)))


# --- eigensolver ---------------------------------------------------------------------
# n <= 2 is the one-slot case of the round-robin ordering; odd n is padded with a zero row and column
@pytest.mark.parametrize("n", [1, 2, 3, 4, 7, 80, 191, 192])
def test_symeig_full_matches_lapack(ctx, n):
    _say(f"case: symeig full n={n}")
    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, n))
    A = A + A.T
    lam, U = _symeig(ctx, A, n, n)
    ref = np.linalg.eigvalsh(A)
    ref = ref[np.argsort(-np.abs(ref), kind="stable")]
    assert np.allclose(lam, ref, atol=1e-12 * np.abs(ref).max())
    assert np.allclose(U.T @ U, np.eye(n), atol=1e-12)
    assert np.allclose(A @ U, U * lam[None, :], atol=1e-11 * np.abs(ref).max())


def _hard_matrix(kind, n, rng):
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    if kind == "graded":        # spectrum over ten decades, as a converged corner matrix's
        lam = np.sign(rng.standard_normal(n)) * np.logspace(0, -10, n)
    elif kind == "degenerate":  # exact four-fold multiplets (symmetry-degenerate spectra)
        lam = np.repeat(rng.standard_normal((n + 3) // 4), 4)[:n]
    elif kind == "clustered":   # many exactly equal and nearly equal values
        lam = np.round(rng.standard_normal(n), 1)
    elif kind == "diagonal":
        return np.diag(rng.standard_normal(n))
    elif kind == "zero":
        return np.zeros((n, n))
    else:                       # path graph: a tridiagonal matrix with unit couplings
        return np.diag(np.ones(n - 1), 1) + np.diag(np.ones(n - 1), -1)
    A = (Q * lam) @ Q.T
    return 0.5 * (A + A.T)


@pytest.mark.parametrize("n", [3, 17, 176, 192])
@pytest.mark.parametrize("kind", ["graded", "degenerate", "clustered", "diagonal", "zero", "path"])
def test_symeig_full_hard_spectra(ctx, kind, n):
    # the full-spectrum solver (Householder tridiagonalisation + divide and conquer) on spectra that
    # exercise its deflation (exact and near multiplets, tiny couplings) against LAPACK: residual,
    # orthogonality and eigenvalues to 1e-12 of the spectral norm
    _say(f"case: symeig full n={n} {kind}")
    A = _hard_matrix(kind, n, np.random.default_rng(100 + n))
    lam, U = _symeig(ctx, A, n, n)
    ref = np.linalg.eigvalsh(A)
    scale = max(np.abs(ref).max(), 1e-300)
    assert np.max(np.abs(np.sort(lam) - ref)) <= 1e-12 * scale
    assert np.allclose(U.T @ U, np.eye(n), atol=1e-12)
    assert np.max(np.abs(A @ U - U * lam[None, :])) <= 1e-12 * scale
    assert np.all(np.diff(np.abs(lam)) <= 0.0)   # sorted by |lambda| descending


def _symeig(ctx, A, want, k):
    import ctypes as C
    n = A.shape[0]
    A = np.asfortranarray(A)
    kk = n if n <= 192 else min(max(k, want), 192)
    ev = np.zeros(kk)
    U = np.zeros((n, kk), order="F")
    ko = C.c_int(0)
    ctx.check(ctx.lib.tnad_symeig(ctx.h, C.c_void_p(A.ctypes.data), n, want, k, C.c_void_p(ev.ctypes.data),
                                  C.c_void_p(U.ctypes.data), C.byref(ko)))
    return ev[:ko.value], U[:, :ko.value]


def test_symeig_topk_subspace(ctx):
    _say("case: symeig top-k n=600 k=96")
    rng = np.random.default_rng(1)
    n, want = 600, 64
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    lam = np.sign(rng.standard_normal(n)) * np.exp(-np.arange(n) / 25.0)
    A = (Q * lam) @ Q.T
    A = 0.5 * (A + A.T)
    ev, U = _symeig(ctx, A, want, 96)
    ref = lam[np.argsort(-np.abs(lam), kind="stable")]
    assert np.allclose(ev[:want], ref[:want], rtol=1e-10, atol=1e-13)
    P = U[:, :want] @ U[:, :want].T
    Qk = Q[:, np.argsort(-np.abs(lam), kind="stable")[:want]]
    assert np.allclose(P, Qk @ Qk.T, atol=1e-9)


# --- svd_back (eigen adjoint) ----------------------------------------------------------
@pytest.mark.parametrize("n,chi", [(40, 10), (216, 24)])
def test_svd_back_sym_matches_oracle(ctx, n, chi):
    import ctypes as C
    _say(f"case: svd_back_sym n={n} chi={chi}")
    rng = np.random.default_rng(n)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    lam = np.sign(rng.standard_normal(n)) * np.exp(-np.arange(n) / 6.0)
    A = np.asfortranarray(0.5 * ((Q * lam) @ Q.T + ((Q * lam) @ Q.T).T))
    H = rng.standard_normal((n, n))
    H = H + H.T
    # gauge-invariant loss tr(U1^T H U1): G = 2 H U1 in each implementation's own gauge
    _, Ug = _symeig(ctx, A, chi, 2 * chi)
    G = np.asfortranarray(2 * H @ Ug[:, :chi])
    out = np.zeros((n, n), order="F")
    ctx.check(ctx.lib.tnad_svd_back_sym(ctx.h, C.c_void_p(A.ctypes.data), n, chi, 2 * chi,
                                        C.c_void_p(G.ctypes.data), C.c_void_p(out.ctypes.data)))
    U, s, Vh = np.linalg.svd(A)
    dU = np.zeros_like(U)
    dU[:, :chi] = 2 * H @ U[:, :chi]
    Mb = O.svd_back(U, s, Vh.T, dU, None, None)
    ref = Mb + Mb.T
    assert np.max(np.abs(out - ref)) < 1e-8 * np.max(np.abs(ref))


# --- one step and the fixed point ---------------------------------------------------------
def test_ctmrgstep_config_b_vals(ctx):
    import tnad
    _say("case: ctmrgstep D=2 chi=20")
    b = O.indexperm_symmetrize(_bulk(2, dist="uniform"))
    ap, a = O.double_layer(b)
    C, T = O.init_raw(a, 20)
    rt = tnad.SquareCTMRGRuntime(a, C, T)
    Co, To, vo = C, T, None
    for it in range(4):
        rt, v = tnad.ctmrgstep(rt, ctx=ctx)
        Co, To, vo = O.ctmrgstep(a, Co, To)
        assert np.allclose(v, vo, atol=1e-12), it
        assert np.allclose(rt.corner, rt.corner.T, atol=1e-14)
        assert np.linalg.norm(rt.corner) == pytest.approx(1.0, abs=1e-14)
        # gauge-invariant content of the environment: spectrum of C and of T^T T
        assert np.allclose(np.sort(np.abs(np.linalg.eigvalsh(rt.corner))),
                           np.sort(np.abs(np.linalg.eigvalsh(Co))), atol=1e-12)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_energy_config_b(ctx, seed):
    # Heisenberg, D=2, d=2, uniform [0,1) bulk, chi=20, tol=1e-6, maxit=100
    import tnad
    _say(f"case: energy config B seed={seed}")
    h = O.hamiltonian_heisenberg()
    b = _bulk(seed, dist="uniform")
    e_o, st = O.energy(h, b, 20, 1e-6, 100, return_state=True)
    e, _, info = tnad.energy_and_grad(h, b, 20, 1e-6, 100, ctx=ctx, want_grad=False, return_info=True)
    assert info["iters"] == st.iters
    assert abs(e - e_o) < 1e-10


def test_product_state_energies_gpu(ctx):
    import tnad
    _say("case: product states")
    h = O.diaglocalhamiltonian([1.0, -1.0])
    rng = np.random.default_rng(5)
    a = np.zeros((2, 2, 2, 2, 2)) + 1e-12 * rng.standard_normal((2, 2, 2, 2, 2))
    a[0, 0, 0, 0, 1] = rng.standard_normal()
    assert tnad.energy(h, a, 4, 1e-12, 100, ctx=ctx) / 2 == pytest.approx(-1.0, abs=1e-9)
    a = np.zeros((2, 2, 2, 2, 2)) + 1e-12 * rng.standard_normal((2, 2, 2, 2, 2))
    a[0, 0, 0, 0, 0] = rng.standard_normal()
    assert tnad.energy(h, a, 10, 0.0, 300, ctx=ctx) / 2 == pytest.approx(1.0, abs=1e-9)


def test_expectationvalue_matches_oracle(ctx):
    import tnad
    _say("case: expectationvalue D=2 chi=8")
    h = O.hamiltonian_heisenberg()
    b = O.indexperm_symmetrize(_bulk(3))
    ap, a = O.double_layer(b)
    C, T = O.init_raw(a, 8)
    for _ in range(3):
        C, T, _v = O.ctmrgstep(a, C, T)
    e = tnad.expectationvalue(h, ap, tnad.SquareCTMRGRuntime(a, C, T), ctx=ctx)
    assert abs(e - O.expectationvalue(h, ap, C, T)) < 1e-13


# --- gradients -------------------------------------------------------------------------
@pytest.mark.parametrize("seed,chi,maxit", [(8, 4, 0), (8, 4, 5), (9, 6, 3), (10, 12, 8)])
def test_energy_gradient_full_spectrum(ctx, seed, chi, maxit):
    import tnad
    _say(f"case: energy grad seed={seed} chi={chi} maxit={maxit}")
    h = O.hamiltonian_heisenberg()
    b = _bulk(seed)
    e_o, g_o, _ = O.energy_and_grad(h, b, chi, 0.0, maxit)
    e, g = tnad.energy_and_grad(h, b, chi, 0.0, maxit, ctx=ctx)
    assert abs(e - e_o) < 1e-12
    assert np.max(np.abs(g - g_o)) < 1e-9 * max(1.0, np.max(np.abs(g_o)))


def test_energy_gradient_truncated_eigensolver(ctx):
    # chi*D = 216 > 192: block subspace iteration + resolvent CG for the rest of the spectrum in the adjoint
    import tnad
    _say("case: energy grad truncated D=3 chi=24")
    h = O.hamiltonian_heisenberg()
    b = _bulk(11, d=3)
    e_o, g_o, _ = O.energy_and_grad(h, b, 24, 0.0, 6)
    e, g = tnad.energy_and_grad(h, b, 24, 0.0, 6, ctx=ctx)
    assert abs(e - e_o) < 1e-10
    assert np.linalg.norm(g - g_o) / np.linalg.norm(g_o) < 1e-6


def test_ctmrg_tape_back_matches_energy_path(ctx):
    # tnad_ctmrg + tnad_ctmrg_back (the rrule pair on ctmrg) against the oracle's reverse sweep
    import tnad
    _say("case: ctmrg tape back D=2 chi=6")
    b = O.indexperm_symmetrize(_bulk(12))
    ap, a = O.double_layer(b)
    C0, T0 = O.init_raw(a, 6)
    rng = np.random.default_rng(0)
    Cbar = rng.standard_normal((6, 6))
    Tbar = rng.standard_normal((6, 4, 6))
    res = O.ctmrg(a, C0, T0, 0.0, 5, keep=True)
    abar_o = np.zeros_like(a)
    Cb, Tb = Cbar, Tbar
    for cache in reversed(res.caches):
        Cb, Tb, ab = O.ctmrgstep_back(a, cache, Cb, Tb)
        abar_o += ab
    rt, info, tape = tnad.ctmrg(tnad.SquareCTMRGRuntime(a, C0, T0), 0.0, 5, ctx=ctx, keep_tape=True)
    # the cotangent is gauge dependent: express it in the oracle's gauge by matching signs of C
    assert info["iters"] == res.iters == 6   # maxit = 5: six steps (StopFunction counter from -1)
    abar = tape.back(*_gauge_match(rt, res, Cbar, Tbar))
    tape.free()
    assert np.max(np.abs(abar - abar_o)) < 1e-9 * np.max(np.abs(abar_o))


def _gauge_match(rt, res, Cbar, Tbar):
    """Transform the cotangents (given in the oracle's gauge) into the GPU's gauge:
    C_gpu = S C_o S, T_gpu = S T_o S with S = diag(+-1) found from the two outputs."""
    Tg, To = rt.edge, res.T
    chi = To.shape[0]
    k0 = int(np.argmax([np.linalg.norm(To[:, :, k]) for k in range(chi)]))
    s = np.ones(chi)
    for i in range(chi):
        j = int(np.argmax(np.abs(To[i, :, k0])))
        s[i] = 1.0 if Tg[i, j, k0] * To[i, j, k0] >= 0 else -1.0
    for _ in range(2):
        for i in range(chi):
            r = np.sum(Tg[i] * To[i] * s[None, :])
            s[i] = 1.0 if r >= 0 else -1.0
    assert np.allclose(s[:, None, None] * To * s[None, None, :], Tg, atol=1e-9)
    assert np.allclose(s[:, None] * res.C * s[None, :], rt.corner, atol=1e-9)
    return s[:, None] * Cbar * s[None, :], s[:, None, None] * Tbar * s[None, None, :]


# --- mixed precision -----------------------------------------------------------------------
def test_mixed_fp32_energy_and_gradient_within_budget(ctx):
    # FP32 contractions while the singular values change by more than switch_tol, then FP64 steps (their
    # eigensolver tolerance tied to the step-to-step change); FP32 reverse sweep.  Energy within 1e-7 and
    # gradient within 1e-5 (relative) of the plain FP64 path, the mixed-precision error budget.
    import tnad
    _say("case: mixed fp32 D=3 chi=24")
    h = O.hamiltonian_heisenberg()
    b = _bulk(13, d=3)
    e64, g64 = tnad.energy_and_grad(h, b, 24, 1e-10, 100, ctx=ctx)
    e32, g32, info = tnad.energy_and_grad(h, b, 24, 1e-10, 100, ctx=ctx, return_info=True, **tnad.MIXED_FP32)
    assert info["low_prec_iters"] > 0
    rel = np.linalg.norm(g32 - g64) / np.linalg.norm(g64)
    _say(f"fp32 mixed: |dE| {abs(e32 - e64):.2e}, gradient relative difference {rel:.2e}")
    assert abs(e32 - e64) < 1e-7
    assert rel < 1e-5
    e_o = O.energy(h, b, 24, 1e-10, 100)
    assert abs(e64 - e_o) < 1e-9


def test_mixed_bf16_energy_and_gradient_within_budget(ctx):
    # BF16 operands on MFMA (FP32 accumulation) while the singular values change by more than 1e-2, then
    # FP64 steps to tol; the FP32 reverse sweep.  Energy within 1e-7 and gradient within 1e-5 (relative)
    # of the plain FP64 path, the mixed-precision error budget.
    import tnad
    _say("case: mixed bf16 D=3 chi=24")
    h = O.hamiltonian_heisenberg()
    b = _bulk(13, d=3)
    e64, g64 = tnad.energy_and_grad(h, b, 24, 1e-10, 100, ctx=ctx)
    e16, g16, info = tnad.energy_and_grad(h, b, 24, 1e-10, 100, ctx=ctx, return_info=True,
                                          bwd_precision="fp32", **tnad.MIXED_BF16)
    assert info["low_prec_iters"] > 0
    rel = np.linalg.norm(g16 - g64) / np.linalg.norm(g64)
    _say(f"bf16 mixed: |dE| {abs(e16 - e64):.2e}, gradient relative difference {rel:.2e}")
    assert abs(e16 - e64) < 1e-7
    assert rel < 1e-5
