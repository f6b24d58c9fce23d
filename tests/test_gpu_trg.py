"""GPU parity of TRG (src/trg.jl) through the C ABI: the reference's golden
values, the oracle's gradient, and the trg_svd split."""
import numpy as np
import pytest

from oracle import tnad_oracle as O

pytestmark = pytest.mark.gpu


def _dlnz(tnad, ctx, beta, chi, niter):
    lnz, abar = tnad.trg_and_grad(tnad.model_tensor(tnad.Ising(), beta), chi, niter, ctx=ctx)
    return lnz, float(np.sum(abar * tnad.dmodel_tensor(tnad.Ising(), beta)))


def test_trg_tensorgrad_golden_gpu(ctx):
    import tnad
    print("case: trg beta=0.4 chi=5 niter=5", flush=True)
    # test/trg.jl:18
    assert tnad.trg(tnad.model_tensor(tnad.Ising(), 0.4), 5, 5, ctx=ctx) == pytest.approx(0.8919788686747141,
                                                                                          rel=1e-12, abs=0)


def test_trg_userguide_gradient_gpu(ctx):
    import tnad
    print("case: trg grad beta=0.5 chi=5 niter=5", flush=True)
    lnz, g = _dlnz(tnad, ctx, 0.5, 5, 5)
    assert g == pytest.approx(1.7502426939979507, rel=1e-10)     # docs/src/userguide.md:28
    lnz_o, g_o = O.trg_and_grad(0.5, 5, 5)
    assert lnz == pytest.approx(lnz_o, rel=1e-13, abs=0)


def test_trg_readme_config_a_gpu(ctx):
    import tnad
    print("case: trg grad beta=0.5 chi=20 niter=20", flush=True)
    lnz, g = _dlnz(tnad, ctx, 0.5, 20, 20)
    assert lnz == pytest.approx(1.0257933734351765, rel=1e-12, abs=0)   # README.md:60
    assert g == pytest.approx(1.7455677143228514, rel=1e-9)            # README.md:70
    assert abs(lnz - O.onsager_lnz(0.5)) < 1e-6


def test_trg_chi32_matches_oracle(ctx):
    # splits of 1024 x 1024 (augmented 2048): the truncated subspace eigensolver, not full Jacobi
    import tnad
    print("case: trg grad beta=0.4 chi=32 niter=8", flush=True)
    lnz, g = _dlnz(tnad, ctx, 0.4, 32, 8)
    lnz_o, g_o = O.trg_and_grad(0.4, 32, 8)
    print(f"lnz {lnz:.15f} (oracle {lnz_o:.15f}), dlnz/dbeta {g:.12f} (oracle {g_o:.12f})", flush=True)
    assert lnz == pytest.approx(lnz_o, rel=1e-12, abs=0)
    assert g == pytest.approx(g_o, rel=1e-8)


def test_trg_chi64_approaches_onsager(ctx):
    # the largest supported bond (splits of 4096 x 4096, augmented 8192); the oracle would take minutes
    # here, so the check is Onsager's exact free energy and its beta-derivative: the oracle's errors at
    # beta = 0.4, 20 iterations are 7.6e-7 / 3.4e-5 at chi = 16 and 3.5e-8 / 1.9e-6 at chi = 32
    import tnad
    print("case: trg grad beta=0.4 chi=64 niter=20", flush=True)
    beta, h = 0.4, 1e-4
    exact = O.onsager_lnz(beta)
    dexact = (O.onsager_lnz(beta + h) - O.onsager_lnz(beta - h)) / (2 * h)
    lnz, g = _dlnz(tnad, ctx, beta, 64, 20)
    print(f"lnz - exact {lnz - exact:.3e}, dlnz/dbeta - exact {g - dexact:.3e}", flush=True)
    assert abs(lnz - exact) < 1.5e-8
    assert abs(g - dexact) < 1e-6


@pytest.mark.parametrize("chi,niter", [(6, 3), (8, 4)])
def test_trg_full_gradient_matches_oracle_direction(ctx, chi, niter):
    # a generic full-rank tensor: every singular value is resolved, so the reference's
    # svd_back is well defined along every direction (for the rank-2 Ising tensor the
    # directional derivative off the physical family depends on LAPACK's arbitrary
    # null-space vectors through dS / (2 sqrt(s)) at s ~ 1e-17)
    import tnad
    print(f"case: trg full gradient random direction generic tensor chi={chi} niter={niter}", flush=True)
    rng = np.random.default_rng(7)
    a = rng.uniform(0.5, 1.5, (3, 3, 3, 3))
    direc = np.random.default_rng(chi).standard_normal(a.shape)
    lnz, abar = tnad.trg_and_grad(a, chi, niter, ctx=ctx)
    lnz_o, g_o = O.trg(a, chi, niter, with_grad_seed=direc)
    assert lnz == pytest.approx(lnz_o, rel=1e-13, abs=0)
    assert float(np.sum(abar * direc)) == pytest.approx(g_o, rel=1e-9)


@pytest.mark.parametrize("shape", [(6, 6, 6, 6), (2, 3, 4, 5)])
def test_trg_svd_reconstruction_gpu(ctx, shape):
    import tnad
    print(f"case: trg_svd {shape}", flush=True)
    rng = np.random.default_rng(1)
    t = rng.standard_normal(shape)
    u, v = tnad.trg_svd(t, 100, 0.0, ctx=ctx)
    assert np.allclose(np.einsum("ija,akl->ijkl", u, v), t, atol=1e-11)


def test_trg_svd_keep_count_gpu(ctx):
    import tnad
    print("case: trg_svd keep count", flush=True)
    rng = np.random.default_rng(1)
    t = np.reshape(rng.standard_normal((4, 2)) @ rng.standard_normal((2, 4)), (2, 2, 2, 2), order="F")
    u, v = tnad.trg_svd(t, 10, 1e-8, ctx=ctx)
    assert u.shape[2] == 3 and v.shape[0] == 3       # src/trg.jl:37 keeps #{s > tol} + 1
    assert np.allclose(np.einsum("ija,akl->ijkl", u, v), t, atol=1e-12)


# pullback of trg_svd on square and non-square splits (the projector terms of src/trg.jl:98-104), with and
# without truncation: a loss invariant under the singular vectors' signs, so the GPU's gauge and the
# oracle's give the same gradient; the oracle's is pinned by central differences in test_oracle.py
@pytest.mark.parametrize("shape,dmax", [((2, 3, 3, 1), 10), ((3, 1, 2, 3), 10), ((2, 3, 3, 1), 2),
                                        ((3, 1, 2, 3), 2), ((2, 3, 2, 3), 4), ((4, 5, 3, 3), 6),
                                        ((6, 6, 6, 6), 12), ((4, 8, 2, 5), 7)])
def test_trg_svd_back_matches_oracle(ctx, shape, dmax):
    import tnad
    print(f"case: trg_svd_back {shape} dmax={dmax}", flush=True)
    rng = np.random.default_rng(5)
    t = rng.standard_normal(shape)
    d1, d2, d3, d4 = shape
    m, mp = d1 * d2, d3 * d4
    A = rng.standard_normal((m, m))
    A = A + A.T
    B = rng.standard_normal((mp, mp))
    B = B + B.T

    def cotangents(u, v):
        k = u.shape[2]
        um, vm = O.reshape_f(u, (m, k)), O.reshape_f(v, (k, mp))
        return O.reshape_f(2 * A @ um, (d1, d2, k)), O.reshape_f(2 * vm @ B, (k, d3, d4))

    uo, vo, cache = O.trg_svd(t, dmax, 0.0, keep=True)
    g_o = O.trg_svd_back(cache, *cotangents(uo, vo))
    ug, vg = tnad.trg_svd(t, dmax, 0.0, ctx=ctx)
    assert ug.shape == uo.shape and vg.shape == vo.shape
    g = tnad.trg_svd_back(t, dmax, 0.0, *cotangents(ug, vg), ctx=ctx)
    err = np.max(np.abs(g - g_o)) / np.max(np.abs(g_o))
    print(f"max |g - g_oracle| / max |g_oracle| = {err:.2e}", flush=True)
    assert err < 1e-10


def test_trg_svd_back_rejects_a_wrong_kept_count(ctx):
    import tnad
    print("case: trg_svd_back wrong k", flush=True)
    t = np.random.default_rng(2).standard_normal((2, 3, 3, 1))
    u, v = tnad.trg_svd(t, 10, 0.0, ctx=ctx)
    with pytest.raises(tnad.TnadError):
        tnad.trg_svd_back(t, 10, 0.0, u[:, :, :2], v[:2], ctx=ctx)
