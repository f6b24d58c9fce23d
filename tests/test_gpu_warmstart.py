"""GPU parity of the learned warm start (tnad_mlp_*, tnad_*_features) against the CPU oracle
(oracle/warmstart_oracle.py), and the iteration gain of a trained network over :raw."""
import numpy as np
import pytest

from oracle import tnad_oracle as O
from oracle import warmstart_oracle as W

pytestmark = pytest.mark.gpu


def _net(ctx, n_in, hidden, layers, n_out, seed=0):
    from tnad import warmstart as ws
    return ws.WarmStartNet(n_in, n_out, hidden=hidden, layers=layers, seed=seed, ctx=ctx)


def test_mlp_forward_and_gradient_match_oracle(ctx):
    print("case: mlp forward/gradient 7-16-16-11 batch 5", flush=True)
    net = _net(ctx, 7, 16, 2, 11)
    shapes = W.mlp_shapes(7, 16, 2, 11)
    rng = np.random.default_rng(0)
    theta = (0.5 * rng.standard_normal(net.num_params)).astype(np.float32)
    net.set_params(theta)
    assert np.array_equal(net.get_params(), theta)
    X = rng.standard_normal((7, 5)).astype(np.float32)
    T = rng.standard_normal((11, 5)).astype(np.float32)
    th64 = theta.astype(np.float64)
    Y_o, _ = W.mlp_forward(th64, shapes, X.astype(np.float64))
    assert np.max(np.abs(net.forward(X) - Y_o)) < 1e-5 * np.max(np.abs(Y_o))
    loss_o, g_o = W.mlp_loss_and_grad(th64, shapes, X.astype(np.float64), T.astype(np.float64))
    loss, g = net.loss_grad(X, T)
    assert loss == pytest.approx(loss_o, rel=1e-5)
    # FP32 arithmetic against the FP64 oracle
    assert np.max(np.abs(g - g_o)) < 2e-5 * np.max(np.abs(g_o))
    net.free()


def test_adam_step_matches_oracle(ctx):
    print("case: mlp adam step", flush=True)
    net = _net(ctx, 6, 8, 1, 4, seed=3)
    shapes = W.mlp_shapes(6, 8, 1, 4)
    theta = net.get_params().astype(np.float64)
    rng = np.random.default_rng(1)
    X = rng.standard_normal((6, 9)).astype(np.float32)
    T = rng.standard_normal((4, 9)).astype(np.float32)
    m = np.zeros_like(theta)
    v = np.zeros_like(theta)
    for t in (1, 2):
        loss_o, g_o = W.mlp_loss_and_grad(theta, shapes, X.astype(np.float64), T.astype(np.float64))
        loss = net.train_step(X, T, 1e-2)
        assert loss == pytest.approx(loss_o, rel=1e-5)
        theta, m, v = W.adam_step(theta, g_o, m, v, t, 1e-2)
    assert np.max(np.abs(net.get_params() - theta)) < 1e-5
    net.free()


def _converged(seed, D=2, chi=6):
    b = O.symmetrized_random_bulk(D, 2, seed, "normal")
    _, a = O.double_layer(O.indexperm_symmetrize(b))
    C0, T0 = O.init_raw(a, chi)
    res = O.ctmrg(a, C0, T0, 1e-12, 200)
    return b, res.C, res.T


def test_features_match_oracle_and_are_gauge_invariant(ctx):
    from tnad import warmstart as ws
    print("case: bulk/env features D=2 chi=6", flush=True)
    b, Cm, Tm = _converged(3)
    assert np.allclose(ws.bulk_features(b, ctx), W.bulk_features(b), atol=1e-7)
    f = ws.env_features(Cm, Tm, ctx)
    assert np.allclose(f, W.env_features(Cm, Tm), atol=1e-6)
    chi = Cm.shape[0]
    S, _ = np.linalg.qr(np.random.default_rng(4).standard_normal((chi, chi)))
    fg = ws.env_features(S.T @ Cm @ S, np.einsum("xi,xay,yk->iak", S, Tm, S), ctx)
    assert np.allclose(fg, f, atol=1e-6)
    fneg = ws.env_features(-Cm, -Tm, ctx)
    assert np.allclose(fneg, f, atol=1e-6)
    C0, T0 = ws.env_from_features(f, chi, Tm.shape[1], ctx)
    C0o, T0o = W.env_from_features(f.astype(np.float64), chi, Tm.shape[1])
    assert np.allclose(C0, C0o, atol=1e-12) and np.allclose(T0, T0o, atol=1e-12)


def test_learned_warm_start_gains_over_raw(ctx):
    # small configuration (D=2, chi=8): trained on 384 converged environments from this library's
    # CTMRG, evaluated on 16 held-out bulks.  The gain asserted is the distance of the first
    # CTMRG step's singular values from the fixed point's (gauge invariant); the iterations to
    # tol=1e-8 from both starts are printed; both starts reach the same fixed point.
    from tnad import warmstart as ws
    import tnad
    print("case: learned warm start D=2 chi=8 (384 train, 16 test)", flush=True)
    D, s, chi, tol = 2, 2, 8, 1e-8
    Dd = D * D
    X, Y, it_train = ws.generate_dataset(range(1000, 1384), D, s, chi, tol, 300, ctx=ctx,
                                         log=lambda m: print(m, flush=True))
    net = ws.WarmStartNet(X.shape[0], Y.shape[0], hidden=256, layers=2, seed=0, ctx=ctx)
    hist = ws.train(net, X, Y, epochs=150, batch=64, lr=1e-3, lr_decay=0.99, log=lambda m: print(m, flush=True))
    assert hist[-1] < 0.8 * hist[0]
    d_raw, d_learned = [], []
    for sd in range(50000, 50016):
        b = ws.random_bulk(D, s, sd)
        a = ws.double_layer_of(b, ctx)
        Cs, Ts, _ = ws.ctmrg_from(a, chi, 1e-12, 500, ctx=ctx)
        _, v_star = tnad.ctmrgstep(tnad.SquareCTMRGRuntime(a, Cs, Ts), ctx=ctx)
        _, v_raw = tnad.ctmrgstep(tnad.SquareCTMRGRuntime.new(a, "raw", chi, ctx=ctx), ctx=ctx)
        _, v_l = tnad.ctmrgstep(tnad.SquareCTMRGRuntime(a, *ws.predict_env(net, b, chi, Dd, ctx)), ctx=ctx)
        d_raw.append(np.linalg.norm(v_raw - v_star))
        d_learned.append(np.linalg.norm(v_l - v_star))
    print("first-step vals distance: raw %.4f learned %.4f" % (np.mean(d_raw), np.mean(d_learned)), flush=True)
    assert np.mean(d_learned) < np.mean(d_raw)
    raw, learned = ws.compare_iterations(net, range(50000, 50016), D, s, chi, tol, 300, ctx=ctx)
    print("iterations to 1e-8: raw", list(raw), "learned", list(learned), "means %.2f %.2f"
          % (raw.mean(), learned.mean()), flush=True)
    # the same fixed point: energies from both starts agree
    h = O.hamiltonian_heisenberg()
    for sd in (50000, 50001):
        b = ws.random_bulk(D, s, sd)
        a = ws.double_layer_of(b, ctx)
        C1, T1, _ = ws.ctmrg_from(a, chi, 1e-12, 500, env0=ws.predict_env(net, b, chi, Dd, ctx), ctx=ctx)
        ap, _ = tnad.double_layer(tnad.indexperm_symmetrize(tnad.SquareIPEPS(b), ctx=ctx), ctx=ctx)
        e_l = tnad.expectationvalue(h, ap, tnad.SquareCTMRGRuntime(a, C1, T1), ctx=ctx)
        e_r = tnad.energy(h, b, chi, 1e-12, 500, ctx=ctx)
        assert e_l == pytest.approx(e_r, abs=1e-8)
    net.free()
