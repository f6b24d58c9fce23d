"""Pins of the learned warm start's CPU oracle (oracle/warmstart_oracle.py): the network
gradient against central finite differences, and the canonical environment gauge against
the symmetries it must remove.  CPU only."""
import numpy as np
import pytest

from oracle import tnad_oracle as O
from oracle import warmstart_oracle as W


def test_mlp_gradient_matches_central_differences():
    rng = np.random.default_rng(0)
    shapes = W.mlp_shapes(5, 4, 2, 3)
    n = sum(int(np.prod(s)) for s in shapes)
    theta = 0.5 * rng.standard_normal(n)
    X = rng.standard_normal((5, 3))
    T = rng.standard_normal((3, 3))
    loss, g = W.mlp_loss_and_grad(theta, shapes, X, T)
    fd = np.zeros(n)
    h = 1e-6
    for i in range(n):
        tp, tm = theta.copy(), theta.copy()
        tp[i] += h / 2
        tm[i] -= h / 2
        fd[i] = (W.mlp_loss_and_grad(tp, shapes, X, T)[0] - W.mlp_loss_and_grad(tm, shapes, X, T)[0]) / h
    assert np.max(np.abs(g - fd)) < 1e-8 * max(1.0, np.max(np.abs(fd)))
    # the loss is half the batch-mean squared error of the forward pass
    Y, _ = W.mlp_forward(theta, shapes, X)
    assert loss == pytest.approx(0.5 * np.sum((Y - T) ** 2) / 3, rel=1e-14)


def test_adam_first_step_is_signed_lr():
    # with m = v = 0 and t = 1 the bias-corrected step is lr * g / (|g| + eps)
    g = np.array([2.0, -3.0, 0.5])
    th, m, v = W.adam_step(np.zeros(3), g, np.zeros(3), np.zeros(3), 1, 0.1)
    assert np.allclose(th, -0.1 * g / (np.abs(g) + 1e-8), rtol=1e-12)


def _converged_env(seed, D=2, chi=6):
    b = O.symmetrized_random_bulk(D, 2, seed, "normal")
    ap, a = O.double_layer(O.indexperm_symmetrize(b))
    C0, T0 = O.init_raw(a, chi)
    res = O.ctmrg(a, C0, T0, 1e-12, 200)
    return b, res.C, res.T


def test_canonical_features_are_gauge_invariant():
    _, C, T = _converged_env(3)
    chi = C.shape[0]
    f0 = W.env_features(C, T)
    S, _ = np.linalg.qr(np.random.default_rng(1).standard_normal((chi, chi)))
    Cg = S.T @ C @ S
    Tg = np.einsum("xi,xay,yk->iak", S, T, S)
    assert np.allclose(W.env_features(Cg, Tg), f0, atol=1e-10)
    # sign gauges too
    s = np.where(np.random.default_rng(2).random(chi) < 0.5, -1.0, 1.0)
    assert np.allclose(W.env_features(s[:, None] * C * s[None, :], s[:, None, None] * T * s[None, None, :]), f0,
                       atol=1e-12)
    # and the global signs of C and T
    for sc, st in ((-1, 1), (1, -1), (-1, -1)):
        assert np.allclose(W.env_features(sc * C, st * T), f0, atol=1e-12)


def test_bulk_features_sign_and_scale_invariant():
    b = O.symmetrized_random_bulk(3, 2, 4, "normal")
    f = W.bulk_features(b)
    assert f.shape == (9 ** 4,)
    assert np.allclose(W.bulk_features(-b), f, atol=1e-15)
    assert np.allclose(W.bulk_features(3.0 * b), f, atol=1e-15)
    assert np.linalg.norm(f) == pytest.approx(1.0, abs=1e-14)


def test_decoded_environment_keeps_the_gauge_invariant_content():
    _, C, T = _converged_env(5)
    chi, Dd, _ = T.shape
    C0, T0 = W.env_from_features(W.env_features(C, T), chi, Dd)
    assert np.allclose(np.sort(np.abs(np.linalg.eigvalsh(C0))), np.sort(np.abs(np.linalg.eigvalsh(C))), atol=1e-12)
    gram = lambda t: np.einsum("iak,ibk->ab", t, t)
    assert np.allclose(gram(T0), gram(T), atol=1e-12)


def test_ctmrg_from_decoded_environment_reaches_the_same_energy():
    b, C, T = _converged_env(6, chi=8)
    h = O.hamiltonian_heisenberg()
    e_raw = O.energy(h, b, 8, 1e-12, 200)
    chi, Dd, _ = T.shape
    env0 = W.env_from_features(W.env_features(C, T), chi, Dd)
    assert O.energy_frozen_init(h, b, 8, 1e-12, 200, env0) == pytest.approx(e_raw, abs=1e-9)
