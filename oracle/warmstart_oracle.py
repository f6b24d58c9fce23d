"""CPU oracle of the learned CTMRG warm start.

TEST INFRASTRUCTURE ONLY.  Only ``tests/`` may import this module, as the
checker of the GPU kernels behind ``tnad_mlp_*`` / ``tnad_env_*``.

The reference has no learned initialisation: its environments start from
``_initializect_square`` (src/ctmrg.jl:62-86, ``:raw`` or ``:random``), which is
``@nograd`` (src/autodiff.jl:5).  The warm start replaces that constant
starting point by a prediction; CTMRG itself still runs to the same tolerance,
so the fixed point (and the gradient, which ignores the initial environment)
is unchanged.  What this module restates is the build's own definition:

  * features of a bulk tensor A: the double layer a = sum_s A (x) conj(A) of
    the ``indexperm_symmetrize``d bulk (src/ipeps.jl:32-39,
    src/variationalipeps.jl:30-34), normalised -- the environment depends on A
    only through a, so this parameter-free first stage removes the A -> -A
    ambiguity and hands the network the quantity CTMRG actually sees;
  * the canonical gauge of an environment (C, T): C = O diag(lam) O^T with
    |lam| descending, T~ = O^T T O on both chi legs, each column of O signed so
    that its weighted coupling to the earlier columns,
    sum_{r<i, a} (a + 1) / (r + 1) T~[r, a, i], is >= 0 (column 0 by the
    largest-|.| component of O[:, 0]), after the global signs of C and T are
    fixed by lam[0] >= 0 and sum_a (a + 1) T~[0, a, 0] >= 0 (a CTMRG fixed point
    is one only up to C -> -C and T -> -T); the features are (lam, T~);
  * decoding: C0 = diag(lam), T0 = (T~ + T~^{kji}) / 2, both normalised, which
    is a valid CTMRG starting environment in any gauge (one step maps gauge
    copies of an environment to the same output);
  * a fully connected network (tanh hidden layers, linear output) trained on
    the mean over the batch of 0.5 |y - t|^2 with Adam.

Pins (tests/test_warmstart_oracle.py): the network gradient against central
finite differences in FP64; the canonical features invariant under a random
orthogonal gauge S (C -> S^T C S, T -> S^T T S) and under A -> -A; decoding
reproduces the spectrum of C and the gauge-invariant Gram matrix of T.
"""
from __future__ import annotations

import numpy as np

from oracle import tnad_oracle as O


# ----------------------------------------------------------------------------- network
def mlp_shapes(n_in, hidden, n_layers, n_out):
    """parameter blocks in storage order: (W1, b1, ..., W_L, b_L, W_out, b_out);
    W matrices column-major (rows = outputs)"""
    shapes = [(hidden, n_in), (hidden,)]
    for _ in range(n_layers - 1):
        shapes += [(hidden, hidden), (hidden,)]
    shapes += [(n_out, hidden), (n_out,)]
    return shapes


def unflatten(theta, shapes):
    out, o = [], 0
    for sh in shapes:
        n = int(np.prod(sh))
        out.append(np.reshape(theta[o:o + n], sh, order="F"))
        o += n
    return out


def flatten(blocks):
    return np.concatenate([np.ravel(b, order="F") for b in blocks])


def mlp_forward(theta, shapes, X):
    """X: (n_in, B).  Returns Y (n_out, B) and the hidden activations."""
    P = unflatten(theta, shapes)
    acts = [X]
    h = X
    nl = len(P) // 2 - 1
    for l in range(nl):
        h = np.tanh(P[2 * l] @ h + P[2 * l + 1][:, None])
        acts.append(h)
    Y = P[-2] @ h + P[-1][:, None]
    return Y, acts


def mlp_loss_and_grad(theta, shapes, X, Tt):
    """loss = mean_b 0.5 |y_b - t_b|^2 and its gradient with respect to theta"""
    P = unflatten(theta, shapes)
    Y, acts = mlp_forward(theta, shapes, X)
    B = X.shape[1]
    R = Y - Tt
    loss = 0.5 * float(np.sum(R * R)) / B
    dY = R / B
    grads = [None] * len(P)
    grads[-2] = dY @ acts[-1].T
    grads[-1] = dY.sum(axis=1)
    dH = P[-2].T @ dY
    nl = len(P) // 2 - 1
    for l in range(nl - 1, -1, -1):
        dZ = dH * (1.0 - acts[l + 1] ** 2)
        grads[2 * l] = dZ @ acts[l].T
        grads[2 * l + 1] = dZ.sum(axis=1)
        dH = P[2 * l].T @ dZ
    return loss, flatten(grads)


def adam_step(theta, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """one Adam update (t = 1-based step count); returns new (theta, m, v)"""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    return theta - lr * mh / (np.sqrt(vh) + eps), m, v


# ----------------------------------------------------------------------------- features
def bulk_features(bulk):
    """network input of a bulk: the double layer a = sum_s A (x) conj(A) of the symmetrised bulk
    (src/variationalipeps.jl:30-34, src/ipeps.jl:32-39), unit norm, column-major flat"""
    _, a = O.double_layer(O.indexperm_symmetrize(bulk))
    x = np.ravel(a, order="F")
    return x / np.linalg.norm(x)


def env_canonical(C, T):
    """(lam, T~) of the canonical gauge (see module docstring); T~ flattened column-major"""
    chi, Dd, _ = T.shape
    C = 0.5 * (C + C.T)
    lam, Ov = np.linalg.eigh(C)
    order = np.argsort(-np.abs(lam), kind="stable")
    lam, Ov = lam[order], Ov[:, order]
    j0 = int(np.argmax(np.abs(Ov[:, 0])))
    if Ov[j0, 0] < 0:
        Ov[:, 0] = -Ov[:, 0]
    Tt = np.einsum("xi,xay,yk->iak", Ov, T, Ov)
    # global signs first: (C, T) -> (-C, T) and (C, -T) are fixed points too (a CTMRG step maps
    # them to each other's images), so fix lam[0] >= 0 and sum_a (a + 1) T~[0, a, 0] >= 0 (column
    # signs leave T~[0, :, 0] alone)
    if lam[0] < 0:
        lam = -lam
    if float(np.arange(1, Dd + 1) @ Tt[0, :, 0]) < 0:
        Tt = -Tt
    # then column i > 0 takes the sign that makes its weighted coupling to the earlier (already
    # signed) columns positive: sum_{r<i, a} (a + 1) / (r + 1) s_r T~[r, a, i] >= 0.  A weighted
    # sum rather than the largest entry: symmetry sectors of the double-layer index give exact
    # ties of opposite sign (T~[r, (x,y), i] = -T~[r, (y,x), i]) that an argmax would split
    # by rounding.
    sgn = np.ones(chi)
    wa = np.arange(1, Dd + 1, dtype=float)
    for i in range(1, chi):
        wr = sgn[:i] / np.arange(1, i + 1, dtype=float)
        if float(wr @ Tt[:i, :, i] @ wa) < 0:
            sgn[i] = -1.0
    Tt = sgn[:, None, None] * Tt * sgn[None, None, :]
    return lam, Tt


def env_features(C, T):
    lam, Tt = env_canonical(C, T)
    return np.concatenate([lam, np.ravel(Tt, order="F")])


def env_from_features(f, chi, Dd):
    lam = f[:chi]
    Tt = np.reshape(f[chi:], (chi, Dd, chi), order="F")
    C0 = np.diag(lam)
    T0 = 0.5 * (Tt + Tt.transpose(2, 1, 0))
    return C0 / np.linalg.norm(C0), T0 / np.linalg.norm(T0)
