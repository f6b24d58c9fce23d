"""Energy gradient at D=3, d=2, chi=64 (corner matrix 576 x 576: block subspace eigensolver in the
forward, resolvent CG in the adjoint) against central finite differences, h = 1e-6, on 32 random entries
of the bulk; plain FP64 and FP32 contractions (FP32 forward steps until the singular values settle, FP32
reverse-sweep contractions).  The function differentiated is the energy after a fixed number of CTMRG steps
(tol = 0, maxit = N: N + 1 steps under the reference's StopFunction counter), so the unrolled gradient
and the finite differences see the same map; N is well past
convergence, so the result is also the fixed-point gradient."""
import sys

import numpy as np
import pytest

from oracle import tnad_oracle as O

pytestmark = pytest.mark.gpu

D, S, CHI, NSTEPS, H = 3, 2, 64, 30, 1e-6
MIXED = dict(precision="fp32", switch_tol=1e-5, eig_warm=True, bwd_precision="fp32")


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def central_differences(ctx):
    import tnad
    h = O.hamiltonian_heisenberg()
    b = O.symmetrized_random_bulk(D, S, 3, "uniform")
    idx = np.random.default_rng(7).choice(b.size, 32, replace=False)
    flat = np.ravel(b, order="F")
    fd = np.zeros(len(idx))
    for q, i in enumerate(idx):
        vals = []
        for sgn in (1.0, -1.0):
            x = flat.copy()
            x[i] += sgn * H
            e, _ = tnad.energy_and_grad(h, np.reshape(x, b.shape, order="F"), CHI, 0.0, NSTEPS, ctx=ctx,
                                        want_grad=False)
            vals.append(e)
        fd[q] = (vals[0] - vals[1]) / (2 * H)
    _say(f"central differences on {len(idx)} entries: |fd| max {np.max(np.abs(fd)):.3e}")
    return h, b, idx, fd


@pytest.mark.parametrize("variant", ["fp64", "fp32_contractions"])
def test_config_d_gradient_against_central_differences(ctx, central_differences, variant):
    import tnad
    h, b, idx, fd = central_differences
    _say(f"case: config D gradient, {variant}")
    opts = {} if variant == "fp64" else MIXED
    e, g = tnad.energy_and_grad(h, b, CHI, 0.0, NSTEPS, ctx=ctx, **opts)
    gsel = np.ravel(g, order="F")[idx]
    err = np.linalg.norm(gsel - fd) / np.linalg.norm(fd)
    _say(f"energy {e:.15f}, gradient vs central differences: relative error {err:.3e}")
    # finite differences of a converged FP64 energy carry ~1e-9 absolute noise (rounding / h);
    # the FP32 variant is held to the north-star bound of 1e-5
    assert err < (1e-6 if variant == "fp64" else 1e-5)
