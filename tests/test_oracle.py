"""Pins of the CPU oracle (oracle/tnad_oracle.py) against the reference's own
golden values, closed forms, finite differences and invariants.  These run on
the CPU only; they establish that the oracle the GPU is later compared against
restates the reference correctly."""
import math

import numpy as np
import pytest

from oracle import tnad_oracle as O


# --- TRG: golden scalars held by the reference -----------------------------------------
def test_trg_tensorgrad_golden():
    # test/trg.jl:12-18 (value agreed with the PyTorch tensorgrad project)
    assert O.trg(O.model_tensor_ising(0.4), 5, 5) == pytest.approx(0.8919788686747141, rel=1e-13, abs=0)


def test_trg_readme_golden_value_and_gradient():
    # README.md:58-60 and README.md:67-70 (chi=20, niter=20, beta=0.5)
    lnz, g = O.trg_and_grad(0.5, 20, 20)
    assert lnz == pytest.approx(1.0257933734351765, rel=1e-13, abs=0)
    assert g == pytest.approx(1.7455677143228514, rel=1e-12, abs=0)


def test_trg_userguide_gradient():
    # docs/src/userguide.md:25-28
    _, g = O.trg_and_grad(0.5, 5, 5)
    assert g == pytest.approx(1.7502426939979507, rel=1e-13, abs=0)


def test_trg_gradient_matches_num_grad():
    # test/trg.jl:19
    f = lambda b: O.trg(O.model_tensor_ising(b), 5, 5)
    _, g = O.trg_and_grad(0.4, 5, 5)
    assert O.num_grad(f, 0.4, 1e-6) == pytest.approx(g, rel=1e-7)


@pytest.mark.parametrize("chi,niter", [(6, 3), (8, 4)])
def test_trg_generic_tensor_gradient_matches_num_grad(chi, niter):
    # full-rank tensor, truncated at chi: the reverse sweep along a random direction against
    # central differences of lnZ (the same configuration the GPU parity test uses)
    a = np.random.default_rng(7).uniform(0.5, 1.5, (3, 3, 3, 3))
    direc = np.random.default_rng(chi).standard_normal(a.shape)
    _, g = O.trg(a, chi, niter, with_grad_seed=direc)
    f = lambda t: O.trg(a + t * direc, chi, niter)
    assert O.num_grad(f, 0.0, 1e-6) == pytest.approx(g, rel=1e-6)


def test_trg_onsager_closed_form():
    # independent pin: Onsager's exact free energy (chi=20 TRG is within 1e-6 of it)
    assert O.onsager_lnz(0.5) == pytest.approx(1.0257928126949176, abs=5e-14)
    assert O.onsager_lnz(0.4) == pytest.approx(0.87936382077494819, abs=5e-14)
    lnz, g = O.trg_and_grad(0.5, 20, 20)
    assert abs(lnz - O.onsager_lnz(0.5)) < 1e-6
    dexact = O.num_grad(O.onsager_lnz, 0.5, 1e-4)
    assert abs(g - dexact) < 1e-5


def test_trg_svd_reconstruction():
    # test/trg.jl:6-10: at tol=0 the split reproduces the tensor
    rng = np.random.default_rng(0)
    t = rng.standard_normal((10, 10, 10, 10))
    u, v = O.trg_svd(t, 100, 0)
    assert np.allclose(np.einsum("ija,akl->ijkl", u, v), t, atol=1e-10)


def test_trg_svd_keep_count_off_by_one():
    # src/trg.jl:37: k = min(#{s > tol} + 1, dmax, len(s))
    rng = np.random.default_rng(1)
    a = rng.standard_normal((4, 2))
    b = rng.standard_normal((2, 4))
    t = np.reshape(a @ b, (2, 2, 2, 2), order="F")      # rank 2 of 4
    u, v = O.trg_svd(t, 10, 1e-8)
    assert u.shape[2] == 3
    u, v = O.trg_svd(t, 2, 1e-8)
    assert u.shape[2] == 2


def test_tensorfromclassical_matches_model_tensor():
    # test/exampletensors.jl:6-9
    b = 0.37
    assert np.allclose(O.model_tensor_ising(b), O.tensorfromclassical(np.array([[b, -b], [-b, b]])))


def test_dmodel_tensor_is_derivative():
    b, h = 0.45, 1e-6
    fd = (O.model_tensor_ising(b + h / 2) - O.model_tensor_ising(b - h / 2)) / h
    assert np.allclose(O.dmodel_tensor_ising(b), fd, atol=1e-8)


# --- svd_back: central FD on small random matrices (test/svd.jl pattern) ---------------
@pytest.mark.parametrize("shape", [(5, 5), (6, 6)])
def test_svd_back_matches_fd(shape):
    rng = np.random.default_rng(3)
    A = rng.standard_normal(shape)
    H = rng.standard_normal((shape[0], shape[0]))
    H = H + H.T
    k = 2

    def loss(M):
        U, s, Vh = np.linalg.svd(M)
        # gauge-invariant use of the first k left vectors and all singular values
        P = U[:, :k] @ U[:, :k].T
        return np.sum(P * H) + np.sum(s[:k] ** 2)

    U, s, Vh = np.linalg.svd(A)
    dU = np.zeros_like(U)
    dU[:, :k] = 2 * H @ U[:, :k]          # d/dU of tr(U_k U_k^T H)
    dS = np.zeros_like(s)
    dS[:k] = 2 * s[:k]
    g = O.svd_back(U, s, Vh.T, dU, dS, None)
    gn = O.num_grad(loss, A, 1e-6)
    assert np.allclose(g, gn, atol=1e-6)


# trg_svd's pullback on non-square splits (the projector terms of src/trg.jl:98-104; test/svd.jl checks
# (6,3) and (3,6) matrices): central FD of a loss that is invariant under the singular vectors' signs
def trg_svd_quadratic_loss(shape, dmax, seed=5):
    rng = np.random.default_rng(seed)
    t = rng.standard_normal(shape)
    d1, d2, d3, d4 = shape
    m, mp = d1 * d2, d3 * d4
    A = rng.standard_normal((m, m))
    A = A + A.T
    B = rng.standard_normal((mp, mp))
    B = B + B.T

    def loss(x):
        u, v = O.trg_svd(x, dmax, 0.0)
        k = u.shape[2]
        um, vm = O.reshape_f(u, (m, k)), O.reshape_f(v, (k, mp))
        return np.trace(um.T @ A @ um) + np.trace(vm @ B @ vm.T)

    u, v = O.trg_svd(t, dmax, 0.0)
    k = u.shape[2]
    um, vm = O.reshape_f(u, (m, k)), O.reshape_f(v, (k, mp))
    ubar = O.reshape_f(2 * A @ um, (d1, d2, k))
    vbar = O.reshape_f(2 * vm @ B, (k, d3, d4))
    return t, loss, ubar, vbar


TRG_SVD_BACK_CASES = [((2, 3, 3, 1), 10), ((3, 1, 2, 3), 10), ((2, 3, 3, 1), 2), ((3, 1, 2, 3), 2),
                      ((2, 3, 2, 3), 4), ((4, 5, 3, 3), 6)]


@pytest.mark.parametrize("shape,dmax", TRG_SVD_BACK_CASES)
def test_trg_svd_back_rectangular_matches_fd(shape, dmax):
    t, loss, ubar, vbar = trg_svd_quadratic_loss(shape, dmax)
    _, _, cache = O.trg_svd(t, dmax, 0.0, keep=True)
    g = O.trg_svd_back(cache, ubar, vbar)
    gn = O.num_grad(loss, t, 1e-6)
    # without the projector terms the error is O(|g|) (checked by planting the omission)
    assert np.max(np.abs(g - gn)) <= 1e-6 * np.max(np.abs(g))


def test_svd_back_all_nothing_returns_none():
    # src/trg.jl:73
    U = np.eye(3)
    assert O.svd_back(U, np.ones(3), U, None, None, None) is None


# --- models -----------------------------------------------------------------------
def test_heisenberg_matches_readme_print():
    # README.md:83-100 prints h[:, :, k, l] blocks of hamiltonian(Heisenberg())
    h = O.hamiltonian_heisenberg()
    assert np.allclose(h[:, :, 0, 0], [[-0.5, 0.0], [0.0, 0.5]])
    assert np.allclose(h[:, :, 1, 0], [[0.0, 0.0], [-1.0, 0.0]])
    assert np.allclose(h[:, :, 0, 1], [[0.0, -1.0], [0.0, 0.0]])
    assert np.allclose(h[:, :, 1, 1], [[0.5, 0.0], [0.0, -0.5]])


@pytest.mark.parametrize("hx", [0.0, 0.5, 1.0, 2.0])
def test_tfising_two_site_spectrum(hx):
    # hamiltonian(TFIsing(hx)) (src/hamiltonianmodels.jl:28-33) is the two-site operator
    # -2 Z Z - hx/2 (X 1 + 1 X) with site one on indices (i, j) and site two on (k, l).  Its spectrum in
    # closed form: the singlet (+2), (|uu> - |dd>)/sqrt2 (-2) and the pair +-sqrt(4 + hx^2) from
    # {(|uu> + |dd>)/sqrt2, triplet 0} coupled by -hx.  A dropped /2, a wrong sign or a transposed site
    # changes it.
    h = O.hamiltonian_tfising(hx)
    H = np.transpose(h, (0, 2, 1, 3)).reshape(4, 4)     # rows (i, k), columns (j, l)
    assert np.allclose(H, H.T)
    ev = np.sort(np.linalg.eigvalsh(H))
    r = math.sqrt(4 + hx * hx)
    assert np.allclose(ev, np.sort([2.0, -2.0, r, -r]), atol=1e-14)
    # the ferromagnetic product state |uu> has <h> = -2 at any hx (the field has no diagonal part)
    assert H[0, 0] == pytest.approx(-2.0)


def test_init_raw_matches_reference_doctest():
    # src/ctmrg.jl:52-60: SquareCTMRGRuntime(randn(2,2,2,2), Val(:raw), 4) has
    # corner[1:2,1:2] = sum(bulk, dims=(3,4)) and edge[1:2,1:2,1:2] = sum(bulk, dims=4); the rest of the
    # chi x chi / chi x D x chi arrays are the zeros they were created as (:77-78).  A non-symmetric
    # bulk, so an index-order mistake shows.
    rng = np.random.default_rng(52)
    a = rng.standard_normal((2, 2, 2, 2))
    C, T = O.init_raw(a, 4)
    assert C.shape == (4, 4) and T.shape == (4, 2, 4)
    Cref = np.zeros((4, 4))
    Tref = np.zeros((4, 2, 4))
    for i in range(2):
        for j in range(2):
            Cref[i, j] = sum(a[i, j, k, l] for k in range(2) for l in range(2))
            for k in range(2):
                Tref[i, j, k] = sum(a[i, j, k, l] for l in range(2))
    assert np.allclose(C, Cref, atol=1e-15) and np.allclose(T, Tref, atol=1e-15)
    assert not np.allclose(C[:2, :2], C[:2, :2].T)   # the input really is non-symmetric


def test_init_raw_truncates_when_D_exceeds_chi():
    # src/ctmrg.jl:79-84: entries of cinit / tinit outside CartesianIndices(corner / edge) are dropped
    rng = np.random.default_rng(53)
    a = rng.standard_normal((5, 5, 5, 5))
    C, T = O.init_raw(a, 3)
    assert C.shape == (3, 3) and T.shape == (3, 5, 3)
    for i in range(3):
        for j in range(3):
            assert C[i, j] == pytest.approx(a[i, j].sum(), abs=1e-13)
        for j in range(5):
            for k in range(3):
                assert T[i, j, k] == pytest.approx(a[i, j, k].sum(), abs=1e-13)


def test_symmetrize_is_d4_invariant_and_normalised():
    rng = np.random.default_rng(4)
    x = O.indexperm_symmetrize(rng.random((2, 2, 2, 2, 2)))
    assert np.linalg.norm(x) == pytest.approx(1.0)
    for p in [(0, 3, 2, 1, 4), (2, 1, 0, 3, 4), (1, 0, 3, 2, 4), (3, 2, 1, 0, 4), (1, 2, 3, 0, 4)]:
        assert np.allclose(np.transpose(x, p), x)


def test_symmetrize_dimension_mismatch():
    # test/ctmrg.jl:13
    with pytest.raises(ValueError):
        O.indexperm_symmetrize(np.zeros((3, 3, 4, 3, 2)))


# --- CTMRG / energy -------------------------------------------------------------------
def _rand_bulk(seed, d=2, s=2, dist="normal"):
    return O.symmetrized_random_bulk(d, s, seed, dist)


def test_ctmrgstep_invariants():
    b = O.indexperm_symmetrize(_rand_bulk(0))
    ap, a = O.double_layer(b)
    C, T = O.init_raw(a, 6)
    for _ in range(4):
        C, T, vals = O.ctmrgstep(a, C, T)
        assert np.allclose(C, C.T)
        assert np.allclose(T, np.transpose(T, (2, 1, 0)))
        assert np.linalg.norm(C) == pytest.approx(1.0)
        assert np.linalg.norm(T) == pytest.approx(1.0)
        assert vals[0] == 1.0 and np.all(np.diff(vals) <= 1e-15)


def test_stopfunction_semantics():
    # src/ctmrg.jl:114 builds StopFunction(oldvals, -1, tol, maxit) and src/fixedpoint.jl:12-14 applies it
    # to every state of iterated(f, guess), the initial one included: with the counter at -1, the state
    # after step m sees counter = m, so tol = 0 runs maxit + 1 steps and maxit = 0 runs one; the first two
    # states never stop (NaN, then Inf)
    b = O.indexperm_symmetrize(_rand_bulk(1))
    ap, a = O.double_layer(b)
    C, T = O.init_raw(a, 4)
    assert O.ctmrg(a, C, T, tol=1e9, maxit=100).iters == 2
    assert O.ctmrg(a, C, T, tol=1e9, maxit=1).iters == 2
    assert O.ctmrg(a, C, T, tol=0.0, maxit=7).iters == 8
    assert O.ctmrg(a, C, T, tol=0.0, maxit=0).iters == 1
    assert O.ctmrg(a, C, T, tol=0.0, maxit=3).iters == 4


def test_product_state_energies():
    # test/variationalipeps.jl:14-25
    h = O.diaglocalhamiltonian([1.0, -1.0])
    rng = np.random.default_rng(5)
    a = np.zeros((2, 2, 2, 2, 2)) + 1e-12 * rng.standard_normal((2, 2, 2, 2, 2))
    a[0, 0, 0, 0, 1] = rng.standard_normal()
    assert O.energy(h, a, 4, 1e-12, 100) / 2 == pytest.approx(-1.0)
    a = np.zeros((2, 2, 2, 2, 2)) + 1e-12 * rng.standard_normal((2, 2, 2, 2, 2))
    a[0, 0, 0, 0, 0] = rng.standard_normal()
    assert O.energy(h, a, 10, 0.0, 300) / 2 == pytest.approx(1.0)
    a = np.zeros((2, 2, 2, 2, 2)) + 1e-12 * rng.standard_normal((2, 2, 2, 2, 2))
    a[0, 0, 0, 0, 1] = a[0, 0, 0, 0, 0] = rng.standard_normal()
    assert abs(O.energy(h, a, 10, 0.0, 300)) < 1e-9


def test_energy_bounded_noninteracting():
    # test/variationalipeps.jl:10-12
    h = O.diaglocalhamiltonian([1.0, -1.0])
    rng = np.random.default_rng(6)
    for _ in range(5):
        e = O.energy(h, rng.random((3, 3, 3, 3, 2)), 5, 0.0, 10) / 2
        assert -1 < e < 1


def test_energy_gauge_and_scale_invariance():
    h = O.hamiltonian_heisenberg()
    b = _rand_bulk(7)
    e0 = O.energy(h, b, 6, 1e-12, 100)
    assert O.energy(h, 3.7 * b, 6, 1e-12, 100) == pytest.approx(e0, abs=1e-12)
    assert O.energy(h, -b, 6, 1e-12, 100) == pytest.approx(e0, abs=1e-12)


@pytest.mark.parametrize("seed,chi,maxit", [(8, 4, 0), (8, 4, 1), (8, 4, 5), (9, 6, 3), (10, 8, 6)])
def test_energy_gradient_matches_num_grad_frozen_init(seed, chi, maxit):
    # the reverse sweep is exact for the function Zygote differentiates: the :raw
    # initial environment is @nograd (src/autodiff.jl:5), so FD holds it fixed.
    # chi=6, 8 exceed the rank of the first corner matrix (null-space columns kept).
    h = O.hamiltonian_heisenberg()
    b = _rand_bulk(seed)
    env0 = O.raw_env_of(b, chi)
    e, g, it = O.energy_and_grad(h, b, chi, 0.0, maxit)
    gn = O.num_grad(lambda x: O.energy_frozen_init(h, x, chi, 0.0, maxit, env0), b, 1e-5)
    assert np.max(np.abs(g - gn)) < 2e-8 * max(1.0, np.max(np.abs(g)))


def test_energy_gradient_approaches_num_grad_at_convergence():
    # test/variationalipeps.jl:121-134: with many iterations the init's influence is gone
    # and the gradient matches FD of the full energy (atol 1e-3 in the reference).
    h = O.hamiltonian_heisenberg()
    b = _rand_bulk(8)
    e, g, it = O.energy_and_grad(h, b, 4, 0.0, 60)
    gn = O.num_grad(lambda x: O.energy(h, x, 4, 0.0, 60), b, 1e-5)
    assert np.max(np.abs(g - gn)) < 1e-8


def test_ctmrg_yang_magnetisation():
    # test/ctmrg.jl:35-42 (random init from NumPy's generator)
    rng = np.random.default_rng(5)
    assert O.magnetisation(0.0, 2, rng) == pytest.approx(O.magofbeta(0.0), abs=1e-6)
    assert O.magnetisation(1.0, 2, rng) == pytest.approx(O.magofbeta(1.0), rel=1e-6)
    assert O.magnetisation(0.2, 10, rng) == pytest.approx(O.magofbeta(0.2), abs=1e-4)
    assert O.magnetisation(0.6, 4, rng) == pytest.approx(O.magofbeta(0.6), rel=1e-6)
    assert O.magnetisation(0.8, 2, rng) == pytest.approx(O.magofbeta(0.8), rel=1e-6)


def test_num_grad_convention():
    # src/autodiff.jl:39-42 doctest
    assert O.num_grad(lambda x: x * x, 3.0) == pytest.approx(6.0)
    assert np.allclose(O.num_grad(np.trace, np.random.default_rng(0).random((2, 2))), np.eye(2))


# --- complex arithmetic (src/trg.jl:85 diagonal term; test/svd.jl, test/trg.jl:22-28,
# test/variationalipeps.jl:136-155) -----------------------------------------------------------------
def _cfd(f, x, idx, h=1e-6):
    """central differences of the real function f along Re and Im of x[idx]: df/dRe + i df/dIm"""
    out = []
    for i in idx:
        parts = []
        for dz in (h, 1j * h):
            xp = x.copy()
            xm = x.copy()
            xp[i] += dz / 2
            xm[i] -= dz / 2
            parts.append((f(xp) - f(xm)) / h)
        out.append(parts[0] + 1j * parts[1])
    return np.array(out)


def _svd_losses(M, H1, H2):
    """the four losses of test/svd.jl (U[:,1], V[:,1], U[1,1] with V[1,1], sum S) and their cotangents"""
    U, s, Vh = np.linalg.svd(M, full_matrices=False)
    V = Vh.conj().T
    out = {}
    psi = U[:, 0]
    dU = np.zeros_like(U)
    dU[:, 0] = 2 * H1 @ psi
    out["U"] = (np.real(psi.conj() @ H1 @ psi), (dU, None, None))
    psi = V[:, 0]
    dV = np.zeros_like(V)
    dV[:, 0] = 2 * H2 @ psi
    out["V"] = (np.real(psi.conj() @ H2 @ psi), (None, None, dV))
    # real(conj(U[1,1]) V[1,1])
    dU = np.zeros_like(U)
    dV = np.zeros_like(V)
    dU[0, 0] = V[0, 0]
    dV[0, 0] = U[0, 0]
    out["UV"] = (np.real(np.conj(U[0, 0]) * V[0, 0]), (dU, None, dV))
    out["S"] = (np.sum(s), (None, np.ones_like(s), None))
    return out, (U, s, V)


@pytest.mark.parametrize("shape", [(6, 3), (3, 6), (3, 3)])
@pytest.mark.parametrize("which", ["U", "V", "UV", "S"])
def test_complex_svd_back_matches_fd(shape, which):
    # test/svd.jl:21-98: every loss is invariant under the singular vectors' phases, so its central
    # differences through LAPACK's SVD pin svd_back, including the i Im diag(U'dU) Sinv term (src/trg.jl:85)
    rng = np.random.default_rng(2)
    m, p = shape
    M = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    H1 = rng.standard_normal((m, m)) + 1j * rng.standard_normal((m, m))
    H1 = H1 + H1.conj().T
    H2 = rng.standard_normal((p, p)) + 1j * rng.standard_normal((p, p))
    H2 = H2 + H2.conj().T
    losses, (U, s, V) = _svd_losses(M, H1, H2)
    g = O.svd_back(U, s, V, *losses[which][1])
    f = lambda X: _svd_losses(X, H1, H2)[0][which][0]
    idx = list(np.ndindex(shape))
    fd = _cfd(f, M, idx)
    gsel = np.array([g[i] for i in idx])
    assert np.max(np.abs(gsel - fd)) < 1e-6 * max(1.0, np.max(np.abs(fd)))


def test_complex_svd_back_imaginary_diagonal_case():
    # test/svd.jl:69-86: A = [-1+1im 2+1im; 1-2im 3+0.8im], loss real(conj(U[1,1]) V[1,1]), entry (2,1)
    A = np.array([[-1 + 1j, 2 + 1j], [1 - 2j, 3 + 0.8j]])
    losses, (U, s, V) = _svd_losses(A, np.eye(2), np.eye(2))
    g = O.svd_back(U, s, V, *losses["UV"][1])
    f = lambda X: _svd_losses(X, np.eye(2), np.eye(2))[0]["UV"][0]
    fd = _cfd(f, A, [(1, 0)], h=1e-4)[0]
    assert abs(g[1, 0] - fd) < 1e-6
    # without the diagonal term the pullback misses (the term the reference adds for complex input)
    U_, s_, V_ = U, s, V
    UdU = U_.conj().T @ losses["UV"][1][0]
    assert np.max(np.abs(np.imag(np.diag(UdU)))) > 1e-3


def test_complex_ctmrgstep_adjoint_matches_fd_on_closed_network():
    # complex CTMRG steps are covariant under the environment's phase gauge, so a closed network of the
    # outputs (every bond pairing a row with a column index) is gauge invariant and its central differences
    # pin the complex reverse sweep (svd_back of the Hermitian corner matrix and the einsum adjoints)
    rng = np.random.default_rng(4)
    cr = lambda *s: rng.standard_normal(s) + 1j * rng.standard_normal(s)
    chi, D = 4, 4
    C = cr(chi, chi)
    C = C + C.conj().T
    T = cr(chi, D, chi)
    T = T + O.ein("ijk->kji", T.conj())
    a = cr(D, D, D, D)

    def run(a_, C_, T_, keep=False):
        caches = []
        for _ in range(2):
            out = O.ctmrgstep(a_, C_, T_, keep=keep)
            C_, T_ = out[0], out[1]
            if keep:
                caches.append(out[3])
        return C_, T_, caches

    def L(a_, C_, T_):
        C2, T2, _ = run(a_, C_, T_)
        return np.real(O.ein("ij,ji->", C2, C2) + O.ein("ij,jak,kl,lai->", C2, T2, C2, T2))

    C2, T2, caches = run(a, C, T, keep=True)
    Cb = 2 * C2.T.conj() + (O.ein("jak,kl,lai->ij", T2, C2, T2) + O.ein("ij,jak,lai->kl", C2, T2, T2)).conj()
    Tb = (O.ein("ij,kl,lai->jak", C2, C2, T2) + O.ein("ij,jak,kl->lai", C2, T2, C2)).conj()
    ab = 0
    for cache in reversed(caches):
        Cb, Tb, x = O.ctmrgstep_back(a, cache, Cb, Tb)
        ab = ab + x
    for x, xb, f in ((C, Cb, lambda v: L(a, v, T)), (T, Tb, lambda v: L(a, C, v)), (a, ab, lambda v: L(v, C, T))):
        idx = [tuple(int(rng.integers(0, n)) for n in x.shape) for _ in range(4)]
        fd = _cfd(f, x, idx)
        gsel = np.array([xb[i] for i in idx])
        assert np.max(np.abs(gsel - fd)) < 2e-6 * np.max(np.abs(xb))


def test_complex_energy_real_copy_and_global_phase():
    # test/variationalipeps.jl:136-155: a real bulk and its complex copy give the same energy and
    # gradient; a global phase leaves the double layer, hence the energy, unchanged
    h = O.hamiltonian_heisenberg()
    b = O.symmetrized_random_bulk(2, 2, 2, "normal")
    e1, g1, it1 = O.energy_and_grad(h, b, 4, 1e-12, 100)
    e2, g2, it2 = O.energy_and_grad(h, b + 0j, 4, 1e-12, 100)
    ph = np.exp(1j * 2.1)
    e3, g3, it3 = O.energy_and_grad(h, b * ph, 4, 1e-12, 100)
    assert it1 == it2 == it3
    assert abs(e1 - e2) < 1e-13 and abs(e1 - e3) < 1e-13
    assert np.max(np.abs(g1 - g2)) < 1e-12
    # d/d(phi x) = phi d/dx (the energy is phase invariant)
    assert np.max(np.abs(g3 - ph * g1)) < 1e-12


def test_complex_phase_rule():
    # real columns stay real (a sign), complex ones end with sum_r w_r U[r, j] real positive
    rng = np.random.default_rng(1)
    X = rng.standard_normal((7, 3))
    Y, ph = O.fix_phases(X + 0j)
    assert np.all(np.abs(np.imag(Y)) == 0) and np.all(np.abs(np.abs(ph) - 1) < 1e-15)
    Z = rng.standard_normal((7, 3)) + 1j * rng.standard_normal((7, 3))
    Zf, _ = O.fix_phases(Z * np.exp(1j * 0.4))
    q = O.phase_weights(7) @ Zf
    assert np.all(np.abs(np.imag(q)) < 1e-14) and np.all(np.real(q) > 0)
    Zg, _ = O.fix_phases(Z)
    assert np.max(np.abs(Zf - Zg)) < 1e-14   # the rule removes any prior phase


def test_complex_trg_matches_real_and_fd():
    # test/trg.jl:22-28: the complex copy of the Ising tensor gives the real lnZ and d lnZ / d beta;
    # a generic complex tensor's directional derivative of Re lnZ matches central differences
    a, da = O.model_tensor_ising(0.4), O.dmodel_tensor_ising(0.4)
    l1, d1 = O.trg(a, 12, 3, with_grad_seed=da)
    l2, d2 = O.trg(a + 0j, 12, 3, with_grad_seed=da + 0j)
    assert abs(l1 - l2) < 1e-14 and abs(d1 - d2) < 1e-12
    rng = np.random.default_rng(1)
    t = rng.standard_normal((3, 3, 3, 3)) + 1j * rng.standard_normal((3, 3, 3, 3)) + 3.0
    dirn = rng.standard_normal(t.shape) + 1j * rng.standard_normal(t.shape)
    _, d = O.trg(t, 6, 3, with_grad_seed=dirn)
    h = 1e-6
    fd = (np.real(O.trg(t + h / 2 * dirn, 6, 3)) - np.real(O.trg(t - h / 2 * dirn, 6, 3))) / h
    assert abs(d - fd) < 1e-7 * abs(d)
