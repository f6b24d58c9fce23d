"""Writes the committed golden fixtures under tests/golden/ from the CPU oracle (oracle/tnad_oracle.py) on
the workloads' seeded inputs.  Each fixture is an .npz of plain arrays (no pickles): inputs, the oracle's
outputs, and the settings they were computed with.

    python tests/golden/make_fixtures.py [b] [c] [d] [stop]

  b: Heisenberg energy on a uniform [0,1) D=2, d=2 bulk (seed 0), chi=20, tol=1e-6, maxit=100: bulk (32),
     energy, gradient (32), step count and the per-step vals traces (80 values each).
  c: the same Hamiltonian on a uniform [0,1) D=4, d=2 bulk (seed 0), chi=128 (corner matrix 2048 x 2048),
     tol=1e-10, maxit=100: bulk (512), energy, gradient (512), step count.
  d: a standard-normal D=3, d=2 bulk (seed 3), chi=64, tol=0, maxit=30 (31 steps): bulk (162), energy,
     gradient (162) and central differences (h = 1e-6) of the energy on 32 entries drawn with seed 7.

  stop: energies and step counts at tol = 1e-8 of four runs with corner matrices above 192 (D=3,
     chi=24 and 64).

The oracle's energy gradient is the reverse sweep of the unrolled program (what Zygote computes for
src/variationalipeps.jl:73); the central differences follow num_grad's convention (src/autodiff.jl:58-63)
with the step of the workload (h = 1e-6).
"""
import sys
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path.insert(0, str(ROOT))

from oracle import tnad_oracle as O  # noqa: E402


def config_b():
    h = O.hamiltonian_heisenberg()
    b = O.symmetrized_random_bulk(2, 2, 0, "uniform")
    chi, tol, maxit = 20, 1e-6, 100
    e, g, it = O.energy_and_grad(h, b, chi, tol, maxit)
    bs = O.indexperm_symmetrize(b)
    ap, a = O.double_layer(bs)
    C0, T0 = O.init_raw(a, chi)
    res = O.ctmrg(a, C0, T0, tol, maxit)
    assert res.iters == it
    vals = np.array(res.vals_trace)
    np.savez(HERE / "config_b.npz", bulk=b, h=h, chi=chi, tol=tol, maxit=maxit, energy=e, grad=g, iters=it,
             vals_trace=vals)
    print(f"config B: e={e:.16f} iters={it} |g|={np.linalg.norm(g):.6e}")


def config_c():
    h = O.hamiltonian_heisenberg()
    b = O.symmetrized_random_bulk(4, 2, 0, "uniform")
    chi, tol, maxit = 128, 1e-10, 100
    e, g, it = O.energy_and_grad(h, b, chi, tol, maxit)
    np.savez(HERE / "config_c.npz", bulk=b, h=h, chi=chi, tol=tol, maxit=maxit, energy=e, grad=g, iters=it)
    print(f"config C: e={e:.16f} iters={it} |g|={np.linalg.norm(g):.6e}")


def config_d():
    h = O.hamiltonian_heisenberg()
    b = O.symmetrized_random_bulk(3, 2, 3, "normal")
    chi, tol, maxit, step = 64, 0.0, 30, 1e-6
    e, g, it = O.energy_and_grad(h, b, chi, tol, maxit)
    idx = np.random.default_rng(7).choice(b.size, 32, replace=False)
    flat = np.ravel(b, order="F")
    fd = np.zeros(len(idx))
    for q, i in enumerate(idx):
        ev = []
        for sgn in (0.5, -0.5):
            x = flat.copy()
            x[i] += sgn * step
            ev.append(O.energy(h, np.reshape(x, b.shape, order="F"), chi, tol, maxit))
        fd[q] = (ev[0] - ev[1]) / step
    np.savez(HERE / "config_d.npz", bulk=b, h=h, chi=chi, tol=tol, maxit=maxit, energy=e, grad=g, iters=it,
             fd_index=idx, fd_step=step, fd=fd)
    err = np.linalg.norm(np.ravel(g, order="F")[idx] - fd) / np.linalg.norm(fd)
    print(f"config D: e={e:.16f} iters={it} gradient vs central differences {err:.3e}")


# CTMRG runs with corner matrices above 192 (where the GPU's stop test sees the leading block of the
# spectrum): Heisenberg energy and the reference's step count from :raw at tol = 1e-8, maxit = 100
STOP_CASES = [(3, 24, 13, "normal"), (3, 24, 14, "uniform"), (3, 64, 3, "normal"), (3, 64, 21, "uniform")]


def stop_counts():
    h = O.hamiltonian_heisenberg()
    rows = []
    for D, chi, seed, dist in STOP_CASES:
        b = O.symmetrized_random_bulk(D, 2, seed, dist)
        e, res = O.energy(h, b, chi, 1e-8, 100, return_state=True)
        rows.append((D, chi, seed, 0 if dist == "normal" else 1, res.iters, e))
        print(f"  D={D} chi={chi} seed={seed} {dist}: iters={res.iters} e={e:.15f}")
    arr = np.array(rows, dtype=float)
    np.savez(HERE / "stop_counts.npz", cases=arr, tol=1e-8, maxit=100,
             columns=np.array(["D", "chi", "seed", "uniform", "iters", "energy"]))


if __name__ == "__main__":
    which = sys.argv[1:] or ["b", "c", "d", "stop"]
    for w in which:
        t0 = time.time()
        {"b": config_b, "c": config_c, "d": config_d, "stop": stop_counts}[w]()
        print(f"  ({time.time() - t0:.1f} s)")
