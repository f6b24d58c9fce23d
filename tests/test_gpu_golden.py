"""GPU parity against the committed golden fixtures (tests/golden/*.npz, written by
tests/golden/make_fixtures.py from the CPU oracle on the workloads' seeded inputs): the CTMRG energy, its
gradient and the step count at the workload shapes, including the chi=128, D=4 headline shape (corner
matrix 2048 x 2048) on the plain FP64 and the mixed-precision paths."""
import sys
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden"


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _load(name):
    return np.load(GOLDEN / name, allow_pickle=False)


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_config_b_energy_gradient_and_vals_trace(ctx):
    # Heisenberg, D=2, d=2, uniform bulk (seed 0), chi=20, tol=1e-6, maxit=100 (corner matrix 80 x 80:
    # full spectrum on the GPU, so vals and the stop test are the reference's exactly)
    import tnad
    f = _load("config_b.npz")
    _say("case: config B fixture")
    e, g, info = tnad.energy_and_grad(f["h"], f["bulk"], int(f["chi"]), float(f["tol"]), int(f["maxit"]),
                                      ctx=ctx, return_info=True)
    _say(f"config B: |dE| {abs(e - float(f['energy'])):.2e}, gradient rel {_rel(g, f['grad']):.2e}, "
         f"iters {info['iters']} (oracle {int(f['iters'])})")
    assert info["iters"] == int(f["iters"])
    assert abs(e - float(f["energy"])) < 1e-12
    assert _rel(g, f["grad"]) < 1e-9
    # per-step vals of the same run, step by step from the :raw start
    bs = tnad.indexperm_symmetrize(tnad.SquareIPEPS(f["bulk"]), ctx=ctx).bulk
    _, a = tnad.double_layer(tnad.SquareIPEPS(bs), ctx=ctx)
    rt = tnad.SquareCTMRGRuntime.new(a, "raw", int(f["chi"]), ctx=ctx)
    for step, vref in enumerate(f["vals_trace"]):
        rt, v = tnad.ctmrgstep(rt, ctx=ctx)
        assert np.max(np.abs(v - vref)) < 1e-12, step


@pytest.mark.parametrize("path", ["fp64", "mixed_fp32"])
def test_config_c_energy_gradient_and_steps(ctx, path):
    # the headline shape: D=4, d=2, uniform bulk (seed 0), chi=128, tol=1e-10, maxit=100.  The
    # mixed-precision budget: |dE| < 1e-7 and gradient relative error < 1e-5 against the FP64 oracle; the
    # plain path is held tighter.  Same step count as the oracle.
    import tnad
    f = _load("config_c.npz")
    _say(f"case: config C fixture, {path}")
    opts = {} if path == "fp64" else dict(tnad.MIXED_FP32, bwd_precision="fp32")
    e, g, info = tnad.energy_and_grad(f["h"], f["bulk"], int(f["chi"]), float(f["tol"]), int(f["maxit"]),
                                      ctx=ctx, return_info=True, **opts)
    de, rel = abs(e - float(f["energy"])), _rel(g, f["grad"])
    _say(f"config C {path}: |dE| {de:.2e}, gradient rel {rel:.2e}, iters {info['iters']} "
         f"(oracle {int(f['iters'])}), low-precision steps {info['low_prec_iters']}")
    assert info["iters"] == int(f["iters"])
    if path == "fp64":
        assert de < 1e-12 and rel < 1e-9
    else:
        assert info["low_prec_iters"] > 0
        assert de < 1e-7 and rel < 1e-5


@pytest.mark.parametrize("variant", ["fp64", "fp32_contractions"])
def test_config_d_gradient_against_oracle(ctx, variant):
    # D=3, d=2, standard-normal bulk (seed 3), chi=64, tol=0, maxit=30 (31 steps): the GPU gradient
    # against the oracle's reverse sweep (what Zygote computes) on all 162 entries, and against the
    # oracle's central differences (h = 1e-6) on 32 entries
    import tnad
    f = _load("config_d.npz")
    _say(f"case: config D fixture, {variant}")
    opts = {} if variant == "fp64" else dict(precision="fp32", switch_tol=1e-5, eig_warm=True,
                                              bwd_precision="fp32")
    e, g, info = tnad.energy_and_grad(f["h"], f["bulk"], int(f["chi"]), float(f["tol"]), int(f["maxit"]),
                                      ctx=ctx, return_info=True, **opts)
    idx = f["fd_index"]
    gsel = np.ravel(g, order="F")[idx]
    rel_o, rel_fd = _rel(g, f["grad"]), _rel(gsel, f["fd"])
    _say(f"config D {variant}: |dE| {abs(e - float(f['energy'])):.2e}, gradient vs oracle {rel_o:.2e}, "
         f"vs oracle central differences {rel_fd:.2e}, iters {info['iters']}")
    assert info["iters"] == int(f["iters"]) == 31
    # the oracle's own gradient meets its central differences at 2.8e-6 (the differences' rounding floor)
    assert rel_fd < 1e-5
    if variant == "fp64":
        assert abs(e - float(f["energy"])) < 1e-11 and rel_o < 1e-8
    else:
        assert abs(e - float(f["energy"])) < 1e-7 and rel_o < 1e-5


def _stop_cases():
    f = _load("stop_counts.npz")
    return [tuple(r) for r in f["cases"]]


@pytest.mark.parametrize("case", _stop_cases() if (GOLDEN / "stop_counts.npz").exists() else [])
def test_step_count_matches_reference_above_192(ctx, case):
    # corner matrices of 216 and 576: the GPU's stop test sees the leading block of the spectrum, the
    # reference's the whole of it; the step counts to tol = 1e-8 and the energies must agree
    import tnad
    from oracle import tnad_oracle as O
    D, chi, seed, uniform, iters, e_ref = case
    D, chi, seed, iters = int(D), int(chi), int(seed), int(iters)
    _say(f"case: step count D={D} chi={chi} seed={seed} uniform={int(uniform)}")
    b = O.symmetrized_random_bulk(D, 2, seed, "uniform" if uniform else "normal")
    e, _, info = tnad.energy_and_grad(O.hamiltonian_heisenberg(), b, chi, 1e-8, 100, ctx=ctx, want_grad=False,
                                      return_info=True)
    _say(f"  iters {info['iters']} (oracle {iters}), |dE| {abs(e - e_ref):.2e}")
    assert info["iters"] == iters
    assert abs(e - e_ref) < 1e-9
