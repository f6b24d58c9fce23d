#!/usr/bin/env python
"""CTMRG forward + adjoint on one MI355X: ms per CTMRG iteration at chi=128, D=4
(iPEPS bond 4, double-layer bond 16, corner matrix 2048 x 2048), mixed precision
vs plain FP64, with the energy-accuracy condition |dE| < 1e-7.

A "step" is one CTMRG iteration: its forward renormalisation step and its share
of the reverse sweep.  The timed region is one fused energy + gradient library
call (tnad_energy with a gradient output) running exactly K iterations (tol=0,
maxit=K) on device-resident inputs, so it also contains the per-call fixed
work (symmetrisation, double layer, :raw init, expectation value and their
adjoints); ms_per_step = elapsed / K.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu-baseline]

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
for p in (str(ROOT), str(ROOT / "tensornetworkad.jl_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

METRIC = "CTMRG fwd+adjoint ms/iter at chi=128,D=4: mixed vs FP64 plain, |dE|<1e-7"

# MI355X published peaks
PEAK = {"fp64": 78.6, "fp32": 157.3, "bf16": 2500.0}   # dense matrix TFLOP/s


def heisenberg_h():
    import numpy as np
    sx = np.array([[0.0, 1.0], [1.0, 0.0]])
    sy = np.array([[0.0, -1.0j], [1.0j, 0.0]])
    sz = np.array([[1.0, 0.0], [0.0, -1.0]])
    o = lambda a, b: np.einsum("ij,kl->ijkl", a, b)
    h = o(sz, sz) - o(sx, sx) - o(sy, sy)
    h = np.einsum("ijcd,kc,ld->ijkl", h, sx, sx.conj().T)
    return np.asfortranarray(np.real(h / 2))


def synthetic_bulk(D, s, seed):
    """uniform [0,1) draw of the (D,D,D,D,s) bulk from NumPy's PCG64 (column-major order)"""
    import numpy as np
    rng = np.random.default_rng(seed)
    return np.reshape(rng.random(D ** 4 * s), (D, D, D, D, s), order="F")


def flops_per_iter(chi, Dd):
    """algorithmic work of one CTMRG iteration (forward contractions + 2x for the adjoint
    contractions + the truncation-aware eigen adjoint), flop counts of the minimal pairwise order"""
    n = chi * Dd
    F_c = 4 * chi ** 3 * Dd + 8 * chi ** 3 * Dd ** 2 + 4 * chi ** 2 * Dd ** 4
    return 3 * F_c + 10 * n * n * chi, F_c


def run_energy(tnad, ctx, h_dev, b_dev, D, s, chi, maxit, tol, opts, grad_dev):
    import ctypes as C
    o = tnad.make_opts(**opts)
    e = C.c_double(0.0)
    it = C.c_int(0)
    ctx.check(ctx.lib.tnad_energy(ctx.h, C.c_void_p(h_dev), C.c_void_p(b_dev), D, s, chi, float(tol), int(maxit),
                                  C.byref(o), C.byref(e), C.c_void_p(grad_dev) if grad_dev else None, C.byref(it)))
    return e.value, it.value


def cpu_baseline(chi, D, s, seed):
    """The FP64 NumPy restatement (oracle) on the host: fwd+adjoint seconds per iteration by the
    slope of one- and two-iteration runs at the same workload."""
    from oracle import tnad_oracle as O
    import numpy as np
    h = O.hamiltonian_heisenberg()
    b = O.symmetrized_random_bulk(D, s, seed, "uniform")
    t0 = time.perf_counter()
    O.energy_and_grad(h, b, chi, 0.0, 1)
    t1 = time.perf_counter()
    O.energy_and_grad(h, b, chi, 0.0, 2)
    t2 = time.perf_counter()
    per_iter = (t2 - t1) - (t1 - t0)
    return per_iter * 1e3, (t2 - t0)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--chi", type=int, default=128)
    ap.add_argument("--D", type=int, default=4)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true")
    a = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if a.gpus != 1 or world != 1:
        print("bench.py: this build measures one GPU only (the CTMRG fixed point is sequential and the "
              "multi-GPU path is not built); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    import tnad

    chi, D, s = a.chi, a.D, 2
    Dd = D * D
    ctx = tnad.Context(0)
    dev = torch.device("cuda:0")
    h = heisenberg_h()
    bulk = synthetic_bulk(D, s, a.seed)
    # inputs resident in HBM: column-major buffers behind torch-owned device memory
    h_t = torch.from_numpy(np.ascontiguousarray(h.ravel(order="F"))).to(dev)
    b_t = torch.from_numpy(np.ascontiguousarray(bulk.ravel(order="F"))).to(dev)
    g_t = torch.zeros_like(b_t)
    torch.cuda.synchronize()

    plain = dict()
    mixed = dict(precision="fp32", switch_tol=1e-5, eig_warm=True, eig_warm_iters=2, bwd_precision="fp32")
    results = {}
    for name, opts in (("plain_fp64", plain), ("mixed", mixed)):
        run_energy(tnad, ctx, h_t.data_ptr(), b_t.data_ptr(), D, s, chi, a.warmup, 0.0, opts, g_t.data_ptr())
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        e, it = run_energy(tnad, ctx, h_t.data_ptr(), b_t.data_ptr(), D, s, chi, a.steps, 0.0, opts, g_t.data_ptr())
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        assert it == a.steps
        results[name] = dict(ms=dt * 1e3 / a.steps, e=e, stats=ctx.stats())

    acc = {}
    if not a.no_accuracy:
        # converged energies and gradients of both paths (tol 1e-10, maxit 100)
        g64 = torch.zeros_like(b_t)
        g_mx = torch.zeros_like(b_t)
        e64, it64 = run_energy(tnad, ctx, h_t.data_ptr(), b_t.data_ptr(), D, s, chi, 100, 1e-10, plain, g64.data_ptr())
        emx, itmx = run_energy(tnad, ctx, h_t.data_ptr(), b_t.data_ptr(), D, s, chi, 100, 1e-10, mixed,
                               g_mx.data_ptr())
        gr = float(torch.linalg.norm(g_mx - g64) / torch.linalg.norm(g64))
        acc = dict(e_fp64=e64, e_mixed=emx, dE=abs(emx - e64), grad_rel_err=gr, iters_fp64=it64, iters_mixed=itmx)

    # roofline: dominant contraction of the step, timed alone with HIP events on the library stream
    spec = "ibcl,jkcb->ijlk"          # enlarged corner, cp = X2 (x) a   (2 chi^2 Dd^4 flops)
    ext = {"i": chi, "l": chi, "b": Dd, "c": Dd, "j": Dd, "k": Dd}
    kflops = 2.0 * chi * chi * Dd ** 4
    roof = {}
    for prec in ("fp64", "fp32"):
        ms = tnad.time_contract(spec, ext, precision=prec, reps=20, ctx=ctx)
        tf = kflops / (ms * 1e-3) / 1e12
        roof[prec] = dict(ms=ms, tflops=tf)
    roofline = {"bound": "mfma", "kernel": f"tc_kernel<float> ein\"{spec}\" (fp32, {kflops:.3e} flop/launch)",
                "achieved": round(roof["fp32"]["tflops"], 3), "peak": PEAK["fp32"], "unit": "TFLOP/s",
                "frac": round(roof["fp32"]["tflops"] / PEAK["fp32"], 4), "traffic": None,
                "fp64_variant": {"achieved": round(roof["fp64"]["tflops"], 3), "peak": PEAK["fp64"],
                                 "frac": round(roof["fp64"]["tflops"] / PEAK["fp64"], 4)}}

    cpu = None
    if not a.no_cpu_baseline:
        per_iter_ms, total_s = cpu_baseline(chi, D, s, a.seed)
        cpu = {"value": round(per_iter_ms, 1), "unit": "ms/iter", "cores": os.cpu_count(), "kind": "port",
               "sample": f"NumPy FP64 restatement, fwd+adjoint, slope of 1- and 2-iteration runs at chi={chi}, "
                         f"D={D} ({total_s:.1f} s of CPU work)"}

    ms_mixed = results["mixed"]["ms"]
    ms_plain = results["plain_fp64"]["ms"]
    out = {
        "metric": METRIC,
        "value": round(ms_mixed, 4),
        "unit": "ms/iter",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_mixed, 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 contractions (MFMA, fp32 accumulate) + fp64 eigensolver/polish",
        "data": "synthetic: uniform [0,1) bulk from numpy PCG64 seed 0, symmetrised; Heisenberg(1,1,1)",
        "config": {"workload": f"CTMRG energy+gradient, iPEPS D={D}, d={s}, chi={chi}, corner matrix "
                               f"{chi * Dd}x{chi * Dd}, tol=0, maxit={a.steps}",
                   "chi": chi, "D": D, "d": s, "parallelism": "single GPU"},
        "plain_fp64_ms_per_step": round(ms_plain, 4),
        "speedup_vs_plain_fp64": round(ms_plain / ms_mixed, 3),
        "accuracy": acc,
        "stats": {k: v["stats"] for k, v in results.items()},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
