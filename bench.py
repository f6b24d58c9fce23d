#!/usr/bin/env python
"""CTMRG forward + adjoint on one MI355X: ms per CTMRG iteration at chi=128, D=4
(iPEPS bond 4, double-layer bond 16, corner matrix 2048 x 2048), mixed precision
vs plain FP64, with the energy-accuracy condition |dE| < 1e-7.

A "step" is one energy + gradient evaluation, the unit an optimiser consumes:
one fused library call (tnad_energy with a gradient output) that symmetrises
the bulk, builds the double layer, runs CTMRG from the :raw environment to
tol = 1e-10 (maxit = 100), evaluates the energy and runs the full reverse
sweep, on device-resident inputs and outputs.  ms/iter = ms per step divided
by the CTMRG iterations the step took; `value` is that figure for the mixed
path.  The plain path is FP64 everywhere with a cold-started eigensolver in
every iteration (the reference's semantics); the mixed path runs the
contractions in FP32 until the singular values settle (switch_tol 1e-5), warm
starts the eigensolver from the previous iteration, and runs the reverse
sweep's contractions in FP32.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu-baseline] [--no-accuracy]

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
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
# HBM bytes per launch of the roofline kernel (FP32 ein"ibcl,jkcb->ijlk" at chi=128, Dd=16) from two
# rocprofv3 --pmc runs of `bench.py --kernel-only` (FETCH_SIZE, then WRITE_SIZE, KiB per dispatch, mean of
# 21 launches), each scaled by what the counter reads for a copy of known size in the same run: the 256 MiB
# device copy reads FETCH_SIZE = 131089 KiB and WRITE_SIZE = 262144 KiB, so fetch x 2.0, write x 1.0
# (the FP64->FP32 cast of 32 MiB gives the same fetch factor).  9317.3 KiB x 2 + 16384 KiB = 35019 KiB.
TRAFFIC = 35019 * 1024


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


def cpu_baseline(chi, D, s, seed, tol, maxit):
    """The FP64 NumPy restatement (oracle) on the host: one energy + gradient evaluation of the
    same workload (same bulk, tol, maxit).  Returns (ms per CTMRG iteration, seconds, iterations,
    energy, gradient)."""
    from oracle import tnad_oracle as O
    h = O.hamiltonian_heisenberg()
    b = O.symmetrized_random_bulk(D, s, seed, "uniform")
    t0 = time.perf_counter()
    e, g, it = O.energy_and_grad(h, b, chi, tol, maxit)
    dt = time.perf_counter() - t0
    return dt * 1e3 / max(it, 1), dt, it, e, g


def blas_threads():
    """threads the host BLAS actually uses (what the CPU leg runs on)"""
    try:
        from threadpoolctl import threadpool_info
        n = [p.get("num_threads", 0) for p in threadpool_info() if p.get("user_api") == "blas"]
        if n:
            return int(max(n))
    except Exception:
        pass
    return os.cpu_count()


def roofline_leg(tnad, ctx, chi, Dd):
    """dominant contraction of the step, timed alone with HIP events on the library stream:
    the enlarged corner cp = X2 (x) a, ein"ibcl,jkcb->ijlk", 2 chi^2 Dd^4 flops per launch"""
    spec = "ibcl,jkcb->ijlk"
    ext = {"i": chi, "l": chi, "b": Dd, "c": Dd, "j": Dd, "k": Dd}
    kflops = 2.0 * chi * chi * Dd ** 4
    roof = {}
    for prec in ("fp64", "fp32"):
        ms = tnad.time_contract(spec, ext, precision=prec, reps=20, ctx=ctx)
        roof[prec] = dict(ms=ms, tflops=kflops / (ms * 1e-3) / 1e12)
    return {"bound": "mfma", "kernel": f"tc_kernel_vec<float, 64, 16> ein\"{spec}\" (fp32, {kflops:.3e} flop/launch)",
            "achieved": round(roof["fp32"]["tflops"], 3), "peak": PEAK["fp32"], "unit": "TFLOP/s",
            "frac": round(roof["fp32"]["tflops"] / PEAK["fp32"], 4), "traffic": TRAFFIC,
            "avg_launch_ms": round(roof["fp32"]["ms"], 5),
            "fp64_variant": {"achieved": round(roof["fp64"]["tflops"], 3), "peak": PEAK["fp64"],
                             "frac": round(roof["fp64"]["tflops"] / PEAK["fp64"], 4),
                             "avg_launch_ms": round(roof["fp64"]["ms"], 5)}}


# FP64 vector peak of one CU (of 256): the eigensolver's one-workgroup kernels run on one CU
CU_FP64_GFLOPS = 78.6e3 / 256


def eigensolver_leg(tnad, ctx, m=192):
    """The exact eigensolver of the Rayleigh-Ritz steps (divide and conquer, m = the subspace block at
    chi = 128), its stages timed alone with HIP events on the library stream.  The Householder
    tridiagonalisation (one workgroup, the matrix in registers) has 4/3 m^3 algorithmic flops; it is bound by
    the per-step latency of its two barriers and dependent LDS / reduction / FP64 divide chain, far below the
    FP64 peak of the one CU it runs on."""
    trd, dc, back = tnad.time_syev(m, reps=10, ctx=ctx)
    flops = 4.0 / 3.0 * m ** 3
    gf = flops / (trd * 1e-3) / 1e9
    return {"bound": "fp64 (one CU)", "kernel": f"trd_kernel (Householder tridiagonalisation, m={m}, one workgroup)",
            "achieved": round(gf, 2), "peak": round(CU_FP64_GFLOPS, 1), "unit": "GFLOP/s (FP64 of one CU)",
            "frac": round(gf / CU_FP64_GFLOPS, 4), "flops": round(flops),
            "avg_launch_ms": round(trd, 4),
            "solve_ms": {"tridiagonalisation": round(trd, 4), "divide_and_conquer": round(dc, 4),
                         "back_transformation_and_sort": round(back, 4)}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--chi", type=int, default=128)
    ap.add_argument("--D", type=int, default=4)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--maxit", type=int, default=100)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--kernel-only", action="store_true",
                    help="only the roofline leg (dominant contraction, 20 launches) and a 256 MiB device copy "
                         "of known size, for hardware-counter runs")
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
    if a.kernel_only:
        roofline = roofline_leg(tnad, ctx, chi, Dd)
        roofline["eigensolver"] = eigensolver_leg(tnad, ctx)
        # a plain copy of known size (256 MiB read + 256 MiB written) to scale FETCH_SIZE / WRITE_SIZE
        x = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device=dev)
        y = torch.empty_like(x)
        y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"roofline": roofline, "copy_bytes_each_way": x.numel() * 4}))
        return
    h = heisenberg_h()
    bulk = synthetic_bulk(D, s, a.seed)
    # inputs resident in HBM: column-major buffers behind torch-owned device memory
    h_t = torch.from_numpy(np.ascontiguousarray(h.ravel(order="F"))).to(dev)
    b_t = torch.from_numpy(np.ascontiguousarray(bulk.ravel(order="F"))).to(dev)
    torch.cuda.synchronize()

    plain = dict()
    mixed = dict(precision="fp32", switch_tol=1e-5, eig_warm=True, bwd_precision="fp32")
    results = {}
    grads = {}
    for name, opts in (("plain_fp64", plain), ("mixed", mixed)):
        g_t = torch.zeros_like(b_t)
        for _ in range(a.warmup):
            run_energy(tnad, ctx, h_t.data_ptr(), b_t.data_ptr(), D, s, chi, a.maxit, a.tol, opts, g_t.data_ptr())
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        its = []
        for _ in range(a.steps):
            e, it = run_energy(tnad, ctx, h_t.data_ptr(), b_t.data_ptr(), D, s, chi, a.maxit, a.tol, opts,
                               g_t.data_ptr())
            its.append(it)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        ms_step = dt * 1e3 / a.steps
        results[name] = dict(ms_step=ms_step, ms_iter=ms_step / (sum(its) / len(its)), iters=its[-1], e=e,
                             stats=ctx.stats())
        grads[name] = np.reshape(g_t.cpu().numpy(), bulk.shape, order="F")
        print(f"[bench] {name}: {ms_step:.2f} ms per energy+gradient, {its[-1]} iterations", file=sys.stderr,
              flush=True)

    acc = {}
    if not a.no_accuracy:
        g64, gmx = grads["plain_fp64"], grads["mixed"]
        acc = dict(e_fp64=results["plain_fp64"]["e"], e_mixed=results["mixed"]["e"],
                   dE=abs(results["mixed"]["e"] - results["plain_fp64"]["e"]),
                   grad_rel_err=float(np.linalg.norm(gmx - g64) / np.linalg.norm(g64)),
                   iters_fp64=results["plain_fp64"]["iters"], iters_mixed=results["mixed"]["iters"])

    roofline = roofline_leg(tnad, ctx, chi, Dd)
    eig_roof = eigensolver_leg(tnad, ctx)

    cpu = None
    if not a.no_cpu_baseline:
        per_iter_ms, total_s, it_o, e_o, g_o = cpu_baseline(chi, D, s, a.seed, a.tol, a.maxit)
        cpu = {"value": round(per_iter_ms, 1), "unit": "ms/iter", "cores": blas_threads(), "kind": "port",
               "sample": f"NumPy/OpenBLAS FP64 restatement, one energy+gradient evaluation of the same "
                         f"workload ({it_o} CTMRG iterations, {total_s:.1f} s of CPU work)"}
        if not a.no_accuracy:
            gn = np.linalg.norm(g_o)
            for name in ("plain_fp64", "mixed"):
                acc[f"dE_vs_oracle_{name}"] = abs(results[name]["e"] - e_o)
                acc[f"grad_rel_err_vs_oracle_{name}"] = float(np.linalg.norm(grads[name] - g_o) / gn)
            acc["iters_oracle"] = it_o

    ms_mixed = results["mixed"]["ms_iter"]
    ms_plain = results["plain_fp64"]["ms_iter"]
    out = {
        "metric": METRIC,
        "value": round(ms_mixed, 4),
        "unit": "ms/iter",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(results["mixed"]["ms_step"], 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 contractions (MFMA, fp32 accumulate) + fp64 eigensolver, eigen adjoint and polish",
        "data": f"synthetic: uniform [0,1) bulk from numpy PCG64 seed {a.seed}, symmetrised; Heisenberg(1,1,1)",
        "config": {"workload": f"energy+gradient of a D={D}, d={s} iPEPS at chi={chi} (corner matrix "
                               f"{chi * Dd}x{chi * Dd}): CTMRG from :raw to tol={a.tol:g} (maxit={a.maxit}), "
                               f"energy, full reverse sweep; one step = one evaluation",
                   "chi": chi, "D": D, "d": s, "tol": a.tol, "maxit": a.maxit, "parallelism": "single GPU"},
        "plain_fp64_ms_per_iter": round(ms_plain, 4),
        "plain_fp64_ms_per_step": round(results["plain_fp64"]["ms_step"], 4),
        "speedup_vs_plain_fp64_per_iter": round(ms_plain / ms_mixed, 3),
        "speedup_vs_plain_fp64_per_step": round(results["plain_fp64"]["ms_step"] / results["mixed"]["ms_step"], 3),
        "accuracy": acc,
        "stats": {k: v["stats"] for k, v in results.items()},
        # device time of one evaluation by phase (HIP events on the library's stream): the CTMRG steps'
        # contractions and epilogues, their eigensolver, the reverse sweep, and the rest (symmetrisation,
        # double layer, expectation value, transfers)
        "phases_ms_per_step": {k: {"forward_contractions": round(v["stats"]["fwd_ms"] - v["stats"]["eig_ms"], 3),
                                   "eigensolver": round(v["stats"]["eig_ms"], 3),
                                   "reverse_sweep": round(v["stats"]["bwd_ms"], 3),
                                   "other": round(v["ms_step"] - v["stats"]["fwd_ms"] - v["stats"]["bwd_ms"], 3)}
                               for k, v in results.items()},
        "roofline": roofline,
        # the dominant kernel by time is the eigensolver's one-workgroup Jacobi rotation kernel, bound by the
        # LDS of the CU it runs on rather than by HBM or the matrix cores: its own roofline
        "eigensolver_roofline": eig_roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
