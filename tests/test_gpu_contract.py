"""GPU parity of the matrix-core contraction kernel (tnad_contract) against NumPy
einsum: plain GEMM shapes (tile edges, K not a multiple of the K-tile) and every
strided contraction pattern of the CTMRG step, its adjoint and the energy chain."""
import zlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = {"fp64": 1e-13, "fp32": 2e-6, "bf16": 2e-2}


def _rel(a, b):
    return np.max(np.abs(a - b)) / max(1e-300, np.max(np.abs(b)))


@pytest.mark.parametrize("prec", ["fp64", "fp32", "bf16"])
@pytest.mark.parametrize("m,n,k", [(16, 16, 4), (64, 64, 16), (70, 33, 45), (1, 5, 3), (130, 200, 300)])
def test_plain_gemm(ctx, prec, m, n, k):
    import tnad
    print(f"case: gemm {prec} {m}x{n}x{k}", flush=True)
    rng = np.random.default_rng(m * 1000 + n + k)
    A = rng.standard_normal((m, k))
    B = rng.standard_normal((k, n))
    Cm = tnad.contract("ik,kj->ij", A, B, precision=prec, ctx=ctx)
    assert _rel(Cm, A @ B) < TOL[prec]


SPECS = [
    "iba,ad->ibd", "ibd,dcl->ibcl", "ibcl,jkcb->ijlk", "aed,abx->edbx", "edbx,bjce->dcxj",
    "dcxj,dcy->xjy", "abx,xjy->abjy", "abjy,bjce->aecy", "aecy,dcy->aed", "abjy,aecy->bjce",
    "ijlk,jkcb->ibcl", "ijlk,ibcl->jkcb", "ibcl,dcl->ibd", "ibd,ibcl->dcl", "iba,ibd->ad",
    "abcij,ijkl->abckl",
]


def _dims(spec, chi, D, s):
    ext = {}
    for ch in "abcdefgijklmxy":
        ext[ch] = D
    lhs, out = spec.split("->")
    ia, ib = lhs.split(",")
    return ia, ib, out


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("spec", SPECS)
def test_step_contractions(ctx, prec, spec):
    import tnad
    print(f"case: contract {prec} {spec}", flush=True)
    rng = np.random.default_rng(zlib.crc32(spec.encode()))   # reproducible across processes
    chi, D = 7, 3
    # extents: indices that carry the environment bond get chi, the others D
    chi_idx = {"iba,ad->ibd": "iad", "ibd,dcl->ibcl": "idl", "ibcl,jkcb->ijlk": "il",
               "aed,abx->edbx": "adx", "edbx,bjce->dcxj": "dx", "dcxj,dcy->xjy": "dxy",
               "abx,xjy->abjy": "axy", "abjy,bjce->aecy": "ay", "aecy,dcy->aed": "ady",
               "abjy,aecy->bjce": "ay", "ijlk,jkcb->ibcl": "il", "ijlk,ibcl->jkcb": "il",
               "ibcl,dcl->ibd": "idl", "ibd,ibcl->dcl": "idl", "iba,ibd->ad": "iad",
               "abcij,ijkl->abckl": "ac"}[spec]
    lhs, out = spec.split("->")
    ia, ib = lhs.split(",")
    ext = {ch: (chi if ch in chi_idx else D) for ch in set(ia + ib + out)}
    if spec == "abcij,ijkl->abckl":
        ext.update({"i": 2, "j": 2, "k": 2, "l": 2})
    A = rng.standard_normal([ext[c] for c in ia])
    B = rng.standard_normal([ext[c] for c in ib])
    got = tnad.contract(spec, A, B, precision=prec, ctx=ctx)
    ref = np.einsum(spec, A, B)
    assert _rel(got, ref) < 10 * TOL[prec]


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("tr", ["NN", "NT", "TN"])
def test_gemm_alpha_beta_accumulate(ctx, prec, tr):
    import tnad
    print(f"case: gemm alpha/beta {prec} {tr}", flush=True)
    rng = np.random.default_rng(7)
    m, n, k = 40, 24, 10
    spec = {"NN": "ik,kj->ij", "NT": "ik,jk->ij", "TN": "ki,kj->ij"}[tr]
    A = rng.standard_normal((m, k) if tr[0] == "N" else (k, m))
    B = rng.standard_normal((k, n) if tr[1] == "N" else (n, k))
    C0 = rng.standard_normal((m, n))
    got = tnad.contract(spec, A, B, precision=prec, ctx=ctx, alpha=-0.5, beta=1.0, Cin=C0)
    ref = -0.5 * np.einsum(spec, A, B) + C0
    assert _rel(got, ref) < 10 * TOL[prec]


@pytest.mark.parametrize("prec", ["fp64", "fp32", "bf16"])
@pytest.mark.parametrize("m,n,k", [(70, 50, 2000), (192, 192, 2048), (300, 20, 777)])
def test_split_k_gemm_alpha_beta(ctx, prec, m, n, k):
    # few output tiles and a long K: the launcher cuts K into slices and reduces them
    import tnad
    print(f"case: split-K gemm {prec} {m}x{n}x{k}", flush=True)
    rng = np.random.default_rng(m + n + k)
    A = rng.standard_normal((k, m))
    B = rng.standard_normal((k, n))
    C0 = rng.standard_normal((m, n))
    got = tnad.contract("ki,kj->ij", A, B, precision=prec, ctx=ctx, alpha=0.75, beta=-1.25, Cin=C0)
    ref = 0.75 * (A.T @ B) - 1.25 * C0
    tol = {"fp64": 1e-13, "fp32": 2e-6, "bf16": 2e-2}[prec]
    assert _rel(got, ref) < 10 * tol


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_split_k_strided_bulk_cotangent(ctx, prec):
    # abar[j,k,c,b] = sum_{i,l} cpbar[i,j,l,k] X2[i,b,c,l]: a 256 x 256 output over K = chi^2
    import tnad
    print(f"case: split-K strided ijlk,ibcl->jkcb {prec}", flush=True)
    rng = np.random.default_rng(3)
    chi, D = 40, 4
    A = rng.standard_normal((chi, D, chi, D))
    B = rng.standard_normal((chi, D, D, chi))
    got = tnad.contract("ijlk,ibcl->jkcb", A, B, precision=prec, ctx=ctx)
    ref = np.einsum("ijlk,ibcl->jkcb", A, B)
    assert _rel(got, ref) < 10 * TOL[prec]


@pytest.mark.parametrize("prec", ["fp64", "fp32", "bf16"])
@pytest.mark.parametrize("tr", ["NN", "NT", "TN", "TT"])
@pytest.mark.parametrize("m,n,k", [(256, 192, 512), (2048, 128, 2048), (136, 72, 48), (64, 64, 16),
                                   (2048, 192, 192), (192, 192, 2048), (96, 40, 64)])
def test_gemm_vector_loads(ctx, prec, tr, m, n, k):
    # sizes whose unit-stride index and K are multiples of the 16-byte vector and of the K-tile take the
    # vector-load kernel: 64 x 64 tiles (also with split-K at 2048 x 128 x 2048, and with partial output
    # tiles at 136 x 72), 32 x 32 tiles for small grids (2048 x 192 x 192; split-K at 192 x 192 x 2048;
    # partial tiles at 96 x 40); BF16 on the double-K instruction wherever K is a multiple of 32;
    # alpha / beta applied to a strided output
    import tnad
    print(f"case: vector-load gemm {prec} {tr} {m}x{n}x{k}", flush=True)
    rng = np.random.default_rng(m + 3 * n + 7 * k + len(tr))
    spec = {"NN": "ik,kj->ij", "NT": "ik,jk->ij", "TN": "ki,kj->ij", "TT": "ki,jk->ij"}[tr]
    A = rng.standard_normal((m, k) if tr[0] == "N" else (k, m))
    B = rng.standard_normal((k, n) if tr[1] == "N" else (n, k))
    C0 = rng.standard_normal((m, n))
    got = tnad.contract(spec, A, B, precision=prec, ctx=ctx, alpha=0.5, beta=-1.0, Cin=C0)
    ref = 0.5 * np.einsum(spec, A, B) - C0
    assert _rel(got, ref) < 10 * TOL[prec]


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("spec", SPECS[:15])
def test_step_contractions_vector_loads(ctx, prec, spec):
    # the CTMRG step's contractions at chi = 32, D = 4 (double-layer bond 16): unit-stride runs and K
    # sub-extents of 16 and 32 take the vector-load kernel where the layout allows
    import tnad
    print(f"case: vector-load contract {prec} {spec}", flush=True)
    rng = np.random.default_rng(zlib.crc32(spec.encode()) + 1)
    chi_idx = {"iba,ad->ibd": "iad", "ibd,dcl->ibcl": "idl", "ibcl,jkcb->ijlk": "il",
               "aed,abx->edbx": "adx", "edbx,bjce->dcxj": "dx", "dcxj,dcy->xjy": "dxy",
               "abx,xjy->abjy": "axy", "abjy,bjce->aecy": "ay", "aecy,dcy->aed": "ady",
               "abjy,aecy->bjce": "ay", "ijlk,jkcb->ibcl": "il", "ijlk,ibcl->jkcb": "il",
               "ibcl,dcl->ibd": "idl", "ibd,ibcl->dcl": "idl", "iba,ibd->ad": "iad"}[spec]
    lhs, out = spec.split("->")
    ia, ib = lhs.split(",")
    ext = {ch: (32 if ch in chi_idx else 16) for ch in set(ia + ib + out)}
    A = rng.standard_normal([ext[c] for c in ia])
    B = rng.standard_normal([ext[c] for c in ib])
    got = tnad.contract(spec, A, B, precision=prec, ctx=ctx)
    ref = np.einsum(spec, A, B)
    assert _rel(got, ref) < 10 * TOL[prec]


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("tr", ["NN", "NT", "TN", "TT"])
def test_gemm_vector_loads_large_tiles(ctx, prec, tr):
    # 4096 x 4096 outputs (1024 tiles of 128 x 128: the large-tile vector-load kernel)
    import tnad
    print(f"case: vector-load gemm 128-tiles {prec} {tr}", flush=True)
    m, n, k = 4096, 4096, 48
    rng = np.random.default_rng(11 + len(tr))
    spec = {"NN": "ik,kj->ij", "NT": "ik,jk->ij", "TN": "ki,kj->ij", "TT": "ki,jk->ij"}[tr]
    A = rng.standard_normal((m, k) if tr[0] == "N" else (k, m))
    B = rng.standard_normal((k, n) if tr[1] == "N" else (n, k))
    got = tnad.contract(spec, A, B, precision=prec, ctx=ctx)
    ref = np.einsum(spec, A, B, optimize=True)
    assert _rel(got, ref) < 10 * TOL[prec]


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_enlarged_corner_at_chi128(ctx, prec):
    # the bench's enlarged corner, ein"ibcl,jkcb->ijlk" at chi = 128 and double-layer bond 16 (a 16384 x 256
    # output over K = 256, unit-stride i runs of 128 and K0 = 16: the vector-load kernel)
    import tnad
    print(f"case: enlarged corner chi=128 {prec}", flush=True)
    rng = np.random.default_rng(5)
    chi, Dd = 128, 16
    X2 = rng.standard_normal((chi, Dd, Dd, chi))
    a = rng.standard_normal((Dd, Dd, Dd, Dd))
    got = tnad.contract("ibcl,jkcb->ijlk", X2, a, precision=prec, ctx=ctx)
    ref = np.einsum("ibcl,jkcb->ijlk", X2, a, optimize=True)
    assert _rel(got, ref) < 10 * TOL[prec]


def test_time_jacobi_runs_the_requested_sweeps(ctx):
    # the Jacobi timing entry: a random 192 x 192 matrix does not converge in 3 sweeps, so the
    # rotation kernel runs 3 x 191 rounds; an odd size is padded to even (101 -> 102, 101 rounds per sweep)
    import tnad
    print("case: time_jacobi", flush=True)
    ms, rounds = tnad.time_jacobi(192, 3, reps=2, ctx=ctx)
    assert rounds == 3 * 191 and 0.0 < ms < 100.0
    ms, rounds = tnad.time_jacobi(101, 2, reps=1, ctx=ctx)
    assert rounds == 2 * 101 and ms > 0.0


def test_time_syev_reports_its_three_stages(ctx):
    # the eigensolver roofline's timing entry: positive device times for the tridiagonalisation, the divide
    # and conquer and the back-transformation of one solve; sizes below 3 are refused
    import tnad
    print("case: time_syev", flush=True)
    for m in (176, 37):
        t = tnad.time_syev(m, reps=2, ctx=ctx)
        assert len(t) == 3 and all(0.0 < x < 100.0 for x in t)
    with pytest.raises(tnad.TnadError):
        tnad.time_syev(2, reps=1, ctx=ctx)
