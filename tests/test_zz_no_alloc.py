"""tnad_symeig_noalloc (the entry point that allocates nothing) against tnad_symeig, through the C ABI:
two calls in a row on device buffers, the second with a smaller matrix, and the host-side checks that
reject what the entry point cannot take before any kernel runs."""
import ctypes as C
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _sym(n, seed):
    r = np.random.default_rng(seed)
    A = r.standard_normal((n, n))
    return np.asfortranarray(A + A.T)


def _symeig_pool(ctx, A):
    # the existing entry point (device work space from the context's pool), host buffers
    n = A.shape[0]
    ev = np.zeros(n)
    U = np.zeros((n, n), order="F")
    ko = C.c_int(0)
    ctx.check(ctx.lib.tnad_symeig(ctx.h, C.c_void_p(A.ctypes.data), n, n, n, C.c_void_p(ev.ctypes.data),
                                  C.c_void_p(U.ctypes.data), C.byref(ko)))
    assert ko.value == n
    return ev, U


def _symeig_noalloc(ctx, A):
    import torch
    n = A.shape[0]
    dev = torch.device("cuda", ctx.device)
    a_t = torch.from_numpy(np.ascontiguousarray(A.ravel(order="F"))).to(dev)
    ev_t = torch.full((n,), np.nan, dtype=torch.float64, device=dev)
    u_t = torch.full((n * n,), np.nan, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    ctx.check(ctx.lib.tnad_symeig_noalloc(ctx.h, C.c_void_p(a_t.data_ptr()), n, C.c_void_p(ev_t.data_ptr()),
                                          C.c_void_p(u_t.data_ptr())))
    ev = ev_t.cpu().numpy()
    U = u_t.cpu().numpy().reshape((n, n), order="F")
    return ev, U


def test_noalloc_matches_symeig_on_two_calls(ctx):
    # first call at the largest size, then a smaller one: the scratch array keeps the first call's
    # rotation record, and the second call must not read any of it
    for n, seed in ((192, 11), (57, 12)):
        _say(f"case: symeig_noalloc n={n} seed={seed}")
        A = _sym(n, seed)
        ev0, U0 = _symeig_pool(ctx, A)
        ev1, U1 = _symeig_noalloc(ctx, A)
        # the same kernels on the same input: identical results
        assert np.array_equal(ev0, ev1), np.max(np.abs(ev0 - ev1))
        assert np.array_equal(U0, U1), np.max(np.abs(U0 - U1))
        # and a valid decomposition (LAPACK eigenvalues sorted by |lambda|)
        ref = np.linalg.eigvalsh(A)
        ref = ref[np.argsort(-np.abs(ref), kind="stable")]
        assert np.max(np.abs(np.abs(ev1) - np.abs(ref))) < 1e-12 * np.abs(ref).max()
        assert np.allclose(A @ U1, U1 * ev1[None, :], atol=1e-11 * np.abs(ref).max())


def test_noalloc_rejects_what_it_cannot_take(ctx):
    import torch
    from tnad import TnadError
    _say("case: symeig_noalloc argument checks")
    dev = torch.device("cuda", ctx.device)
    a_t = torch.zeros(200 * 200, dtype=torch.float64, device=dev)
    e_t = torch.zeros(200, dtype=torch.float64, device=dev)
    u_t = torch.zeros(200 * 200, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    p = lambda t: C.c_void_p(t.data_ptr())
    # n above the Jacobi size: refused on the host
    with pytest.raises(TnadError) as ei:
        ctx.check(ctx.lib.tnad_symeig_noalloc(ctx.h, p(a_t), 193, p(e_t), p(u_t)))
    assert ei.value.code == -6
    # host memory where device memory is required
    A = _sym(8, 1)
    with pytest.raises(TnadError) as ei:
        ctx.check(ctx.lib.tnad_symeig_noalloc(ctx.h, C.c_void_p(A.ctypes.data), 8, p(e_t), p(u_t)))
    assert ei.value.code == -1
