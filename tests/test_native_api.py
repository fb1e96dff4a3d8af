"""The convenience C entry points a binding of the reference would call
(gpu): one-call module compile with a named Const, and the single-system
PCG, which gives the batch engine's result bit for bit and the direct
solution within its tolerance."""

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def test_module_compile_and_const(hip):
    import torch
    from pyop2_amd import _native as nat, jit
    src = r'''
__constant__ double scale;
extern "C" __global__ void fill(double* y, long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = scale * (double)i;
}'''
    opts = [o.encode() for o in jit.HIPRTC_OPTS]
    copts = (ctypes.c_char_p * len(opts))(*opts)
    mod, fn = ctypes.c_void_p(), ctypes.c_void_p()
    nat.check(nat.lib.op2hip_module_compile(src.encode(), b'fill', copts, len(opts),
                                            jit.cache_dir().encode(), ctypes.byref(mod),
                                            ctypes.byref(fn)), 'compile')
    try:
        s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        v = np.array([2.5])
        nat.check(nat.lib.op2hip_module_set_const(mod, b'scale', v.ctypes.data_as(ctypes.c_void_p),
                                                  8, s), 'const')
        n = 1000
        y = torch.zeros(n, dtype=torch.float64, device='cuda')
        argv = (ctypes.c_uint64 * 2)(y.data_ptr(), n)
        nat.check(nat.lib.op2hip_launch(fn, (n + 255) // 256, 256, 0, s, argv, 2), 'launch')
        torch.cuda.synchronize()
        assert np.array_equal(y.cpu().numpy(), 2.5 * np.arange(n))
        # a Const larger than the symbol is refused
        big = np.zeros(4)
        assert nat.lib.op2hip_module_set_const(mod, b'scale', big.ctypes.data_as(ctypes.c_void_p),
                                               32, s) != 0
    finally:
        nat.lib.op2hip_module_destroy(mod)


def test_single_pcg_matches_batch_and_direct(hip):
    import torch
    from pyop2_amd import _native as nat, cg
    n = 3000
    # SPD tridiagonal with a varying diagonal
    d = 2.0 + np.linspace(0.0, 1.0, n)
    rows = [0]
    cols, vals = [], []
    for i in range(n):
        for j, v in ((i - 1, -1.0), (i, d[i]), (i + 1, -1.0)):
            if 0 <= j < n:
                cols.append(j)
                vals.append(v)
        rows.append(len(cols))
    b = np.sin(np.arange(n) * 0.01)
    dev = lambda a, t: torch.tensor(np.asarray(a), dtype=t, device='cuda')
    rp, ci, va, bb = dev(rows, torch.int32), dev(cols, torch.int32), dev(vals, torch.float64), \
        dev(b, torch.float64)
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    x1 = torch.zeros(n, dtype=torch.float64, device='cuda')
    work = torch.zeros(5 * n, dtype=torch.float64, device='cuda')
    its, rn, reason = ctypes.c_int(), ctypes.c_double(), ctypes.c_int()
    nat.check(nat.lib.op2hip_pcg(n, _ptr(rp), _ptr(ci), _ptr(va), _ptr(bb), _ptr(x1), _ptr(work),
                                 1e-12, 1e-50, 1e5, 10000, ctypes.byref(its), ctypes.byref(rn),
                                 ctypes.byref(reason), s), 'pcg')
    assert reason.value > 0 and its.value > 0
    # the batch engine with one system gives the same x bit for bit
    x2 = torch.zeros(n, dtype=torch.float64, device='cuda')
    w2 = torch.zeros((5, n), dtype=torch.float64, device='cuda')
    sysd = cg.make_system(rp.data_ptr(), ci.data_ptr(), va.data_ptr(), bb.data_ptr(),
                          x2.data_ptr(), [t.data_ptr() for t in w2], 0, n, 1e-12, 1e-50, 1e5,
                          10000)
    out = cg.CGBatch(1, n).solve([sysd], s.value)
    assert out[0].its == its.value
    assert np.array_equal(x1.cpu().numpy(), x2.cpu().numpy())
    A = np.zeros((n, n))
    for i in range(n):
        A[i, cols[rows[i]:rows[i + 1]]] = vals[rows[i]:rows[i + 1]]
    xd = np.linalg.solve(A, b)
    assert np.abs(x1.cpu().numpy() - xd).max() <= 1e-9 * np.abs(xd).max()
