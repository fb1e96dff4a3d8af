"""op2hip_pcg_noalloc, the single CG solve that allocates no device memory
(its state lives in a static per-GPU array), against op2hip_pcg through the
C ABI (gpu): two calls in a row, the second on a smaller system, each bit
for bit the allocating entry point's solution, iteration count, residual
norm and reason; a system beyond the array's capacity is refused."""

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _system(n, seed):
    """An SPD pentadiagonal system with a varying diagonal (CSR arrays)."""
    rng = np.random.default_rng(seed)
    d = 4.5 + rng.random(n)
    rows, cols, vals = [0], [], []
    for i in range(n):
        for j, v in ((i - 2, -0.5), (i - 1, -1.0), (i, d[i]), (i + 1, -1.0), (i + 2, -0.5)):
            if 0 <= j < n:
                cols.append(j)
                vals.append(v)
        rows.append(len(cols))
    return (np.array(rows, dtype=np.int32), np.array(cols, dtype=np.int32),
            np.array(vals), rng.standard_normal(n))


def _solve(fn, n, arrs, torch, s):
    rp, ci, va, bb = arrs
    x = torch.full((n,), 7.0, dtype=torch.float64, device='cuda')      # overwritten
    work = torch.full((5 * n,), -3.0, dtype=torch.float64, device='cuda')
    its, rn, reason = ctypes.c_int(), ctypes.c_double(), ctypes.c_int()
    st = fn(n, _ptr(rp), _ptr(ci), _ptr(va), _ptr(bb), _ptr(x), _ptr(work), 1e-11, 1e-50, 1e5,
            5000, ctypes.byref(its), ctypes.byref(rn), ctypes.byref(reason), s)
    return st, x.cpu().numpy(), its.value, rn.value, reason.value


def test_noalloc_pcg_matches_pcg(hip):
    import torch
    from pyop2_amd import _native as nat
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    for n, seed in ((40000, 1), (2500, 2)):
        rows, cols, vals, b = _system(n, seed)
        dev = lambda a, t: torch.tensor(a, dtype=t, device='cuda')
        arrs = (dev(rows, torch.int32), dev(cols, torch.int32), dev(vals, torch.float64),
                dev(b, torch.float64))
        ref = _solve(nat.lib.op2hip_pcg, n, arrs, torch, s)
        got = _solve(nat.lib.op2hip_pcg_noalloc, n, arrs, torch, s)
        assert ref[0] == 0 and got[0] == 0, nat.last_error()
        assert ref[4] > 0 and ref[2] > 0
        assert got[2:] == ref[2:]
        assert np.array_equal(got[1], ref[1])


def test_noalloc_pcg_refuses_oversize(hip):
    import torch
    from pyop2_amd import _native as nat
    mx = nat.lib.op2hip_pcg_noalloc_max_rows()
    assert mx == 1 << 24
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    one = torch.zeros(8, dtype=torch.float64, device='cuda')
    idx = torch.zeros(8, dtype=torch.int32, device='cuda')
    its, rn, reason = ctypes.c_int(), ctypes.c_double(), ctypes.c_int()
    st = nat.lib.op2hip_pcg_noalloc(mx + 1, _ptr(idx), _ptr(idx), _ptr(one), _ptr(one),
                                    _ptr(one), _ptr(one), 1e-10, 1e-50, 1e5, 10,
                                    ctypes.byref(its), ctypes.byref(rn), ctypes.byref(reason), s)
    assert st != 0 and 'at most' in nat.last_error()
