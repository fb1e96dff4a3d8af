"""Global reductions, as the reference's test/unit/test_global_reduction.py
states them (:40-400): direct MIN/MAX over uint32, int32, float32 and
float64 Dats; READ, INC, MIN and MAX Globals of dimension 1 and 2, with the
Global's initial value either beaten by the data or not; repeated loops
into one Global, with the Global reset in between, and into two Globals.
4096 elements, so device backends use many workgroups."""

import numpy as np
import pytest
from numpy.testing import assert_allclose

from pyop2_amd import op2

NELEMS = 4096

CTYPE = {np.uint32: 'unsigned int', np.int32: 'int', np.float32: 'float',
         np.float64: 'double'}


@pytest.fixture
def s1():
    return op2.Set(NELEMS, 1, 's1')


@pytest.fixture
def s2():
    return op2.Set(NELEMS, 2, 's2')


@pytest.fixture
def d1(s1):
    return op2.Dat(s1, np.arange(NELEMS) + 1, dtype=np.uint32)


@pytest.fixture
def d2(s2):
    return op2.Dat(s2, np.arange(2 * NELEMS) + 1, dtype=np.uint32)


def _k(body, name='k', t='unsigned int'):
    return op2.Kernel("void %s(%s *x, %s *g) { %s }" % (name, t, t, body), name)


MIN1 = "if (*x < *g) *g = *x;"
MAX1 = "if (*x > *g) *g = *x;"
INC1 = "*g += *x;"
MIN2 = "if (x[0] < g[0]) g[0] = x[0]; if (x[1] < g[1]) g[1] = x[1];"
MAX2 = "if (x[0] > g[0]) g[0] = x[0]; if (x[1] > g[1]) g[1] = x[1];"
INC2 = "g[0] += x[0]; g[1] += x[1];"


@pytest.mark.parametrize('dtype,value,mode,g0,expect', [
    (np.uint32, 12, 'MIN', 8, 8),
    (np.int32, -12, 'MIN', 8, -12),
    (np.int32, -12, 'MAX', -42, -12),
    (np.float32, -12.0, 'MIN', -0.8, -12.0),
    (np.float32, -12.0, 'MAX', -42.8, -12.0),
    (np.float64, -12.0, 'MIN', -0.8, -12.0),
    (np.float64, -12.0, 'MAX', -42.8, -12.0)])
def test_direct_min_max(backend, s1, dtype, value, mode, g0, expect):
    d = op2.Dat(s1, [value] * NELEMS, dtype)
    g = op2.Global(1, g0, dtype, "g")
    t = CTYPE[dtype]
    op2.par_loop(_k(MIN1 if mode == 'MIN' else MAX1, 'kernel_' + mode.lower(), t), s1,
                 d(op2.IdentityMap, op2.READ), g(getattr(op2, mode)))
    assert_allclose(g.data[0], expect)


def test_1d_read(backend, s1, d1):
    g = op2.Global(1, 1, dtype=np.uint32)
    op2.par_loop(_k("*x = *g;"), s1, d1(op2.IdentityMap, op2.WRITE), g(op2.READ))
    assert all(d1.data_ro == g.data)


def test_2d_read(backend, s1, d1):
    g = op2.Global(2, (1, 2), dtype=np.uint32)
    op2.par_loop(_k("*x = g[0] + g[1];"), s1, d1(op2.IdentityMap, op2.WRITE), g(op2.READ))
    assert all(d1.data_ro == g.data.sum())


def test_1d_inc(backend, s1, d1):
    g = op2.Global(1, 0, dtype=np.uint32)
    op2.par_loop(_k(INC1), s1, d1(op2.IdentityMap, op2.READ), g(op2.INC))
    assert g.data[0] == d1.data_ro.sum()


@pytest.mark.parametrize('mode,dat_wins', [('MIN', True), ('MIN', False),
                                           ('MAX', True), ('MAX', False)])
def test_1d_min_max(backend, s1, d1, mode, dat_wins):
    """The result is the data's extremum when it beats the Global's initial
    value, else that initial value."""
    if mode == 'MIN':
        if not dat_wins:
            d1.data[:] += 10
        ext = int(d1.data_ro.min())
        val = ext + 1 if dat_wins else ext - 1
    else:
        ext = int(d1.data_ro.max())
        val = ext - 1 if dat_wins else ext + 1
    g = op2.Global(1, val, dtype=np.uint32)
    op2.par_loop(_k(MIN1 if mode == 'MIN' else MAX1), s1, d1(op2.IdentityMap, op2.READ),
                 g(getattr(op2, mode)))
    assert g.data[0] == (ext if dat_wins else val)


def test_2d_inc(backend, s1, d2):
    g = op2.Global(2, (0, 0), dtype=np.uint32)
    op2.par_loop(_k(INC2), s1, d2(op2.IdentityMap, op2.READ), g(op2.INC))
    assert g.data[0] == d2.data_ro[:, 0].sum()
    assert g.data[1] == d2.data_ro[:, 1].sum()


@pytest.mark.parametrize('mode,dat_wins', [('MIN', True), ('MIN', False),
                                           ('MAX', True), ('MAX', False)])
def test_2d_min_max(backend, s1, d2, mode, dat_wins):
    if mode == 'MIN' and not dat_wins:
        d2.data[:, 0] += 10
        d2.data[:, 1] += 10
    red = np.min if mode == 'MIN' else np.max
    ext = [int(red(d2.data_ro[:, c])) for c in range(2)]
    step = 1 if (mode == 'MIN') == dat_wins else -1
    val = [e + step for e in ext]
    g = op2.Global(2, tuple(val), dtype=np.uint32)
    op2.par_loop(_k(MIN2 if mode == 'MIN' else MAX2), s1, d2(op2.IdentityMap, op2.READ),
                 g(getattr(op2, mode)))
    assert list(g.data) == (ext if dat_wins else val)


def test_1d_multi_inc_same_global(backend, s1, d1):
    g = op2.Global(1, 0, dtype=np.uint32)
    for k in (1, 2):
        op2.par_loop(_k(INC1), s1, d1(op2.IdentityMap, op2.READ), g(op2.INC))
        assert g.data[0] == k * d1.data_ro.sum()


def test_1d_multi_inc_same_global_reset(backend, s1, d1):
    g = op2.Global(1, 0, dtype=np.uint32)
    op2.par_loop(_k(INC1), s1, d1(op2.IdentityMap, op2.READ), g(op2.INC))
    assert g.data[0] == d1.data_ro.sum()
    g.data = 10
    op2.par_loop(_k(INC1), s1, d1(op2.IdentityMap, op2.READ), g(op2.INC))
    assert g.data[0] == d1.data_ro.sum() + 10


def test_1d_multi_inc_diff_global(backend, s1, d1):
    g = op2.Global(1, 0, dtype=np.uint32)
    g2 = op2.Global(1, 10, dtype=np.uint32)
    op2.par_loop(_k(INC1), s1, d1(op2.IdentityMap, op2.READ), g(op2.INC))
    assert g.data[0] == d1.data_ro.sum()
    op2.par_loop(_k(INC1), s1, d1(op2.IdentityMap, op2.READ), g2(op2.INC))
    assert g2.data[0] == d1.data_ro.sum() + 10
    assert g.data[0] == d1.data_ro.sum()


def test_new_global_each_call(backend, s1, d1):
    """A driver building a fresh Global per call (demo/airfoil.py:124-130
    makes ``rms`` inside its loop): each call reduces into its own Global
    and reads its own READ Global; a repeated call may reuse the planned
    launch, but never a previous call's Global."""
    k = op2.Kernel("void kg(unsigned int *x, unsigned int *a, unsigned int *g) "
                   "{ *g += *x * *a; }", 'kg')
    total = int(d1.data_ro.sum())
    for it in range(5):
        a = op2.Global(1, it + 1, dtype=np.uint32)
        g = op2.Global(1, 7 * it, dtype=np.uint32)
        op2.par_loop(k, s1, d1(op2.IdentityMap, op2.READ), a(op2.READ), g(op2.INC))
        assert g.data[0] == (total * (it + 1) + 7 * it) % 2 ** 32
        assert a.data[0] == it + 1
    if backend == 'hip':
        from pyop2_amd import hip as H
        assert any(e[1].globals_red for e in H._fast_cache.values())


def test_new_global_aliasing(backend, s1, d1):
    """One Global at two positions, then two Globals there: the second
    call is not taken for the first one's repeat."""
    k = op2.Kernel("void kh(unsigned int *x, unsigned int *g, unsigned int *h) "
                   "{ *g += *x; *h += 1; }", 'kh')
    total = int(d1.data_ro.sum())
    for _ in range(2):
        g = op2.Global(1, 0, dtype=np.uint32)
        op2.par_loop(k, s1, d1(op2.IdentityMap, op2.READ), g(op2.INC), g(op2.INC))
        assert g.data[0] == total + NELEMS
        g1 = op2.Global(1, 0, dtype=np.uint32)
        g2 = op2.Global(1, 3, dtype=np.uint32)
        op2.par_loop(k, s1, d1(op2.IdentityMap, op2.READ), g1(op2.INC), g2(op2.INC))
        assert g1.data[0] == total and g2.data[0] == NELEMS + 3
