"""Dat arithmetic and norm, as the reference's test/unit/test_linalg.py
states them (:60-300): binary operators return a new Dat with the left
operand's dtype, in-place operators keep it, Dats of different shape are
refused with ValueError, scalars broadcast, ``norm`` is the 2-norm.  The
reference sets ``_data`` directly; these tests write through ``.data``,
which every backend keeps coherent with its device copy."""

import numpy as np
import pytest

from pyop2_amd import op2

NELEMS = 8


@pytest.fixture
def s():
    return op2.Set(NELEMS)


@pytest.fixture
def x(s):
    return op2.Dat(s, None, np.float64, "x")


@pytest.fixture
def y(s):
    return op2.Dat(s, np.arange(1, NELEMS + 1), np.float64, "y")


@pytest.fixture
def yi(s):
    return op2.Dat(s, np.arange(1, NELEMS + 1), np.int64, "y")


@pytest.fixture
def x2():
    return op2.Dat(op2.Set(NELEMS, (1, 2)), np.zeros(2 * NELEMS), np.float64, "x")


@pytest.fixture
def y2():
    return op2.Dat(op2.Set(NELEMS, (2, 1)), np.zeros(2 * NELEMS), np.float64, "y")


def _set(d, v):
    d.data[:] = v


class TestLinAlgOp:

    def test_ops(self, backend, x, y):
        _set(x, 2 * y.data_ro)
        assert all((x + y).data_ro == 3 * y.data_ro)
        assert all((x - y).data_ro == y.data_ro)
        assert all((x * y).data_ro == 2 * y.data_ro * y.data_ro)
        assert all((x / y).data_ro == 2.0)

    @pytest.mark.parametrize('op', ['__add__', '__sub__', '__mul__', '__truediv__'])
    def test_shape_mismatch(self, backend, x2, y2, op):
        with pytest.raises(ValueError):
            getattr(x2, op)(y2)

    def test_scalar(self, backend, x, y):
        _set(x, y.data_ro + 1.0)
        assert all(x.data_ro == (y + 1.0).data_ro)
        _set(x, y.data_ro - 1.0)
        assert all(x.data_ro == (y - 1.0).data_ro)
        _set(x, 2 * y.data_ro)
        assert all(x.data_ro == (y * 2.0).data_ro)
        assert all((x / 2.0).data_ro == y.data_ro)

    @pytest.mark.parametrize('op', ['__add__', '__sub__', '__mul__', '__truediv__'])
    def test_result_dtype(self, backend, y, yi, op):
        """The result has the left operand's dtype."""
        assert getattr(y, op)(yi).data_ro.dtype == np.float64
        assert getattr(yi, op)(y).data_ro.dtype == np.int64


class TestLinAlgIop:

    def test_iops(self, backend, x, y):
        _set(x, 2 * y.data_ro)
        x += y
        assert all(x.data_ro == 3 * y.data_ro)
        _set(x, 2 * y.data_ro)
        x -= y
        assert all(x.data_ro == y.data_ro)
        _set(x, 2 * y.data_ro)
        x *= y
        assert all(x.data_ro == 2 * y.data_ro * y.data_ro)
        _set(x, 2 * y.data_ro)
        x /= y
        assert all(x.data_ro == 2.0)

    @pytest.mark.parametrize('op', ['__iadd__', '__isub__', '__imul__', '__itruediv__'])
    def test_shape_mismatch(self, backend, x2, y2, op):
        with pytest.raises(ValueError):
            getattr(x2, op)(y2)

    def test_scalar(self, backend, x, y):
        _set(x, y.data_ro + 1.0)
        y += 1.0
        assert all(x.data_ro == y.data_ro)
        _set(x, y.data_ro - 1.0)
        y -= 1.0
        assert all(x.data_ro == y.data_ro)
        _set(x, 2 * y.data_ro)
        y *= 2.0
        assert all(x.data_ro == y.data_ro)
        _set(x, 2 * y.data_ro)
        x /= 2.0
        assert all(x.data_ro == y.data_ro)

    @pytest.mark.parametrize('op', ['__iadd__', '__isub__', '__imul__', '__itruediv__'])
    def test_dtype_kept(self, backend, y, yi, op):
        a = getattr(y, op)(yi)
        assert a is y and y.data_ro.dtype == np.float64
        b = getattr(yi, op)(y)
        assert b is yi and yi.data_ro.dtype == np.int64


class TestLinAlgScalar:

    def test_norm(self, backend):
        n = op2.Dat(op2.Set(2), [3, 4], np.float64, "n")
        assert abs(n.norm - 5) < 1e-12

    def test_ops_after_loop(self, backend, s, y):
        """Arithmetic sees what a par_loop wrote on the device."""
        k = op2.Kernel("void k(double *v) { *v = 2.0 * *v; }", 'k')
        op2.par_loop(k, s, y(op2.IdentityMap, op2.RW))
        assert all((y + 1.0).data_ro.ravel() == 2.0 * np.arange(1, NELEMS + 1) + 1.0)
        assert abs(y.norm - 2.0 * np.sqrt((np.arange(1, NELEMS + 1) ** 2).sum())) < 1e-12
