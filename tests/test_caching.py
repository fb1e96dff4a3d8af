"""Caching semantics, as the reference's test/unit/test_caching.py states
them (:310-650): kernels are cached by code and name, compiled loops by
(kernel, argument signature, Const shapes), not by which Dats of a given
shape are passed, and sparsities by their map pairs, unordered.  Run on the
CPU oracle and (marked gpu) on hip, whose compiled-loop cache is the
code-object table the par_loop launcher fills.

The hip-only tests at the end check the plan caches the owner-computes
loops add: one plan per (iteration set, race columns), reused across Dats
of the same layout, and none for loops without a race."""

import numpy as np
import pytest

from pyop2_amd import op2
from pyop2_amd import backends as _backends

NELEMS = 8


def _code_cache():
    """The active backend's table of compiled loops."""
    if _backends.get_backend() == 'hip':
        from pyop2_amd import hip
        return hip._loop_cache
    from oracle import sequential
    return sequential._compiled


def _clear():
    if _backends.get_backend() == 'hip':
        from pyop2_amd import hip
        hip._call_cache.clear()
        hip._loop_cache.clear()
    else:
        from oracle import sequential
        sequential._compiled.clear()


def _backend_class(name):
    return _backends._BackendSelector._backend.__dict__[name]


@pytest.fixture
def iterset():
    return op2.Set(NELEMS, 1, "iterset")


@pytest.fixture
def indset():
    return op2.Set(NELEMS, 1, "indset")


@pytest.fixture
def indset2():
    return op2.Set(NELEMS, 2, "indset2")


@pytest.fixture
def x(indset):
    return op2.Dat(indset, np.arange(NELEMS), np.uint32, "x")


@pytest.fixture
def y(indset):
    return op2.Dat(indset, np.zeros(NELEMS), np.uint32, "y")


@pytest.fixture
def x2(indset2):
    return op2.Dat(indset2, np.arange(2 * NELEMS), np.uint32, "x2")


@pytest.fixture
def a(iterset):
    return op2.Dat(iterset, np.arange(NELEMS), np.uint32, "a")


@pytest.fixture
def b(iterset):
    return op2.Dat(iterset, np.arange(NELEMS), np.uint32, "b")


@pytest.fixture
def iter2ind1(iterset, indset):
    return op2.Map(iterset, indset, 1, np.arange(NELEMS), "iter2ind1")


@pytest.fixture
def iter2ind22(iterset, indset2):
    return op2.Map(iterset, indset2, 2, np.arange(2 * NELEMS) % NELEMS, "iter2ind22")


SWAP = """
void kernel_swap(unsigned int* x, unsigned int* y)
{
  unsigned int t;
  t = *x;
  *x = *y;
  *y = t;
}
"""


class TestGeneratedCodeCache:

    def test_same_args(self, backend, iterset, iter2ind1, x, a):
        _clear()
        cpy = "void kernel_cpy(unsigned int* dst, unsigned int* src) { *dst = *src; }"
        for _ in range(2):
            op2.par_loop(op2.Kernel(cpy, "kernel_cpy"), iterset,
                         a(op2.IdentityMap, op2.WRITE), x(iter2ind1[0], op2.READ))
            assert len(_code_cache()) == 1
        assert list(a.data_ro) == list(range(NELEMS))

    def test_diff_kernel(self, backend, iterset, iter2ind1, x, a):
        _clear()
        op2.par_loop(op2.Kernel("void kernel_cpy(unsigned int* dst, unsigned int* src) "
                                "{ *dst = *src; }", "kernel_cpy"), iterset,
                     a(op2.IdentityMap, op2.WRITE), x(iter2ind1[0], op2.READ))
        assert len(_code_cache()) == 1
        op2.par_loop(op2.Kernel("void kernel_cpy(unsigned int* DST, unsigned int* SRC) "
                                "{ *DST = *SRC; }", "kernel_cpy"), iterset,
                     a(op2.IdentityMap, op2.WRITE), x(iter2ind1[0], op2.READ))
        assert len(_code_cache()) == 2

    def test_invert_arg_similar_shape(self, backend, iterset, iter2ind1, x, y):
        _clear()
        op2.par_loop(op2.Kernel(SWAP, "kernel_swap"), iterset,
                     x(iter2ind1[0], op2.RW), y(iter2ind1[0], op2.RW))
        assert len(_code_cache()) == 1
        op2.par_loop(op2.Kernel(SWAP, "kernel_swap"), iterset,
                     y(iter2ind1[0], op2.RW), x(iter2ind1[0], op2.RW))
        assert len(_code_cache()) == 1
        # swapped twice: back where they started
        assert list(x.data_ro) == list(range(NELEMS))
        assert list(y.data_ro) == [0] * NELEMS

    def test_dloop_ignore_scalar(self, backend, iterset, a, b):
        _clear()
        op2.par_loop(op2.Kernel(SWAP, "kernel_swap"), iterset,
                     a(op2.IdentityMap, op2.RW), b(op2.IdentityMap, op2.RW))
        assert len(_code_cache()) == 1
        op2.par_loop(op2.Kernel(SWAP, "kernel_swap"), iterset,
                     b(op2.IdentityMap, op2.RW), a(op2.IdentityMap, op2.RW))
        assert len(_code_cache()) == 1

    def test_vector_map(self, backend, iterset, x2, iter2ind22):
        _clear()
        k = op2.Kernel("""
void kernel_swap(unsigned int* x[2])
{
  unsigned int t;
  t = x[0][0];
  x[0][0] = x[0][1];
  x[0][1] = t;
}""", "kernel_swap")
        for _ in range(2):
            op2.par_loop(k, iterset, x2(iter2ind22, op2.RW))
            assert len(_code_cache()) == 1

    def test_map_index_order_matters(self, backend, iterset, x2, iter2ind22):
        _clear()
        k = op2.Kernel("void k(unsigned int *x, unsigned int *y) {}", 'k')
        op2.par_loop(k, iterset, x2(iter2ind22[0], op2.INC), x2(iter2ind22[1], op2.INC))
        assert len(_code_cache()) == 1
        op2.par_loop(k, iterset, x2(iter2ind22[1], op2.INC), x2(iter2ind22[0], op2.INC))
        assert len(_code_cache()) == 2

    def test_same_iteration_space_works(self, backend, iterset, x2, iter2ind22):
        _clear()
        k = op2.Kernel("void k(unsigned int *x, int i) {}", 'k')
        for _ in range(2):
            op2.par_loop(k, iterset(2), x2(iter2ind22[op2.i[0]], op2.INC))
            assert len(_code_cache()) == 1

    def test_change_const_dim_matters(self, backend, iterset):
        d = op2.Dat(iterset, np.arange(NELEMS), np.uint32)
        _clear()
        k = op2.Kernel("void k(unsigned int *x) {}", 'k')
        c = op2.Const(1, 1, name='c', dtype=np.uint32)
        op2.par_loop(k, iterset, d(op2.IdentityMap, op2.WRITE))
        assert len(_code_cache()) == 1
        c.remove_from_namespace()
        c = op2.Const(2, (1, 1), name='c', dtype=np.uint32)
        op2.par_loop(k, iterset, d(op2.IdentityMap, op2.WRITE))
        assert len(_code_cache()) == 2
        c.remove_from_namespace()

    def test_change_const_data_doesnt_matter(self, backend, iterset):
        d = op2.Dat(iterset, np.zeros(NELEMS), np.uint32)
        _clear()
        k = op2.Kernel("void k(unsigned int *x) { *x = c; }", 'k')
        c = op2.Const(1, 1, name='c', dtype=np.uint32)
        op2.par_loop(k, iterset, d(op2.IdentityMap, op2.WRITE))
        assert len(_code_cache()) == 1
        assert list(d.data_ro) == [1] * NELEMS
        c.data = 2
        op2.par_loop(k, iterset, d(op2.IdentityMap, op2.WRITE))
        assert len(_code_cache()) == 1
        # the new value reaches the compiled loop without a recompile
        assert list(d.data_ro) == [2] * NELEMS
        c.remove_from_namespace()

    def test_change_dat_dtype_matters(self, backend, iterset):
        d = op2.Dat(iterset, np.arange(NELEMS), np.uint32)
        _clear()
        k = op2.Kernel("void k(void *x) {}", 'k')
        op2.par_loop(k, iterset, d(op2.IdentityMap, op2.WRITE))
        assert len(_code_cache()) == 1
        d = op2.Dat(iterset, np.arange(NELEMS), np.int32)
        op2.par_loop(k, iterset, d(op2.IdentityMap, op2.WRITE))
        assert len(_code_cache()) == 2

    def test_change_global_dtype_matters(self, backend, iterset):
        g = op2.Global(1, 0, dtype=np.uint32)
        _clear()
        k = op2.Kernel("void k(void *x) {}", 'k')
        op2.par_loop(k, iterset, g(op2.INC))
        assert len(_code_cache()) == 1
        g = op2.Global(1, 0, dtype=np.float64)
        op2.par_loop(k, iterset, g(op2.INC))
        assert len(_code_cache()) == 2


class TestKernelCache:

    def test_kernels_same_code_same_name(self, backend):
        cache = _backend_class('Kernel')._cache
        cache.clear()
        k1 = op2.Kernel("void k(void *x) {}", 'k')
        k2 = op2.Kernel("void k(void *x) {}", 'k')
        assert k1 is k2 and len(cache) == 1

    def test_kernels_same_code_differing_name(self, backend):
        cache = _backend_class('Kernel')._cache
        cache.clear()
        k1 = op2.Kernel("void k(void *x) {}", 'k')
        k2 = op2.Kernel("void k(void *x) {}", 'l')
        assert k1 is not k2 and len(cache) == 2

    def test_kernels_differing_code_same_name(self, backend):
        cache = _backend_class('Kernel')._cache
        cache.clear()
        k1 = op2.Kernel("void k(void *x) {}", 'k')
        k2 = op2.Kernel("void l(void *x) {}", 'k')
        assert k1 is not k2 and len(cache) == 2

    def test_kernels_differing_code_differing_name(self, backend):
        cache = _backend_class('Kernel')._cache
        cache.clear()
        k1 = op2.Kernel("void k(void *x) {}", 'k')
        k2 = op2.Kernel("void l(void *x) {}", 'l')
        assert k1 is not k2 and len(cache) == 2


class TestSparsityCache:

    @pytest.fixture
    def m1(self):
        return op2.Map(op2.Set(5), op2.Set(5), 1, [0, 1, 2, 3, 4])

    @pytest.fixture
    def m2(self, m1):
        return op2.Map(m1.iterset, m1.dataset, 1, [1, 2, 3, 4, 0])

    def test_sparsities_differing_maps_not_cached(self, backend, m1, m2):
        assert op2.Sparsity(m1) is not op2.Sparsity(m2)

    def test_sparsities_differing_map_pairs_not_cached(self, backend, m1, m2):
        assert op2.Sparsity((m1, m2)) is not op2.Sparsity((m2, m1))

    def test_sparsities_differing_map_tuples_not_cached(self, backend, m1, m2):
        assert op2.Sparsity(((m1, m1), (m2, m2))) is not op2.Sparsity(((m2, m2), (m2, m2)))

    def test_sparsities_same_map_cached(self, backend, m1):
        assert op2.Sparsity(m1) is op2.Sparsity(m1)

    def test_sparsities_same_map_pair_cached(self, backend, m1):
        assert op2.Sparsity((m1, m1)) is op2.Sparsity((m1, m1))

    def test_sparsities_same_map_tuple_cached(self, backend, m1, m2):
        assert op2.Sparsity(((m1, m1), (m2, m2))) is op2.Sparsity(((m1, m1), (m2, m2)))

    def test_sparsities_different_ordered_map_tuple_cached(self, backend, m1, m2):
        assert op2.Sparsity(((m1, m1), (m2, m2))) is op2.Sparsity(((m2, m2), (m1, m1)))

    @pytest.mark.gpu
    def test_two_mats_on_same_sparsity_share_data(self, hip, m1):
        """Two Mats on one Sparsity share its device CSR (reference
        test_caching.py:642-650, device backends)."""
        sp = hip.Sparsity((m1, m1))
        mat1 = hip.Mat(sp, 'float64')
        mat2 = hip.Mat(sp, 'float64')
        assert mat1.sparsity is mat2.sparsity
        assert mat1.sparsity._dev_rowptr.data_ptr() == mat2.sparsity._dev_rowptr.data_ptr()
        assert mat1.sparsity._dev_colidx.data_ptr() == mat2.sparsity._dev_colidx.data_ptr()
        assert mat1._dev_vals.data_ptr() != mat2._dev_vals.data_ptr()


@pytest.mark.gpu
class TestOwnerPlanCache:
    """Owner-computes plans of the hip backend: keyed by iteration set,
    size and race columns (map and index), not by the Dats passed."""

    INC2 = "void k(unsigned int *v, unsigned int *a, unsigned int *b) { *a += *v; *b += 2 * *v; }"

    def test_same_columns_share_plan(self, hip, iterset, indset2, iter2ind22):
        from pyop2_amd import hip as H
        _clear()
        H._owner_cache.clear()
        v = hip.Dat(iterset, np.arange(NELEMS), np.uint32)
        p = hip.Dat(indset2, np.zeros(2 * NELEMS), np.uint32)
        q = hip.Dat(indset2, np.zeros(2 * NELEMS), np.uint32)
        k = hip.Kernel(self.INC2, 'k')
        hip.par_loop(k, iterset, v(hip.IdentityMap, hip.READ),
                     p(iter2ind22[0], hip.INC), p(iter2ind22[1], hip.INC))
        n = len([x for x in H._owner_cache.values() if x is not None])
        hip.par_loop(k, iterset, v(hip.IdentityMap, hip.READ),
                     q(iter2ind22[0], hip.INC), q(iter2ind22[1], hip.INC))
        assert len([x for x in H._owner_cache.values() if x is not None]) == n
        assert np.array_equal(p.data_ro, q.data_ro)
        # a different race column order is a different plan
        hip.par_loop(k, iterset, v(hip.IdentityMap, hip.READ),
                     q(iter2ind22[1], hip.INC), q(iter2ind22[0], hip.INC))
        assert len(H._owner_cache) > n

    def test_read_only_indirect_needs_no_plan(self, hip, iterset, iter2ind1, x, a):
        from pyop2_amd import hip as H
        _clear()
        H._owner_cache.clear()
        H._plan_cache.clear()
        cpy = "void kernel_cpy(unsigned int* dst, unsigned int* src) { *dst = *src; }"
        hip.par_loop(hip.Kernel(cpy, "kernel_cpy"), iterset,
                     a(hip.IdentityMap, hip.WRITE), x(iter2ind1[0], hip.READ))
        assert len(H._owner_cache) == 0 and len(H._plan_cache) == 0
