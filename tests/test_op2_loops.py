"""par_loop semantics on every backend: direct and indirect loops, Global
reductions, vector maps and iteration-space Dats.

Assertions are the closed-form known answers of the reference's unit tests
(test/unit/test_direct_loop.py:79-213, test_indirect_loop.py:70-145,
test_global_reduction.py:146-399, test_vector_map.py:95-211,
test_iteration_space_dats.py:95-211), rewritten for Python 3.  They pin the
CPU oracle (run with -m "not gpu") and check the hip backend (-m gpu).
"""

import numpy as np
import pytest

from pyop2_amd import op2

NELEMS = 4096  # several blocks, and more than one thread per element


def shuffled(n):
    """The reference tests' fixed shuffle: random.shuffle with a constant
    random() of 0.02041724, i.e. for i = n-1 .. 1 swap(i, int(0.02041724*(i+1)))."""
    a = np.arange(n, dtype=np.uint32)
    for i in range(n - 1, 0, -1):
        j = int(0.02041724 * (i + 1))
        a[i], a[j] = a[j], a[i]
    return a


@pytest.fixture
def elems(backend):
    return op2.Set(NELEMS, 1, "elems")


@pytest.fixture
def x(elems):
    return op2.Dat(elems, np.arange(NELEMS, dtype=np.uint32), np.uint32, "x")


class TestDirectLoop:

    def test_wo(self, elems, x):
        op2.par_loop(op2.Kernel("void kernel_wo(unsigned int* x) { *x = 42; }", "kernel_wo"),
                     elems, x(op2.IdentityMap, op2.WRITE))
        assert (x.data == 42).all()

    def test_rw(self, elems, x):
        op2.par_loop(op2.Kernel("void kernel_rw(unsigned int* x) { (*x) = (*x) + 1; }",
                                "kernel_rw"), elems, x(op2.IdentityMap, op2.RW))
        assert x.data.sum() == NELEMS * (NELEMS + 1) // 2

    def test_global_inc(self, elems, x):
        g = op2.Global(1, 0, np.uint32, "g")
        k = "void k(unsigned int* x, unsigned int* inc) { (*x) = (*x) + 1; (*inc) += (*x); }"
        op2.par_loop(op2.Kernel(k, "k"), elems, x(op2.IdentityMap, op2.RW), g(op2.INC))
        assert g.data[0] == NELEMS * (NELEMS + 1) // 2

    def test_global_inc_init_not_zero(self, elems):
        g = op2.Global(1, 0, np.uint32, "g")
        g.data[0] = 10
        op2.par_loop(op2.Kernel("void k(unsigned int* inc) { (*inc) += 1; }", 'k'), elems,
                     g(op2.INC))
        assert g.data[0] == NELEMS + 10

    def test_global_max_min(self, elems, x):
        kmax = op2.Kernel("void kmax(unsigned int *x, unsigned int *g) { if (*g < *x) *g = *x; }",
                          'kmax')
        kmin = op2.Kernel("void k2(unsigned int *x, unsigned int *g) { if (*g > *x) *g = *x; }",
                          'k2')
        g = op2.Global(1, 0, np.uint32, "g")
        op2.par_loop(kmax, elems, x(op2.IdentityMap, op2.READ), g(op2.MAX))
        assert g.data[0] == NELEMS - 1
        g.data[0] = NELEMS * 2
        op2.par_loop(kmax, elems, x(op2.IdentityMap, op2.READ), g(op2.MAX))
        assert g.data[0] == NELEMS * 2
        g.data[0] = 1000
        op2.par_loop(kmin, elems, x(op2.IdentityMap, op2.READ), g(op2.MIN))
        assert g.data[0] == 0
        g.data[0] = 10
        x.data[:] = 11
        op2.par_loop(kmin, elems, x(op2.IdentityMap, op2.READ), g(op2.MIN))
        assert g.data[0] == 10

    def test_global_read(self, elems, x):
        h = op2.Global(1, 1, np.uint32, "h")
        op2.par_loop(op2.Kernel("void kr(unsigned int* x, unsigned int* h) { (*x) += (*h); }",
                                "kr"), elems, x(op2.IdentityMap, op2.RW), h(op2.READ))
        assert x.data.sum() == NELEMS * (NELEMS + 1) // 2

    def test_2d_dat(self, elems):
        s2 = op2.Set(NELEMS, 2, "elems2")
        y = op2.Dat(s2, np.zeros((NELEMS, 2), dtype=np.uint32), np.uint32, "y")
        op2.par_loop(op2.Kernel("void k2d(unsigned int* x) { x[0] = 42; x[1] = 43; }", "k2d"),
                     elems, y(op2.IdentityMap, op2.WRITE))
        assert (y.data == [42, 43]).all()

    def test_2d_dat_soa(self, elems):
        s2 = op2.Set(NELEMS, 2, "elems2")
        soa = op2.Dat(s2, np.zeros((NELEMS, 2), dtype=np.uint32), np.uint32, "soa", soa=True)
        k = "void ksoa(unsigned int * x) { OP2_STRIDE(x, 0) = 42; OP2_STRIDE(x, 1) = 43; }"
        op2.par_loop(op2.Kernel(k, "ksoa"), elems, soa(op2.IdentityMap, op2.WRITE))
        assert (soa.data[:, 0] == 42).all() and (soa.data[:, 1] == 43).all()
        assert soa.data.flags['C_CONTIGUOUS']

    def test_parloop_sets_ro_flag(self, elems, x):
        x_data = x.data
        op2.par_loop(op2.Kernel("void kro(unsigned int *x) { *x = 1; }", 'kro'), elems,
                     x(op2.IdentityMap, op2.WRITE))
        with pytest.raises((RuntimeError, ValueError)):
            x_data[0] = 1

    def test_host_write_works(self, elems, x):
        g = op2.Global(1, 0, np.uint32, "g")
        k = op2.Kernel("void kh(unsigned int *x, unsigned int *g) { *g += *x; }", 'kh')
        x.data[:] = 1
        g.data[:] = 0
        op2.par_loop(k, elems, x(op2.IdentityMap, op2.READ), g(op2.INC))
        assert g.data[0] == NELEMS
        x.data[:] = 2
        g.data[:] = 0
        op2.par_loop(k, elems, x(op2.IdentityMap, op2.READ), g(op2.INC))
        assert g.data[0] == 2 * NELEMS

    def test_zero(self, elems, x):
        x.data[:] = 10
        assert (x.data == 10).all()
        x.zero()
        assert (x.data == 0).all()
        s2 = op2.Set(NELEMS, 2, "elems2")
        y = op2.Dat(s2, np.full((NELEMS, 2), 10, dtype=np.uint32), np.uint32, "y")
        y.zero()
        assert (y.data == 0).all()


class TestIndirectLoop:

    @pytest.fixture
    def iterset2indset(self, elems):
        indset = op2.Set(NELEMS, 1, "indset")
        return op2.Map(elems, indset, 1, shuffled(NELEMS), "iterset2indset")

    @pytest.fixture
    def xi(self, iterset2indset):
        return op2.Dat(iterset2indset.dataset, np.arange(NELEMS, dtype=np.uint32), np.uint32)

    def test_onecolor_wo(self, elems, xi, iterset2indset):
        op2.par_loop(op2.Kernel("void kwo(unsigned int* x) { *x = 42; }", "kwo"), elems,
                     xi(iterset2indset[0], op2.WRITE))
        assert (xi.data == 42).all()

    def test_onecolor_rw(self, elems, xi, iterset2indset):
        op2.par_loop(op2.Kernel("void krw(unsigned int* x) { (*x) = (*x) + 1; }", "krw"), elems,
                     xi(iterset2indset[0], op2.RW))
        assert xi.data.sum() == NELEMS * (NELEMS + 1) // 2

    def test_indirect_inc(self, elems):
        unitset = op2.Set(1, 1, "unitset")
        u = op2.Dat(unitset, np.array([0], dtype=np.uint32), np.uint32, "u")
        m = op2.Map(elems, unitset, 1, np.zeros(NELEMS, dtype=np.uint32), "iterset2unitset")
        op2.par_loop(op2.Kernel("void kinc(unsigned int* x) { (*x) = (*x) + 1; }", "kinc"),
                     elems, u(m[0], op2.INC))
        assert u.data[0] == NELEMS

    def test_global_read(self, elems, xi, iterset2indset):
        g = op2.Global(1, 2, np.uint32, "g")
        op2.par_loop(op2.Kernel("void kg(unsigned int* x, unsigned int* g) { (*x) /= (*g); }",
                                "kg"), elems, xi(iterset2indset[0], op2.RW), g(op2.READ))
        assert xi.data.sum() == sum(v // 2 for v in range(NELEMS))

    def test_global_inc(self, elems, xi, iterset2indset):
        g = op2.Global(1, 0, np.uint32, "g")
        k = "void kgi(unsigned int *x, unsigned int *inc) { (*x) = (*x) + 1; (*inc) += (*x); }"
        op2.par_loop(op2.Kernel(k, "kgi"), elems, xi(iterset2indset[0], op2.RW), g(op2.INC))
        assert xi.data.sum() == NELEMS * (NELEMS + 1) // 2
        assert g.data[0] == NELEMS * (NELEMS + 1) // 2

    def test_2d_map(self, backend):
        nedges = NELEMS - 1
        nodes = op2.Set(NELEMS, 1, "nodes")
        edges = op2.Set(nedges, 1, "edges")
        node_vals = op2.Dat(nodes, np.arange(NELEMS, dtype=np.uint32), np.uint32, "node_vals")
        edge_vals = op2.Dat(edges, np.zeros(nedges, dtype=np.uint32), np.uint32, "edge_vals")
        e_map = np.array([(i, i + 1) for i in range(nedges)], dtype=np.uint32)
        edge2node = op2.Map(edges, nodes, 2, e_map, "edge2node")
        k = """void ksum(unsigned int *nodes1, unsigned int *nodes2, unsigned int *edge)
               { *edge = *nodes1 + *nodes2; }"""
        op2.par_loop(op2.Kernel(k, "ksum"), edges, node_vals(edge2node[0], op2.READ),
                     node_vals(edge2node[1], op2.READ), edge_vals(op2.IdentityMap, op2.WRITE))
        assert (np.arange(1, nedges * 2 + 1, 2).reshape(nedges, 1) == edge_vals.data).all()


class TestGlobalReductions:

    @pytest.mark.parametrize('dtype', [np.uint32, np.int32, np.float32, np.float64])
    def test_inc(self, elems, dtype):
        d = op2.Dat(elems, np.arange(NELEMS) + 1, dtype=dtype)
        ct = d.ctype
        g = op2.Global(1, 0, dtype, "g")
        op2.par_loop(op2.Kernel("void kinc(%s *x, %s *g) { *g += *x; }" % (ct, ct), "kinc"),
                     elems, d(op2.IdentityMap, op2.READ), g(op2.INC))
        assert g.data[0] == NELEMS * (NELEMS + 1) // 2

    @pytest.mark.parametrize('dtype', [np.uint32, np.float64])
    def test_min_max_2d(self, backend, dtype):
        s2 = op2.Set(NELEMS, 2, 'set2')
        d2 = op2.Dat(s2, np.arange(2 * NELEMS) + 1, dtype=dtype)
        ct = d2.ctype
        kmin = """void kmin2(%s *x, %s *g) {
                  if (x[0] < g[0]) g[0] = x[0]; if (x[1] < g[1]) g[1] = x[1]; }""" % (ct, ct)
        kmax = """void kmax2(%s *x, %s *g) {
                  if (x[0] > g[0]) g[0] = x[0]; if (x[1] > g[1]) g[1] = x[1]; }""" % (ct, ct)
        gmin = op2.Global(2, [1000 * NELEMS, 1000 * NELEMS], dtype, "gmin")
        gmax = op2.Global(2, [0, 0], dtype, "gmax")
        op2.par_loop(op2.Kernel(kmin, 'kmin2'), s2, d2(op2.IdentityMap, op2.READ), gmin(op2.MIN))
        op2.par_loop(op2.Kernel(kmax, 'kmax2'), s2, d2(op2.IdentityMap, op2.READ), gmax(op2.MAX))
        assert list(gmin.data) == [1, 2]
        assert list(gmax.data) == [2 * NELEMS - 1, 2 * NELEMS]

    def test_repeated_inc(self, elems):
        d = op2.Dat(elems, np.ones(NELEMS), dtype=np.float64)
        g = op2.Global(1, 0.0, np.float64, "g")
        k = op2.Kernel("void krep(double *x, double *g) { *g += *x; }", "krep")
        for r in range(1, 4):
            op2.par_loop(k, elems, d(op2.IdentityMap, op2.READ), g(op2.INC))
            assert g.data[0] == r * NELEMS

    def test_read_into_dat(self, elems):
        d = op2.Dat(elems, np.zeros(NELEMS), dtype=np.uint32)
        g = op2.Global(2, [3, 4], np.uint32, "g")
        op2.par_loop(op2.Kernel("void kw(unsigned int *x, unsigned int *g) { *x = g[0] + g[1]; }",
                                "kw"), elems, d(op2.IdentityMap, op2.WRITE), g(op2.READ))
        assert (d.data == 7).all()


NNODES = 4096
NELE = NNODES // 2


class TestVectorMap:

    @pytest.fixture
    def sets(self, backend):
        node = op2.Set(NNODES, 1, 'node')
        ele = op2.Set(NELE, 1, 'ele')
        node2 = op2.Set(NNODES, 2, 'node2')
        ele2 = op2.Set(NELE, 2, 'ele2')
        vals = np.arange(NNODES) // 2
        return (node, ele, node2, ele2, op2.Map(node, ele, 1, vals, 'node2ele'),
                op2.Map(node2, ele2, 1, vals, 'node2ele2'))

    def test_sum_nodes_to_edges(self, backend):
        nedges = NNODES - 1
        nodes = op2.Set(NNODES, 1, "nodes")
        edges = op2.Set(nedges, 1, "edges")
        node_vals = op2.Dat(nodes, np.arange(NNODES, dtype=np.uint32), np.uint32)
        edge_vals = op2.Dat(edges, np.zeros(nedges, dtype=np.uint32), np.uint32)
        edge2node = op2.Map(edges, nodes, 2, np.array([(i, i + 1) for i in range(nedges)]))
        k = "void ks(unsigned int* nodes[1], unsigned int *edge) { *edge = nodes[0][0] + nodes[1][0]; }"
        op2.par_loop(op2.Kernel(k, "ks"), edges, node_vals(edge2node, op2.READ),
                     edge_vals(op2.IdentityMap, op2.WRITE))
        assert (np.arange(1, nedges * 2 + 1, 2).reshape(nedges, 1) == edge_vals.data).all()

    def test_read_write_inc_1d(self, sets):
        node, ele, _, _, n2e, _ = sets
        d1 = op2.Dat(node, np.zeros(NNODES), dtype=np.int32)
        vd1 = op2.Dat(ele, np.zeros(NELE), dtype=np.int32)
        vd1.data[:] = np.arange(NELE).reshape(NELE, 1)
        op2.par_loop(op2.Kernel("void kr(int *d, int *vd[1]) { *d = vd[0][0]; }", 'kr'), node,
                     d1(op2.IdentityMap, op2.WRITE), vd1(n2e, op2.READ))
        assert (d1.data[::2] == vd1.data).all() and (d1.data[1::2] == vd1.data).all()
        op2.par_loop(op2.Kernel("void kw(int *vd[1]) { vd[0][0] = 2; }", 'kw'), node,
                     vd1(n2e, op2.WRITE))
        assert (vd1.data == 2).all()
        vd1.data[:] = 3
        d1.data[:] = np.arange(NNODES).reshape(d1.data.shape)
        op2.par_loop(op2.Kernel("void ki(int *d, int *vd[1]) { vd[0][0] += *d; }", 'ki'), node,
                     d1(op2.IdentityMap, op2.READ), vd1(n2e, op2.INC))
        expected = 3 + np.arange(0, NNODES, 2) + np.arange(1, NNODES, 2)
        assert (vd1.data.reshape(-1) == expected).all()

    def test_inc_2d(self, sets):
        _, _, node2, ele2, _, n2e2 = sets
        d2 = op2.Dat(node2, np.arange(2 * NNODES).reshape(NNODES, 2), dtype=np.int32)
        vd2 = op2.Dat(ele2, np.zeros((NELE, 2)), dtype=np.int32)
        vd2.data[:, 0] = 3
        vd2.data[:, 1] = 4
        k = "void ki2(int *d, int *vd[2]) { vd[0][0] += d[0]; vd[0][1] += d[1]; }"
        op2.par_loop(op2.Kernel(k, 'ki2'), node2, d2(op2.IdentityMap, op2.READ),
                     vd2(n2e2, op2.INC))
        assert (vd2.data[:, 0] == 3 + np.arange(0, 2 * NNODES, 4) + np.arange(2, 2 * NNODES, 4)).all()
        assert (vd2.data[:, 1] == 4 + np.arange(1, 2 * NNODES, 4) + np.arange(3, 2 * NNODES, 4)).all()


class TestIterationSpaceDats:

    def test_sum_nodes_to_edges(self, backend):
        nedges = NNODES - 1
        nodes = op2.Set(NNODES, 1, "nodes")
        edges = op2.Set(nedges, 1, "edges")
        node_vals = op2.Dat(nodes, np.arange(NNODES, dtype=np.uint32), np.uint32)
        edge_vals = op2.Dat(edges, np.zeros(nedges, dtype=np.uint32), np.uint32)
        edge2node = op2.Map(edges, nodes, 2, np.array([(i, i + 1) for i in range(nedges)]))
        k = "void kis(unsigned int* nodes, unsigned int *edge, int i) { *edge += *nodes; }"
        op2.par_loop(op2.Kernel(k, "kis"), edges(edge2node.dim),
                     node_vals(edge2node[op2.i[0]], op2.READ),
                     edge_vals(op2.IdentityMap, op2.INC))
        assert (np.arange(1, nedges * 2 + 1, 2).reshape(nedges, 1) == edge_vals.data).all()

    def test_inc_itspace(self, backend):
        node = op2.Set(NNODES, 1, 'node')
        ele = op2.Set(NELE, 1, 'ele')
        e2n = op2.Map(ele, node, 2, np.arange(NNODES).reshape(NELE, 2), 'e2n')
        d = op2.Dat(node, np.arange(NNODES), dtype=np.int32)
        vd = op2.Dat(ele, np.zeros(NELE), dtype=np.int32)
        k = "void kii(int *d, int *vd, int i) { vd[0] += d[0]; }"
        op2.par_loop(op2.Kernel(k, 'kii'), ele(2), d(e2n[op2.i[0]], op2.READ),
                     vd(op2.IdentityMap, op2.INC))
        assert (vd.data.reshape(-1) == np.arange(0, NNODES, 2) + np.arange(1, NNODES, 2)).all()
        # increments through the iteration space into the node set
        k2 = "void kio(int *vd, int *d, int i) { d[0] += vd[0] * (i + 1); }"
        d.data[:] = 0
        op2.par_loop(op2.Kernel(k2, 'kio'), ele(2), vd(op2.IdentityMap, op2.READ),
                     d(e2n[op2.i[0]], op2.INC))
        exp = np.zeros(NNODES, dtype=np.int64)
        exp[0::2] = vd.data.reshape(-1)
        exp[1::2] = 2 * vd.data.reshape(-1)
        assert (d.data.reshape(-1) == exp).all()


class TestConstants:

    def test_1d_and_2d_const(self, elems):
        c1 = op2.Const(1, 100, 'myconstant', dtype=np.int32)
        c2 = op2.Const(2, [100, 200], 'myconstant2', dtype=np.int32)
        try:
            d = op2.Dat(elems, np.zeros(NELEMS, dtype=np.int32), np.int32)
            op2.par_loop(op2.Kernel("void kc(int *x) { *x = myconstant + myconstant2[1]; }",
                                    'kc'), elems, d(op2.IdentityMap, op2.WRITE))
            assert (d.data == 300).all()
            # changing a Const's value takes effect without new code
            c1.data = 7
            op2.par_loop(op2.Kernel("void kc(int *x) { *x = myconstant + myconstant2[1]; }",
                                    'kc'), elems, d(op2.IdentityMap, op2.WRITE))
            assert (d.data == 207).all()
        finally:
            c1.remove_from_namespace()
            c2.remove_from_namespace()

    def test_unique_name(self, backend):
        c = op2.Const(1, 1, 'dupname', dtype=np.int32)
        try:
            with pytest.raises(op2.Const.NonUniqueNameError if hasattr(op2.Const, 'NonUniqueNameError')
                               else ValueError):
                op2.Const(1, 2, 'dupname', dtype=np.int32)
        finally:
            c.remove_from_namespace()
