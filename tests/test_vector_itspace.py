"""Indirect READ, WRITE and INC of 1- and 2-component Dats through a map of
arity 1, passed as a vector map (``double *v[1]``) or as an
iteration-space map (``map[op2.i[0]]`` with a trailing ``int i``), as the
reference's test/unit/test_vector_map.py and test_iteration_space_dats.py
state them (:95-250): 4096 nodes, two per element, so every element is
named twice; INC sums both nodes' values into it, WRITE leaves the value
both write, READ gives each node its element's value."""

import numpy as np
import pytest

from pyop2_amd import op2

NNODES = 4096
NELE = NNODES // 2


@pytest.fixture
def node():
    return op2.Set(NNODES, 1, 'node')


@pytest.fixture
def node2():
    return op2.Set(NNODES, 2, 'node2')


@pytest.fixture
def ele():
    return op2.Set(NELE, 1, 'ele')


@pytest.fixture
def ele2():
    return op2.Set(NELE, 2, 'ele2')


@pytest.fixture
def node2ele(node, ele):
    return op2.Map(node, ele, 1, np.arange(NNODES) // 2, 'node2ele')


@pytest.fixture
def node2ele2(node2, ele2):
    return op2.Map(node2, ele2, 1, np.arange(NNODES) // 2, 'node2ele2')


def _arg(dat, m, form, access):
    return dat(m, access) if form == 'vector' else dat(m[op2.i[0]], access)


def _iter(s, m, form):
    return s if form == 'vector' else s(m.dim)


def _kernel(body_vec, body_its, args, form):
    if form == 'vector':
        return op2.Kernel("void k(%s) { %s }" % (args % {'v': '*vd[1]'}, body_vec), 'k')
    return op2.Kernel("void k(%s, int i) { %s }" % (args % {'v': '*vd'}, body_its), 'k')


FORMS = ['vector', 'itspace']


@pytest.mark.parametrize('form', FORMS)
def test_read_1d(backend, node, ele, node2ele, form):
    d = op2.Dat(node, np.zeros(NNODES), dtype=np.int32)
    vd = op2.Dat(ele, np.arange(NELE), dtype=np.int32)
    k = _kernel("*d = vd[0][0];", "*d = vd[0];", "int *d, int %(v)s", form)
    op2.par_loop(k, _iter(node, node2ele, form), d(op2.IdentityMap, op2.WRITE),
                 _arg(vd, node2ele, form, op2.READ))
    assert all(d.data_ro[::2] == vd.data_ro)
    assert all(d.data_ro[1::2] == vd.data_ro)


@pytest.mark.parametrize('form', FORMS)
def test_write_1d(backend, node, ele, node2ele, form):
    vd = op2.Dat(ele, np.zeros(NELE), dtype=np.int32)
    k = _kernel("vd[0][0] = 2;", "vd[0] = 2;", "int %(v)s", form)
    op2.par_loop(k, _iter(node, node2ele, form), _arg(vd, node2ele, form, op2.WRITE))
    assert all(vd.data_ro == 2)


@pytest.mark.parametrize('form', FORMS)
def test_inc_1d(backend, node, ele, node2ele, form):
    d = op2.Dat(node, np.arange(NNODES), dtype=np.int32)
    vd = op2.Dat(ele, np.full(NELE, 3), dtype=np.int32)
    k = _kernel("vd[0][0] += *d;", "vd[0] += *d;", "int *d, int %(v)s", form)
    op2.par_loop(k, _iter(node, node2ele, form), d(op2.IdentityMap, op2.READ),
                 _arg(vd, node2ele, form, op2.INC))
    expected = 3 + np.arange(0, NNODES, 2) + np.arange(1, NNODES, 2)
    assert all(vd.data_ro.ravel() == expected)


@pytest.mark.parametrize('form', FORMS)
def test_read_2d(backend, node2, ele2, node2ele2, form):
    d = op2.Dat(node2, np.zeros(2 * NNODES), dtype=np.int32)
    vd = op2.Dat(ele2, np.arange(2 * NELE), dtype=np.int32)
    k = _kernel("d[0] = vd[0][0]; d[1] = vd[0][1];", "d[0] = vd[0]; d[1] = vd[1];",
                "int *d, int %(v)s", form)
    op2.par_loop(k, _iter(node2, node2ele2, form), d(op2.IdentityMap, op2.WRITE),
                 _arg(vd, node2ele2, form, op2.READ))
    for c in range(2):
        assert all(d.data_ro[::2, c] == vd.data_ro[:, c])
        assert all(d.data_ro[1::2, c] == vd.data_ro[:, c])


@pytest.mark.parametrize('form', FORMS)
def test_write_2d(backend, node2, ele2, node2ele2, form):
    vd = op2.Dat(ele2, np.zeros(2 * NELE), dtype=np.int32)
    k = _kernel("vd[0][0] = 2; vd[0][1] = 3;", "vd[0] = 2; vd[1] = 3;", "int %(v)s", form)
    op2.par_loop(k, _iter(node2, node2ele2, form), _arg(vd, node2ele2, form, op2.WRITE))
    assert all(vd.data_ro[:, 0] == 2)
    assert all(vd.data_ro[:, 1] == 3)


@pytest.mark.parametrize('form', FORMS)
def test_inc_2d(backend, node2, ele2, node2ele2, form):
    d = op2.Dat(node2, np.arange(2 * NNODES), dtype=np.int32)
    vd = op2.Dat(ele2, np.tile([3, 4], NELE), dtype=np.int32)
    k = _kernel("vd[0][0] += d[0]; vd[0][1] += d[1];", "vd[0] += d[0]; vd[1] += d[1];",
                "int *d, int %(v)s", form)
    op2.par_loop(k, _iter(node2, node2ele2, form), d(op2.IdentityMap, op2.READ),
                 _arg(vd, node2ele2, form, op2.INC))
    e0 = 3 + np.arange(0, 2 * NNODES, 4) + np.arange(2, 2 * NNODES, 4)
    e1 = 4 + np.arange(1, 2 * NNODES, 4) + np.arange(3, 2 * NNODES, 4)
    assert all(vd.data_ro[:, 0] == e0)
    assert all(vd.data_ro[:, 1] == e1)


def test_sum_nodes_to_edges_itspace(backend):
    """A 1-D chain of edges: each edge sums its two nodes through an
    iteration space over the edge's map (INC of a direct Dat)."""
    nedges = NNODES - 1
    nodes = op2.Set(NNODES, 1, "nodes")
    edges = op2.Set(nedges, 1, "edges")
    node_vals = op2.Dat(nodes, np.arange(NNODES, dtype=np.uint32), np.uint32, "node_vals")
    edge_vals = op2.Dat(edges, np.zeros(nedges, dtype=np.uint32), np.uint32, "edge_vals")
    e_map = np.array([(i, i + 1) for i in range(nedges)], dtype=np.uint32)
    edge2node = op2.Map(edges, nodes, 2, e_map, "edge2node")
    k = op2.Kernel("void kernel_sum(unsigned int* nodes, unsigned int *edge, int i) "
                   "{ *edge += nodes[0]; }", "kernel_sum")
    op2.par_loop(k, edges(edge2node.dim), node_vals(edge2node[op2.i[0]], op2.READ),
                 edge_vals(op2.IdentityMap, op2.INC))
    assert all(edge_vals.data_ro.ravel() == np.arange(1, nedges * 2 + 1, 2))
