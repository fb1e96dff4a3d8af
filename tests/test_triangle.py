"""Triangle mesh files (.node/.ele) read as the reference's demos read them
(demo/triangle_reader.py:39-89), then the mass2d_triangle demo
(demo/mass2d_triangle.py:88-120): P1 mass matrix and right-hand side
M f, solved for x = f."""

import numpy as np
import pytest

from pyop2_amd import op2
from pyop2_amd.workloads import fem, meshes


def write_triangle(prefix, coords, elem_node, shuffle_seed=None):
    """A .node/.ele pair with 1-based ids, comments, and (optionally) the
    lines in a shuffled order, which the ids put right."""
    rng = np.random.default_rng(shuffle_seed)
    order_n = rng.permutation(len(coords)) if shuffle_seed is not None else range(len(coords))
    with open(prefix + '.node', 'w') as h:
        h.write('%d 2 0 0\n# node x y\n' % len(coords))
        for i in order_n:
            h.write('%d %.17g %.17g\n' % (i + 1, coords[i, 0], coords[i, 1]))
    order_e = rng.permutation(len(elem_node)) if shuffle_seed is not None else range(len(elem_node))
    with open(prefix + '.ele', 'w') as h:
        h.write('%d 3 0\n' % len(elem_node))
        for e in order_e:
            a, b, c = elem_node[e] + 1
            h.write('%d %d %d %d  # triangle\n' % (e + 1, a, b, c))


def test_read_triangle(backend, tmp_path):
    m = meshes.TriMesh(9, seed=2)
    prefix = str(tmp_path / 'mesh')
    write_triangle(prefix, m.coords, m.elem_node, shuffle_seed=5)
    nodes, vnodes, coords, elements, elem_node, elem_vnode = meshes.read_triangle(op2, prefix)
    assert nodes.size == m.nvert and vnodes.dim == (2,) and elements.size == m.nelem
    assert np.array_equal(coords.data_ro, m.coords)
    assert np.array_equal(elem_node.values, m.elem_node)
    assert np.array_equal(elem_vnode.values, m.elem_node)


def test_read_triangle_rejects_quadratic(backend, tmp_path):
    prefix = str(tmp_path / 'q')
    with open(prefix + '.node', 'w') as h:
        h.write('3 2 0 0\n1 0 0\n2 1 0\n3 0 1\n')
    with open(prefix + '.ele', 'w') as h:
        h.write('1 6 0\n1 1 2 3 1 2 3\n')
    with pytest.raises(ValueError):
        meshes.read_triangle(op2, prefix)


def test_mass2d_triangle_demo(backend, tmp_path):
    m = meshes.TriMesh(16, seed=4)
    prefix = str(tmp_path / 'sq')
    write_triangle(prefix, m.coords, m.elem_node)
    nodes, vnodes, coords, elements, elem_node, elem_vnode = meshes.read_triangle(op2, prefix)
    fv = np.random.default_rng(4).random(m.nvert)
    mat = op2.Mat(op2.Sparsity((elem_node, elem_node), "sparsity"), np.float64, "mat")
    f = op2.Dat(nodes, fv.copy(), np.float64, "f")
    b = op2.Dat(nodes, np.zeros(m.nvert), np.float64, "b")
    x = op2.Dat(nodes, np.zeros(m.nvert), np.float64, "x")
    op2.par_loop(op2.Kernel(fem.MASS_TRI, "mass_tri"), elements(3, 3),
                 mat((elem_node[op2.i[0]], elem_node[op2.i[1]]), op2.INC),
                 coords(elem_vnode, op2.READ))
    op2.par_loop(op2.Kernel(fem.RHS_TRI, "rhs_tri"), elements(3),
                 b(elem_node[op2.i[0]], op2.INC), coords(elem_vnode, op2.READ),
                 f(elem_node, op2.READ))
    op2.Solver(relative_tolerance=1e-12).solve(mat, x, b)
    assert np.abs(x.data_ro.ravel() - fv).max() <= 1e-8
