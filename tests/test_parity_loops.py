"""hip backend against the CPU oracle on small loops of every argument kind:
direct, indirect READ/WRITE/INC, vector maps, iteration spaces, Global
READ/INC/MIN/MAX (int32, uint32, float64), Mat INC/WRITE (scalar and
vector fields), sparsity, solve."""

import numpy as np
import pytest

from oracle import sequential as seq

pytestmark = pytest.mark.gpu


def mk(W, n_elem=1000, n_node=300, seed=0):
    rng = np.random.default_rng(seed)
    elems = W.Set(n_elem, 1, 'elems')
    nodes = W.Set(n_node, 1, 'nodes')
    vnodes = W.Set(n_node, 2, 'vnodes')
    m = rng.integers(0, n_node, size=(n_elem, 3)).astype(np.int32)
    return elems, nodes, vnodes, m


def test_indirect_inc_scalar(hip):
    rng = np.random.default_rng(1)
    code = "void k(double *x, double *a, double *b, double *c) { *a += x[0]; *b += 2.0 * x[0]; *c -= x[0] * x[0]; }"
    xv = rng.random(1000)
    out = {}
    for name, W in (('hip', hip), ('seq', seq)):
        elems, nodes, _, m = mk(W)
        emap = W.Map(elems, nodes, 3, m, 'em')
        x = W.Dat(elems, xv.copy(), np.float64)
        y = W.Dat(nodes, np.zeros(300), np.float64)
        W.par_loop(W.Kernel(code, 'k'), elems, x(W.IdentityMap, W.READ),
                   y(emap[0], W.INC), y(emap[1], W.INC), y(emap[2], W.INC))
        out[name] = y.data_ro.copy()
    np.testing.assert_allclose(out['hip'], out['seq'], rtol=1e-12, atol=1e-12)


def test_global_reductions(hip):
    rng = np.random.default_rng(2)
    xv = rng.random(5000)
    code = """void k(double *x, double *s, double *mn, double *mx, int *cnt) {
      *s += *x; if (*x < *mn) *mn = *x; if (*x > *mx) *mx = *x; *cnt += 1; }"""
    res = {}
    for name, W in (('hip', hip), ('seq', seq)):
        S = W.Set(5000)
        x = W.Dat(S, xv.copy(), np.float64)
        s = W.Global(1, 1.5, np.float64)
        mn = W.Global(1, 10.0, np.float64)
        mx = W.Global(1, -10.0, np.float64)
        c = W.Global(1, 7, np.int32)
        W.par_loop(W.Kernel(code, 'k'), S, x(W.IdentityMap, W.READ), s(W.INC), mn(W.MIN),
                   mx(W.MAX), c(W.INC))
        res[name] = (s.data[0], mn.data[0], mx.data[0], c.data[0])
    h, o = res['hip'], res['seq']
    assert abs(h[0] - o[0]) <= 1e-12 * abs(o[0])
    assert h[1:] == o[1:]


def test_mass_matrix_and_solve(hip):
    from pyop2_amd.workloads import fem, meshes
    mesh = meshes.TriMesh(16, seed=3)
    res = {}
    for name, W in (('hip', hip), ('seq', seq)):
        r = fem.mass2d(W, mesh, f=np.linspace(0, 1, mesh.nvert), solve_rtol=1e-12)
        res[name] = r
    assert np.array_equal(res['hip']['rowptr'], res['seq']['rowptr'])
    assert np.array_equal(res['hip']['colidx'], res['seq']['colidx'])
    np.testing.assert_allclose(res['hip']['mat'], res['seq']['mat'], rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(res['hip']['b'], res['seq']['b'], rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(res['hip']['x'], res['seq']['x'], rtol=1e-10, atol=1e-12)
