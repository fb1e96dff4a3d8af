"""Edge cases of the hip loop templates against the CPU oracle (gpu):
empty and one-element iteration sets, targets no element names, INC through
vector maps with MIN/MAX Globals in the same owner loop (a redundantly run
element counts in its primary block only), two INC Dats (the coloured
fallback), Mat WRITE (insert), and loops whose maps name the same target
twice in one element."""

import numpy as np
import pytest

from oracle import sequential as seq

pytestmark = pytest.mark.gpu

INC3 = """void k(double *x, double *a[3], double *s, double *mn, double *mx) {
  for (int j = 0; j < 3; ++j) { a[j][0] += x[0] * (j + 1); a[j][1] -= x[0]; }
  *s += x[0]; if (x[0] < *mn) *mn = x[0]; if (x[0] > *mx) *mx = x[0]; }"""


def _inc3(W, n, nt, m, xv):
    E = W.Set(n)
    T = W.Set(nt, 2)
    mp = W.Map(E, T, 3, m)
    x = W.Dat(E, xv.copy(), np.float64)
    a = W.Dat(T, np.arange(2 * nt, dtype=np.float64).reshape(nt, 2) * 0.25, np.float64)
    s = W.Global(1, 0.5, np.float64)
    mn = W.Global(1, 2.0, np.float64)
    mx = W.Global(1, -2.0, np.float64)
    W.par_loop(W.Kernel(INC3, 'k'), E, x(W.IdentityMap, W.READ), a(mp, W.INC), s(W.INC),
               mn(W.MIN), mx(W.MAX))
    return a.data_ro.copy(), s.data[0], mn.data[0], mx.data[0]


@pytest.mark.parametrize('n,nt,seed', [(0, 10, 0), (1, 10, 1), (7, 3, 2), (3000, 500, 3),
                                        (3000, 20000, 4)])
def test_vector_inc_with_min_max(hip, n, nt, seed):
    """Owner loop through an arity-3 vector map, increments bit-exact (one
    increment per element and target column, summed in element order);
    MIN/MAX exact; the sum within 1e-12.  nt=20000 leaves most targets
    unnamed; (7, 3) names targets many times per element."""
    rng = np.random.default_rng(seed)
    m = rng.integers(0, nt, size=(n, 3)).astype(np.int32) if n else np.zeros((0, 3), np.int32)
    xv = rng.random(n) - 0.5
    h = _inc3(hip, n, nt, m, xv)
    o = _inc3(seq, n, nt, m, xv)
    np.testing.assert_allclose(h[0], o[0], rtol=1e-12, atol=1e-13)
    assert h[2] == o[2] and h[3] == o[3]
    assert abs(h[1] - o[1]) <= 1e-12 * max(1.0, abs(o[1]))


@pytest.mark.parametrize('colouring', ['blocks', 'elements'])
def test_two_inc_dats_coloured(hip, colouring):
    """Two INC Dats take the coloured template, with the reference's
    block/thread colours or with element colours: within 1e-12."""
    from pyop2_amd import configuration as cfg
    with cfg.override(hip_colouring=colouring):
        _two_inc_dats(hip)


def test_mat_coloured_assembly(hip):
    """Mat assembly forced onto the coloured template (element colours by
    default for Mat loops, and block/thread colours): the P1 mass matrix
    within 1e-12 of the oracle's."""
    from pyop2_amd import configuration as cfg
    from pyop2_amd.workloads import fem, meshes
    mesh = meshes.TriMesh(12, seed=4)
    ref = None
    for colouring in (None, 'auto', 'blocks', 'elements'):
        W = seq if colouring is None else hip
        nodes = W.Set(mesh.nvert)
        vn = W.Set(mesh.nvert, 2)
        el = W.Set(mesh.nelem)
        en = W.Map(el, nodes, 3, mesh.elem_node)
        ev = W.Map(el, vn, 3, mesh.elem_node)
        x = W.Dat(vn, mesh.coords.copy(), np.float64)
        A = W.Mat(W.Sparsity((en, en)), np.float64)
        with cfg.override(hip_owner=False, hip_colouring=colouring or 'auto'):
            W.par_loop(W.Kernel(fem.MASS_TRI, 'mass_tri'), el(3, 3),
                       A((en[W.i[0]], en[W.i[1]]), W.INC), x(ev, W.READ))
        vals = np.array(A.array)
        if ref is None:
            ref = vals
        else:
            assert np.abs(vals - ref).max() <= 1e-12 * np.abs(ref).max()


def _two_inc_dats(hip):
    code = """void k(double *x, double *a, double *b) { *a += x[0]; *b += x[0] * x[0]; }"""
    rng = np.random.default_rng(5)
    n, nt = 4000, 700
    m = rng.integers(0, nt, size=(n, 2)).astype(np.int32)
    xv = rng.random(n)
    out = {}
    for name, W in (('hip', hip), ('seq', seq)):
        E, T = W.Set(n), W.Set(nt)
        mp = W.Map(E, T, 2, m)
        x = W.Dat(E, xv.copy(), np.float64)
        a = W.Dat(T, np.zeros(nt), np.float64)
        b = W.Dat(T, np.ones(nt), np.float64)
        W.par_loop(W.Kernel(code, 'k'), E, x(W.IdentityMap, W.READ), a(mp[0], W.INC),
                   b(mp[1], W.INC))
        out[name] = (a.data_ro.copy(), b.data_ro.copy())
    for u, v in zip(out['hip'], out['seq']):
        np.testing.assert_allclose(u, v, rtol=1e-12, atol=1e-14)


def test_mat_insert(hip):
    """Mat WRITE: each entry ends with the value of the last element (in
    sequential order) that writes it."""
    from pyop2_amd.workloads import meshes
    mesh = meshes.TriMesh(9, seed=6)
    code = """void k(double A[1][1], double *x[2], int i, int j) {
      A[0][0] = x[i][0] * 10.0 + x[j][1] + (double)(i * 3 + j); }"""
    out = {}
    for name, W in (('hip', hip), ('seq', seq)):
        nodes = W.Set(mesh.nvert)
        vn = W.Set(mesh.nvert, 2)
        el = W.Set(mesh.nelem)
        en = W.Map(el, nodes, 3, mesh.elem_node)
        ev = W.Map(el, vn, 3, mesh.elem_node)
        x = W.Dat(vn, mesh.coords.copy(), np.float64)
        A = W.Mat(W.Sparsity((en, en)), np.float64)
        W.par_loop(W.Kernel(code, 'k'), el(3, 3), A((en[W.i[0]], en[W.i[1]]), W.WRITE),
                   x(ev, W.READ))
        out[name] = A.values.copy()
    np.testing.assert_array_equal(out['hip'], out['seq'])


def test_repeated_target_in_element(hip):
    """An element naming one target in both columns increments it twice, in
    column order, as the sequential loop does."""
    code = "void k(double *x, double *a, double *b) { *a += x[0]; *b += 3.0 * x[0]; }"
    n, nt = 500, 40
    rng = np.random.default_rng(8)
    m = rng.integers(0, nt, size=(n, 2)).astype(np.int32)
    m[::5, 1] = m[::5, 0]
    xv = rng.random(n)
    out = {}
    for name, W in (('hip', hip), ('seq', seq)):
        E, T = W.Set(n), W.Set(nt)
        mp = W.Map(E, T, 2, m)
        x = W.Dat(E, xv.copy(), np.float64)
        a = W.Dat(T, np.zeros(nt), np.float64)
        W.par_loop(W.Kernel(code, 'k'), E, x(W.IdentityMap, W.READ), a(mp[0], W.INC),
                   a(mp[1], W.INC))
        out[name] = a.data_ro.copy()
    np.testing.assert_array_equal(out['hip'], out['seq'])
