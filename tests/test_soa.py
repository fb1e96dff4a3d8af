"""Structure-of-arrays Dats (``soa=True``): stored component-major on the
device, component k of row r at k * rows + r (reference: device.py:153-174,
cuda_direct_loop.jinja2:80-82), while user-visible data stay rows of
components and kernels see a row (``OP2_STRIDE(a, k)`` is ``a[k]``).

CPU: the generated sources of every loop form (direct, owner-computes,
coloured; gathered, written, incremented, vector-map and iteration-space
arguments) compile for gfx950.  GPU: the airfoil step with q, q_old and
res stored as SoA is bit-exact to the oracle (and to the AoS step), in the
owner and coloured forms and through vector maps; SoA coordinates in Mat
assembly give the AoS result bit for bit; host round trips keep values.
"""

import numpy as np
import pytest

from oracle import sequential as seq
from pyop2_amd.workloads import airfoil, meshes


def _sources(W, soa):
    from pyop2_amd import base, codegen
    from pyop2_amd.codegen import LoopSource
    mesh = meshes.QuadMesh(8, 4, seed=1)
    A = airfoil.Airfoil(W, mesh, soa=soa)
    consts = base.Const._definitions()
    out = []
    specs = [
        (A.k_save_soln, (), [A.p_q(W.IdentityMap, W.READ), A.p_qold(W.IdentityMap, W.WRITE)],
         ['stream']),
        (A.k_res_calc, (), [A.p_x(A.pedge[0], W.READ), A.p_x(A.pedge[1], W.READ),
                            A.p_q(A.pevcell[0], W.READ), A.p_q(A.pevcell[1], W.READ),
                            A.p_adt(A.pecell[0], W.READ), A.p_adt(A.pecell[1], W.READ),
                            A.p_res(A.pevcell[0], W.INC), A.p_res(A.pevcell[1], W.INC)],
         ['owner', 'coloured']),
        (A.k_update, (), [A.p_qold(W.IdentityMap, W.READ), A.p_q(W.IdentityMap, W.WRITE),
                          A.p_res(W.IdentityMap, W.RW), A.p_adt(W.IdentityMap, W.READ),
                          A.rms(W.INC)], ['stream']),
    ]
    for kernel, ext, args, modes in specs:
        for mode in modes:
            src = LoopSource(kernel, ext, args, consts, mode, 256)
            src.map_of = {}
            out.append(src.source())
    # vector-map READ and INC, iteration-space READ and INC
    vk = W.Kernel("""
void vsum(double *q[4], double *r[4]) {
  for (int j = 0; j < 2; ++j) for (int k = 0; k < 4; ++k) r[j][k] += OP2_STRIDE(q[j], k);
}""", 'vsum')
    out.append(LoopSource(vk, (), [A.p_q(A.pevcell, W.READ), A.p_res(A.pevcell, W.INC)],
                          consts, 'owner', 256).source())
    ik = W.Kernel("""
void isum(double *q, double *r, int j) { for (int k = 0; k < 4; ++k) r[k] += OP2_STRIDE(q, k); }
""", 'isum')
    out.append(LoopSource(ik, (2,), [A.p_q(A.pevcell[W.i[0]], W.READ),
                                     A.p_res(A.pevcell[W.i[0]], W.INC)],
                          consts, 'owner', 256).source())
    A.remove_consts()
    return out


def test_soa_sources_compile():
    from pyop2_amd import hip, jit
    for src in _sources(hip, True):
        assert 'OP2_STRIDE' in src
        assert len(jit.compile_source(src)) > 0


@pytest.fixture
def consts(hip):
    """The airfoil Consts, declared once and read by both worlds."""
    cs = [hip.Const(1, v, k, np.float64) for k, v in airfoil.CONSTS.items()]
    cs.append(hip.Const(4, airfoil.free_stream(), 'qinf', np.float64))
    yield
    for c in cs:
        c.remove_from_namespace()


def _run(W, mesh, q0, soa, opts=None, vector=False, steps=2):
    from pyop2_amd import configuration as cfg
    cls = airfoil.AirfoilVector if vector else airfoil.Airfoil
    A = cls(W, mesh, q0=q0, soa=soa, consts=False)
    with cfg.override(**(opts or {})):
        for _ in range(steps):
            A.iteration()
    return (np.array(A.p_q.data_ro), np.array(A.p_res.data_ro), float(A.rms.data[0]),
            np.array(A.p_adt.data_ro))


@pytest.mark.gpu
@pytest.mark.parametrize('vector', [False, True])
def test_airfoil_soa_bit_exact(hip, consts, vector):
    mesh = meshes.QuadMesh(96, 48, seed=13)
    q0 = airfoil.initial_q(mesh.ncell, seed=4)
    h = _run(hip, mesh, q0, True, vector=vector)
    o = _run(seq, mesh, q0, False, vector=vector)
    aos = _run(hip, mesh, q0, False, vector=vector)
    assert np.array_equal(h[0], o[0]) and np.array_equal(h[1], o[1])
    assert np.array_equal(h[3], o[3])
    assert abs(h[2] - o[2]) <= 1e-12 * abs(o[2])
    assert np.array_equal(h[0], aos[0]) and h[2] == aos[2]


@pytest.mark.gpu
def test_airfoil_soa_coloured(hip, consts):
    mesh = meshes.QuadMesh(64, 32, seed=14)
    q0 = airfoil.initial_q(mesh.ncell, seed=5)
    h = _run(hip, mesh, q0, True, opts={'hip_owner': False}, steps=1)
    a = _run(hip, mesh, q0, False, opts={'hip_owner': False}, steps=1)
    o = _run(seq, mesh, q0, False, steps=1)
    assert np.array_equal(h[0], a[0])       # same colours, same order
    assert np.abs(h[0] - o[0]).max() <= 1e-12 * np.abs(o[0]).max()


@pytest.mark.gpu
def test_soa_roundtrip_and_mat_coordinates(hip):
    from pyop2_amd.workloads import fem
    rng = np.random.default_rng(3)
    s = hip.Set(1000, 3, 'rt3')
    v = rng.random((1000, 3))
    d = hip.Dat(s, v.copy(), np.float64, 'rt', soa=True)
    k = hip.Kernel("void twice(double *a) { for (int j = 0; j < 3; ++j) OP2_STRIDE(a, j) *= 2.0; }",
                   'twice')
    hip.par_loop(k, s, d(hip.IdentityMap, hip.RW))
    assert np.array_equal(d.data_ro, 2.0 * v)
    # P1 mass with structure-of-arrays coordinates (vector-map reads)
    mesh = meshes.TriMesh(30, seed=2)
    out = []
    for soa in (False, True):
        nodes = hip.Set(mesh.nvert, 1, 'mn%d' % soa)
        vn = hip.Set(mesh.nvert, 2, 'mv%d' % soa)
        el = hip.Set(mesh.nelem, 1, 'me%d' % soa)
        en = hip.Map(el, nodes, 3, mesh.elem_node, 'men%d' % soa)
        ev = hip.Map(el, vn, 3, mesh.elem_node, 'mev%d' % soa)
        x = hip.Dat(vn, mesh.coords.copy(), np.float64, 'mx%d' % soa, soa=soa)
        A = hip.Mat(hip.Sparsity((en, en), 'msp%d' % soa), np.float64, 'mA%d' % soa)
        hip.par_loop(hip.Kernel(fem.MASS_TRI, 'mass_tri'), el(3, 3),
                     A((en[hip.i[0]], en[hip.i[1]]), hip.INC), x(ev, hip.READ))
        out.append(A.array.copy())
    assert np.array_equal(out[0], out[1])


@pytest.mark.gpu
def test_solve_on_soa_dat_refused(hip):
    s = hip.Set(4, 2, 'sv')
    en = hip.Map(hip.Set(2, 1, 'se'), hip.Set(4, 1, 'sn'), 2, np.array([[0, 1], [2, 3]]), 'sm')
    vm = hip.Map(en.iterset, s, 2, np.array([[0, 1], [2, 3]]), 'svm')
    A = hip.Mat(hip.Sparsity((vm, vm), 'ssp'), np.float64, 'sA')
    x = hip.Dat(s, np.zeros((4, 2)), np.float64, 'sx', soa=True)
    b = hip.Dat(s, np.ones((4, 2)), np.float64, 'sb', soa=True)
    with pytest.raises(NotImplementedError):
        hip.solve(A, x, b)
