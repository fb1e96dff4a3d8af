"""Deferred Global reductions on the hip backend.

A loop reducing into one FP64 Global (INC, MIN or MAX) leaves its combine
pass pending; the next loop launch that neither reduces nor names that
Global runs it in one extra workgroup, whose waves stand in for the 16
wavefronts of the combine kernel, so the Global gets the value the combine
kernel would have given, bit for bit.  Reading the Global, a loop that
uses it or reduces itself, and graph capture run a pending pass first.

Each case runs with ``hip_defer_reductions`` on and off and compares the
Globals and Dats bit for bit: partial counts from 1 to beyond 4096 (the
combine kernel's per-thread stride), dimensions 1 and 3, each consumer
template (direct loops, owner-computes increments, row-block Mat assembly
with 64-thread blocks, coloured increments), the airfoil step (its rms)
eagerly, as the reference's demo drives it and replayed as a graph, and
the flush points."""

import numpy as np
import pytest

from pyop2_amd import configuration as cfg
from pyop2_amd.workloads import airfoil, fem, meshes

pytestmark = pytest.mark.gpu

OPS = {'INC': 'g[k] += x[k];', 'MIN': 'if (x[k] < g[k]) g[k] = x[k];',
       'MAX': 'if (x[k] > g[k]) g[k] = x[k];'}


def _reduce_loop(op2, s, x, g, mode, cdim):
    k = op2.Kernel('void red_%s(double *x, double *g) { for (int k = 0; k < %d; ++k) { %s } }'
                   % (mode.lower(), cdim, OPS[mode]), 'red_%s' % mode.lower())
    op2.par_loop(k, s, x(op2.IdentityMap, op2.READ), g(getattr(op2, mode)))


def _consumer(op2, kind, tri):
    """A loop of template ``kind`` touching neither Global."""
    if kind == 'stream':
        d = op2.Dat(tri['nodes'], np.arange(tri['mesh'].nvert, dtype=np.float64), np.float64)
        op2.par_loop(op2.Kernel('void dbl(double *a) { a[0] *= 2.0; }', 'dbl'), tri['nodes'],
                     d(op2.IdentityMap, op2.RW))
        return d
    if kind in ('owner', 'coloured'):
        d = op2.Dat(tri['nodes'], np.zeros(tri['mesh'].nvert), np.float64)
        k = op2.Kernel('void cnt(double *a, double *b, double *c) '
                       '{ a[0] += 1.0; b[0] += 2.0; c[0] += 3.0; }', 'cnt')
        en = tri['en']
        with cfg.override(hip_owner=(kind == 'owner')):
            op2.par_loop(k, tri['elems'], d(en[0], op2.INC), d(en[1], op2.INC), d(en[2], op2.INC))
        return d
    A = op2.Mat(tri['sp'], np.float64)
    en, ev = tri['en'], tri['ev']
    op2.par_loop(op2.Kernel(fem.MASS_TRI, 'mass_tri'), tri['elems'](3, 3),
                 A((en[op2.i[0]], en[op2.i[1]]), op2.INC), tri['x'](ev, op2.READ))
    return A


@pytest.fixture(scope='module')
def tri(hip):
    op2 = hip
    mesh = meshes.TriMesh(40, seed=3)
    nodes = op2.Set(mesh.nvert)
    vnodes = op2.Set(mesh.nvert, 2)
    elems = op2.Set(mesh.nelem)
    en = op2.Map(elems, nodes, 3, mesh.elem_node)
    ev = op2.Map(elems, vnodes, 3, mesh.elem_node)
    return {'mesh': mesh, 'nodes': nodes, 'elems': elems, 'en': en, 'ev': ev,
            'x': op2.Dat(vnodes, mesh.coords.copy(), np.float64),
            'sp': op2.Sparsity((en, en))}


def _case(op2, tri, n, mode, cdim, kind, defer, blocks):
    from pyop2_amd import hip as H
    rng = np.random.default_rng(n + cdim)
    xs = rng.standard_normal((n, cdim)) * 10.0 ** rng.uniform(-3, 3, (n, 1))
    s = op2.Set(n, cdim)
    x = op2.Dat(s, xs, np.float64)
    g0 = {'INC': 0.25, 'MIN': 1e300, 'MAX': -1e300}[mode]
    g = op2.Global(cdim, [g0] * cdim, np.float64)
    with cfg.override(hip_defer_reductions=defer, hip_stream_blocks=blocks):
        _reduce_loop(op2, s, x, g, mode, cdim)
        assert (H._pending[0] is not None) == defer
        out = _consumer(op2, kind, tri)
        # the consumer's launch took the pending combine pass along
        assert H._pending[0] is None
    vals = out.array if kind == 'rows' else out.data_ro
    return np.array(g.data), np.array(vals)


@pytest.mark.parametrize('n,blocks', [(300, 4096), (70000, 4096), (3000000, 8192)])
@pytest.mark.parametrize('mode', ['INC', 'MIN', 'MAX'])
def test_deferred_equals_combine_kernel(hip, tri, n, blocks, mode):
    a = _case(hip, tri, n, mode, 1, 'stream', True, blocks)
    b = _case(hip, tri, n, mode, 1, 'stream', False, blocks)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


@pytest.mark.parametrize('kind', ['stream', 'owner', 'coloured', 'rows'])
@pytest.mark.parametrize('cdim', [1, 3])
def test_deferred_each_consumer_template(hip, tri, kind, cdim):
    a = _case(hip, tri, 200000, 'INC', cdim, kind, True, 4096)
    b = _case(hip, tri, 200000, 'INC', cdim, kind, False, 4096)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    # the values themselves: the partial sums of the data, within rounding
    rng = np.random.default_rng(200000 + cdim)
    xs = rng.standard_normal((200000, cdim)) * 10.0 ** rng.uniform(-3, 3, (200000, 1))
    assert np.allclose(a[0], 0.25 + xs.sum(axis=0), rtol=1e-12, atol=1e-9)


def test_flush_points(hip, tri):
    from pyop2_amd import hip as H
    op2 = hip
    n = 5000
    s = op2.Set(n, 1)
    x = op2.Dat(s, np.arange(n, dtype=np.float64), np.float64)
    want = 0.25 + float(np.arange(n).sum())
    # reading the value
    g = op2.Global(1, 0.25, np.float64)
    _reduce_loop(op2, s, x, g, 'INC', 1)
    assert H._pending[0] is not None
    assert g.data[0] == want and H._pending[0] is None
    # a loop reading the Global runs the pass first
    g = op2.Global(1, 0.25, np.float64)
    _reduce_loop(op2, s, x, g, 'INC', 1)
    y = op2.Dat(s, np.zeros(n), np.float64)
    op2.par_loop(op2.Kernel('void cp(double *y, double *g) { y[0] = g[0]; }', 'cp'), s,
                 y(op2.IdentityMap, op2.WRITE), g(op2.READ))
    assert (y.data_ro == want).all()
    # the same loop twice: the first pass runs before the second overwrites
    # the partials
    g = op2.Global(1, 0.25, np.float64)
    _reduce_loop(op2, s, x, g, 'INC', 1)
    _reduce_loop(op2, s, x, g, 'INC', 1)
    assert g.data[0] == 0.25 + 2 * float(np.arange(n).sum())
    # a host write of the Global after the loop is not overwritten later
    g = op2.Global(1, 0.25, np.float64)
    _reduce_loop(op2, s, x, g, 'INC', 1)
    g.data = np.array([7.0])
    _consumer(op2, 'stream', tri)
    assert g.data[0] == 7.0


def _airfoil_run(op2, defer, how):
    from pyop2_amd import hip as H
    mesh = meshes.QuadMesh(64, 32, seed=5)
    q0 = airfoil.initial_q(mesh.ncell, seed=2)
    with cfg.override(hip_defer_reductions=defer):
        A = airfoil.Airfoil(op2, mesh, q0=q0)
        try:
            rms = []
            if how == 'graph':
                A.iteration()
                g = H.capture(A.iteration)
                for _ in range(3):
                    g.replay()
                rms.append(A.rms.data[0])
            elif how == 'demo':
                for _ in range(3):
                    rms.append(A.demo_iteration())
            else:
                for _ in range(4):
                    A.iteration()
                rms.append(A.rms.data[0])
            return np.array(rms), np.array(A.p_q.data_ro)
        finally:
            A.remove_consts()


@pytest.mark.parametrize('how', ['eager', 'demo', 'graph'])
def test_airfoil_rms_deferred(hip, how):
    a = _airfoil_run(hip, True, how)
    b = _airfoil_run(hip, False, how)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
