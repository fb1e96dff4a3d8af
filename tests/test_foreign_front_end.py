"""The backend driven by another op2 front end, as the reference's would
drive it once registered (pyop2/backends.py:103-146).

The reference's ``op2`` module hands user code its own descriptors
(pyop2/op2.py:43: READ ... MAX, IdentityMap and i from pyop2/base.py:56-90,
981-1011, 1127) and looks backend classes up by name in the backend
module's ``__dict__`` (backends.py:103-114).  The front end below is a
stand-in for it: descriptor classes with the reference's names and
attributes (distinct objects from this package's), and a selector
metaclass that dispatches by class name.  Programs written against it must
give the results of the same programs written against this package's own
``op2``: on the oracle here, on the hip backend in the GPU tests.
"""

import numpy as np
import pytest

from oracle import sequential as seq
from pyop2_amd import base, codegen


class _RefBase(object):
    """Descriptor classes shaped like the reference's (same class names,
    attributes and identity semantics)."""

    class Access(object):
        def __init__(self, mode):
            self._mode = mode

    class IterationIndex(object):
        def __init__(self, index=None):
            self._index = index

        @property
        def index(self):
            return self._index

        def __getitem__(self, idx):
            return type(self)(idx)

    class Set(object):
        def __init__(self, size):
            self.size = size

    class Map(object):
        def __init__(self, iterset, dataset, dim, values, name):
            self.iterset, self.dataset, self.dim, self.values, self.name = \
                iterset, dataset, dim, values, name


def front_end(backend_module):
    """A namespace like the reference's op2 with ``backend_module``
    selected: classes dispatch by name, descriptors are the front end's."""

    class _Selector(type):
        def __call__(cls, *args, **kwargs):
            return backend_module.__dict__[cls.__name__](*args, **kwargs)

    ns = {}
    for name in ('Set', 'Map', 'Dat', 'Global', 'Kernel', 'Sparsity', 'Mat'):
        ns[name] = _Selector(name, (object,), {})
    R = _RefBase
    ns.update(READ=R.Access('READ'), WRITE=R.Access('WRITE'), RW=R.Access('RW'),
              INC=R.Access('INC'), MIN=R.Access('MIN'), MAX=R.Access('MAX'),
              i=R.IterationIndex(),
              IdentityMap=R.Map(R.Set(0), R.Set(0), 1, [], 'identity'),
              par_loop=lambda k, it, *args: backend_module.par_loop(k, it, *args))
    return type('op2', (object,), ns)


DIRECT = """
void bump(double *x, double *s) { x[0] += 1.0; *s += x[0]; }
"""
DEGREE = """
void degree(double *d, double *w, int j) { d[0] += w[0]; }
"""
EDGE_MAT = """
void edge_mat(double A[1][1], double *w, int j, int k) { A[0][0] += (j == k ? w[0] : -w[0]); }
"""


def edge_mesh(seed=0, nnode=40, nedge=90):
    rng = np.random.default_rng(seed)
    e = rng.integers(0, nnode, size=(nedge, 2)).astype(np.int32)
    e[:, 1] = np.where(e[:, 1] == e[:, 0], (e[:, 0] + 1) % nnode, e[:, 1])
    return e, rng.random(nedge)


def program(op2, edges_nodes, w, tag):
    """Direct RW loop with a Global INC, itspace INC through a map and a Mat
    assembly, written against ``op2``."""
    nnode, nedge = int(edges_nodes.max()) + 1, len(edges_nodes)
    nodes = op2.Set(nnode, 1, tag + 'n')
    edges = op2.Set(nedge, 1, tag + 'e')
    en = op2.Map(edges, nodes, 2, edges_nodes, tag + 'en')
    x = op2.Dat(nodes, np.arange(nnode, dtype=np.float64), np.float64, tag + 'x')
    s = op2.Global(1, 0.0, np.float64, tag + 's')
    op2.par_loop(op2.Kernel(DIRECT, 'bump'), nodes, x(op2.IdentityMap, op2.RW), s(op2.INC))
    d = op2.Dat(nodes, np.zeros(nnode), np.float64, tag + 'd')
    wd = op2.Dat(edges, w.copy(), np.float64, tag + 'w')
    op2.par_loop(op2.Kernel(DEGREE, 'degree'), edges(2), d(en[op2.i[0]], op2.INC),
                 wd(op2.IdentityMap, op2.READ))
    sp = op2.Sparsity((en, en), tag + 'sp')
    A = op2.Mat(sp, np.float64, tag + 'A')
    op2.par_loop(op2.Kernel(EDGE_MAT, 'edge_mat'), edges(2, 2),
                 A((en[op2.i[0]], en[op2.i[1]]), op2.INC), wd(op2.IdentityMap, op2.READ))
    return (np.array(x.data_ro).reshape(-1), float(s.data[0]), np.array(d.data_ro).reshape(-1),
            A.values.copy())


def expected(edges_nodes, w):
    nnode = int(edges_nodes.max()) + 1
    x = np.arange(nnode, dtype=np.float64) + 1.0
    d = np.zeros(nnode)
    np.add.at(d, edges_nodes[:, 0], w)
    np.add.at(d, edges_nodes[:, 1], w)
    A = np.zeros((nnode, nnode))
    for (a, b), v in zip(edges_nodes, w):
        A[a, a] += v
        A[b, b] += v
        A[a, b] -= v
        A[b, a] -= v
    return x, x.sum(), d, A


def test_descriptors_are_matched_by_role():
    ref = front_end(seq)
    assert base.canonical(ref.READ) is base.READ
    assert base.canonical(ref.INC) is base.INC
    assert base.canonical(ref.IdentityMap) is base.IdentityMap
    assert base.canonical(ref.i[1]).index == 1
    assert base.canonical(ref.i) is base.i
    assert codegen._mode(ref.MAX) == 'MAX'
    # other objects pass through unchanged
    other = object()
    assert base.canonical(other) is other


def test_oracle_through_foreign_front_end():
    """The oracle backend driven by the stand-in front end gives the closed
    forms (and what this package's own op2 names give)."""
    e, w = edge_mesh(1)
    x, s, d, A = program(front_end(seq), e, w, 'ff')
    xe, se, de, Ae = expected(e, w)
    assert np.array_equal(x, xe)
    assert s == se
    assert np.abs(d - de).max() <= 1e-14 * np.abs(de).max()
    assert np.abs(A - Ae).max() <= 1e-14 * np.abs(Ae).max()
    own = program(seq, e, w, 'fo')
    assert np.array_equal(own[0], x) and own[1] == s
    assert np.array_equal(own[2], d) and np.array_equal(own[3], A)


def test_hip_classifies_foreign_arguments():
    """hip's loop classification (stream or owner-computes) sees the
    front end's descriptors as its own (no GPU: arguments only)."""
    from pyop2_amd import hip
    ref = front_end(hip)
    e, w = edge_mesh(2)
    nodes = ref.Set(int(e.max()) + 1, 1, 'cn')
    edges = ref.Set(len(e), 1, 'ce')
    en = ref.Map(edges, nodes, 2, e, 'cen')
    d = ref.Dat(nodes, np.zeros(nodes.size), np.float64, 'cd')
    wd = ref.Dat(edges, w, np.float64, 'cw')
    k = ref.Kernel(DEGREE, 'degree')
    a_inc = d(en[ref.i[0]], ref.INC)
    a_rd = wd(ref.IdentityMap, ref.READ)
    assert a_inc.access is base.INC and a_rd.map is base.IdentityMap
    assert isinstance(a_inc.idx, base.IterationIndex) and a_inc.idx.index == 0
    assert hip.ParLoop(k, edges(2), a_inc, a_rd)._mode() == 'owner'
    assert hip.ParLoop(k, edges, wd(ref.IdentityMap, ref.RW))._mode() == 'stream'


@pytest.mark.gpu
def test_hip_through_foreign_front_end(hip):
    """The hip backend driven by the stand-in front end matches the oracle
    driven by it: direct loop and Global bit-exact, owner-computes INC and
    Mat assembly within 1e-12 of the array's maximum."""
    from pyop2_amd import hip as hipmod
    e, w = edge_mesh(3, nnode=300, nedge=900)
    xh, sh, dh, Ah = program(front_end(hipmod), e, w, 'gh')
    xo, so, do, Ao = program(front_end(seq), e, w, 'go')
    assert np.array_equal(xh, xo)
    assert sh == so
    assert np.abs(dh - do).max() <= 1e-12 * np.abs(do).max()
    assert np.abs(Ah - Ao).max() <= 1e-12 * np.abs(Ao).max()
