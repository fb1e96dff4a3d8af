"""Device data layout: locality renumbering of Sets.

The reference stores every Dat in the user's element order.  On meshes
whose numbering is scattered, every gather through a map then touches a
different cache line, and an MI355X loop moves many times its unique bytes.
The hip backend therefore stores each Set's device data in a *layout order*
chosen for locality, decided once per Set at its first device use:

0. an iteration Set of an owner-computes loop is sorted by the position of
   the target its race map names first (edges follow their first cell);
1. a Set that maps into an already laid-out Set is sorted by the first
   position its elements name there (edges follow their cells);
2. a Set that an already laid-out Set maps into takes the order in which
   those elements first name it (nodes follow the cells that use them);
3. otherwise a recursive bisection of the Set's adjacency through the
   largest map into it (cells adjacent across edges): breadth-first cuts in
   halves down to ranges of 32, so every aligned range of positions is a
   compact patch of the mesh;
4. otherwise the identity.

Sets used directly in one loop (a Dat on ``vcells`` read in a loop over
``cells``) must share their layout, so they are linked into one group that
is decided together.  ``perm[pos]`` is the user index stored at device
position ``pos``; ``iperm`` the inverse.  User-visible data (``.data``,
``Sparsity.rowptr/colidx``, ``Mat.array``) stay in user numbering: transfers
and the host views translate.
"""

import ctypes

import numpy as np

from . import _native as nat
from . import configuration as cfg
from .exceptions import MapValueError


class Layout(object):

    def __init__(self, perm):
        self.perm = None if perm is None else np.ascontiguousarray(perm, dtype=np.int64)
        if self.perm is None:
            self.iperm = None
        else:
            self.iperm = np.empty_like(self.perm)
            self.iperm[self.perm] = np.arange(len(self.perm), dtype=np.int64)

    @property
    def identity(self):
        return self.perm is None

    def to_device_order(self, a):
        """Rows of a user-ordered array in device order."""
        return a if self.perm is None else a[self.perm]

    def to_user_order(self, a, out):
        """Scatter device-ordered rows ``a`` into user-ordered ``out``."""
        if self.perm is None:
            out[...] = a
        else:
            out[self.perm] = a


IDENTITY = Layout(None)


class _Group(object):
    """Sets that share one layout (union-find node)."""

    def __init__(self, s):
        self.parent = self
        self.sets = [s]
        self.layout = None

    def root(self):
        g = self
        while g.parent is not g:
            g = g.parent
        return g


def _group(s):
    g = getattr(s, '_layout_group', None)
    if g is None:
        g = s._layout_group = _Group(s)
    return g.root()


def link(a, b):
    """Give Sets ``a`` and ``b`` (same total size) one layout."""
    ga, gb = _group(a), _group(b)
    if ga is gb or a.total_size != b.total_size:
        return
    if ga.layout is not None and gb.layout is not None:
        if not _same(ga.layout, gb.layout):
            raise RuntimeError("Sets %s and %s are accessed directly in one loop but were laid "
                               "out differently on the device" % (a.name, b.name))
        gb.parent = ga
        ga.sets += gb.sets
        return
    if gb.layout is not None:
        ga, gb = gb, ga
    gb.parent = ga
    ga.sets += gb.sets


def _same(x, y):
    if x.perm is None or y.perm is None:
        return x.perm is None and y.perm is None
    return np.array_equal(x.perm, y.perm)


def register_map(m):
    """Record a map on its sets (called when a hip Map is built).  Two maps
    from one Set with identical values into two Sets of one size (a mesh's
    element-to-node map onto scalar, vector and coordinate node Sets) give
    their target Sets one layout, so the maps' device copies are identical
    too (one buffer, one gather pattern)."""
    for other in getattr(m.iterset, '_maps_from', []):
        if (other.dataset is not m.dataset and other.dim == m.dim and
                other.dataset.total_size == m.dataset.total_size and
                np.shape(other.values) == np.shape(m.values) and
                np.array_equal(other.values, m.values)):
            ga, gb = _group(other.dataset), _group(m.dataset)
            if ga is not gb and (ga.layout is None or gb.layout is None):
                link(other.dataset, m.dataset)
            break
    for s, attr in ((m.iterset, '_maps_from'), (m.dataset, '_maps_to')):
        lst = getattr(s, attr, None)
        if lst is None:
            lst = []
            setattr(s, attr, lst)
        lst.append(m)


def values(m):
    """The values of Map ``m`` after a one-time check that every entry
    names an element of its target Set.  Every device use of a map (the
    layout decisions, which index permutations with its values, and the
    upload) reads them through here: an entry out of range would otherwise
    wrap to another element on the host (a negative numpy index) or gather
    outside the target's device array."""
    if m.values is not None and not getattr(m, '_range_checked', False):
        v = np.asarray(m.values)
        if v.size:
            lo, hi = int(v.min()), int(v.max())
            if lo < 0 or hi >= m.dataset.total_size:
                bad = lo if lo < 0 else hi
                raise MapValueError("Map %s has the value %d, outside its target Set %s of "
                                    "size %d" % (m.name, bad, m.dataset.name,
                                                 m.dataset.total_size))
        m._range_checked = True
    return m.values


def decided(s):
    return _group(s).layout is not None


def prefer(s, m, col=0):
    """Ask that Set ``s``, if still undecided, be sorted by the position
    of the target its map ``m`` names in column ``col`` (an owner-computes
    loop over ``s`` then finds each block's primary elements consecutive)."""
    if not decided(s):
        s._layout_pref = (m, col)


def of(s):
    """The layout of Set ``s``, deciding it (for its whole group) if needed."""
    g = _group(s)
    if g.layout is None:
        g.layout = IDENTITY  # guards recursion while deciding
        g.layout = _decide(g)
    return g.layout


def _decide(g):
    if not cfg.get('hip_layout', True):
        return IDENTITY
    sets = g.sets
    n = sets[0].total_size
    if n <= 1:
        return IDENTITY
    maps_from = [m for s in sets for m in getattr(s, '_maps_from', [])
                 if len(m.values) and m.iterset.total_size == n]
    maps_to = [m for s in sets for m in getattr(s, '_maps_to', []) if len(m.values)]
    # 0. a preferred map (set by an owner-computes loop): sort by the
    #    position of the target in the preferred column
    for s in sets:
        pref = getattr(s, '_layout_pref', None)
        if pref is not None and len(pref[0].values) and pref[0].iterset.total_size == n:
            m, col = pref
            T = of(m.dataset)
            if _group(m.dataset) is not g and not T.identity:
                key = T.iperm[np.asarray(values(m))[:, col]]
                return Layout(np.argsort(key, kind='stable'))
    # 1. follow the first laid-out target set
    for m in maps_from:
        T = m.dataset
        if _group(T) is not g and decided(T) and not of(T).identity:
            key = of(T).iperm[values(m)].min(axis=1)
            return Layout(np.argsort(key, kind='stable'))
    # 2. first touch from a laid-out source set
    for m in maps_to:
        U = m.iterset
        if _group(U) is not g and decided(U) and not of(U).identity:
            src = np.ascontiguousarray(of(U).perm, dtype=np.int32)
            return Layout(_first_touch(n, m, src))
    # 3. recursive bisection (or breadth-first order) of the adjacency
    #    through the largest map into the set
    if maps_to:
        m = max(maps_to, key=lambda x: x.values.size)
        return Layout(adjacency_order(n, values(m), cfg.get('hip_layout_order', 'bisect')))
    # 4. an iteration set only: lay out its largest target first, then follow it
    if maps_from:
        m = max(maps_from, key=lambda x: x.dataset.total_size)
        if _group(m.dataset) is not g:
            T = of(m.dataset)
            if not T.identity:
                key = T.iperm[values(m)].min(axis=1)
                return Layout(np.argsort(key, kind='stable'))
    return IDENTITY


def adjacency_order(n, values, how='bisect', leaf=32):
    """perm of ``n`` targets ordered for locality through map ``values``
    (nelem x arity): 'bisect' (recursive breadth-first bisection) or 'bfs'."""
    v = np.ascontiguousarray(values, dtype=np.int32)
    perm = np.empty(n, dtype=np.int32)
    vp, pp = v.ctypes.data_as(ctypes.c_void_p), perm.ctypes.data_as(ctypes.c_void_p)
    if how == 'bfs':
        nat.check(nat.lib.op2hip_order_bfs(n, v.shape[0], vp, v.shape[1], pp), 'bfs ordering')
    else:
        nat.check(nat.lib.op2hip_order_bisect(n, v.shape[0], vp, v.shape[1], leaf, pp),
                  'bisection ordering')
    return perm


def _first_touch(n, m, src_order):
    v = np.ascontiguousarray(values(m), dtype=np.int32)
    perm = np.empty(n, dtype=np.int32)
    nat.check(nat.lib.op2hip_order_first_touch(
        n, len(src_order), v.ctypes.data_as(ctypes.c_void_p), v.shape[1],
        src_order.ctypes.data_as(ctypes.c_void_p), perm.ctypes.data_as(ctypes.c_void_p)),
        'first-touch ordering')
    return perm


def map_values(m):
    """Map values in device numbering: rows in the iteration set's layout
    order, entries as positions in the target set's layout (int32)."""
    v = getattr(m, '_layout_vals', None)
    if v is None:
        v = np.asarray(values(m))
        src, dst = of(m.iterset), of(m.dataset)
        if not src.identity:
            v = v[src.perm]
        if not dst.identity:
            v = dst.iperm[v]
        v = np.ascontiguousarray(v, dtype=np.int32)
        m._layout_vals = v
    return v


def csr_to_user(rp, ci, rmult, cmult, rl, cl):
    """A CSR pattern in device numbering -> (rowptr, colidx, nnz, slotmap)
    in user numbering.  Rows are ``rmult*node + r`` of the row Set (layout
    ``rl``), columns ``cmult*node + c`` of the column Set (layout ``cl``);
    slotmap[k] is the device slot of user slot k (None when both layouts
    are the identity).  Each user row's columns come out sorted."""
    lsize = len(rp) - 1
    if rl.identity and cl.identity:
        return rp, ci, np.diff(rp).astype(np.int32), None
    drow = np.repeat(np.arange(lsize, dtype=np.int64), np.diff(rp))
    urow = drow if rl.identity else rmult * rl.perm[drow // rmult] + drow % rmult
    dcol = np.asarray(ci, dtype=np.int64)
    ucol = dcol if cl.identity else cmult * cl.perm[dcol // cmult] + dcol % cmult
    order = np.lexsort((ucol, urow))
    counts = np.bincount(urow, minlength=lsize)
    urp = np.zeros(lsize + 1, dtype=np.int32)
    np.cumsum(counts, out=urp[1:])
    return urp, ucol[order].astype(np.int32), counts.astype(np.int32), order
