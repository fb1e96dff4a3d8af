"""Seeded synthetic meshes with randomly permuted numbering.

All generators return plain numpy arrays (int32 maps, float64 coordinates);
numbering of every set is permuted with ``np.random.default_rng(seed)`` so
that gathers through maps are irregular, as on real unstructured meshes.
"""

import numpy as np


def _perm(rng, n, permute):
    return rng.permutation(n).astype(np.int64) if permute else np.arange(n, dtype=np.int64)


class QuadMesh(object):
    """Airfoil-style 2-D quad mesh of nx x ny cells on [0, lx] x [0, ly].

    Sets: nodes ((nx+1)(ny+1)), cells (nx*ny), interior edges
    ((nx-1)ny + nx(ny-1)), boundary edges (2(nx+ny)).  Maps (int32):
    ``pedge`` (edge -> 2 nodes), ``pecell`` (edge -> 2 cells),
    ``pbedge`` (bedge -> 2 nodes), ``pbecell`` (bedge -> 1 cell),
    ``pcell`` (cell -> 4 nodes, counter-clockwise).  An interior edge's node
    order makes its normal (dy, -dx) point from its first cell to its
    second; a boundary edge's normal points out of the domain.  ``bound``
    is a seeded int32 flag in {0, 1} per boundary edge."""

    def __init__(self, nx, ny, seed=0, permute=True, lx=2.0, ly=1.0):
        rng = np.random.default_rng(seed)
        self.nx, self.ny = nx, ny
        nnode = (nx + 1) * (ny + 1)
        ncell = nx * ny
        npn = _perm(rng, nnode, permute)   # old node id -> new id
        npc = _perm(rng, ncell, permute)

        def node(i, j):
            return npn[np.asarray(j) * (nx + 1) + np.asarray(i)]

        def cell(i, j):
            return npc[np.asarray(j) * nx + np.asarray(i)]

        xs = np.linspace(0.0, lx, nx + 1)
        ys = np.linspace(0.0, ly, ny + 1)
        X, Y = np.meshgrid(xs, ys)  # (ny+1, nx+1)
        x = np.zeros((nnode, 2))
        x[npn, 0] = X.reshape(-1)
        x[npn, 1] = Y.reshape(-1)
        # cells -> nodes
        I, J = np.meshgrid(np.arange(nx), np.arange(ny))
        I, J = I.reshape(-1), J.reshape(-1)
        pcell = np.zeros((ncell, 4), dtype=np.int64)
        pcell[cell(I, J)] = np.stack([node(I, J), node(I + 1, J), node(I + 1, J + 1),
                                      node(I, J + 1)], axis=1)
        # interior edges: vertical interfaces (left/right cells), then horizontal
        vi, vj = np.meshgrid(np.arange(nx - 1), np.arange(ny))
        vi, vj = vi.reshape(-1), vj.reshape(-1)
        ev_n = np.stack([node(vi + 1, vj + 1), node(vi + 1, vj)], axis=1)
        ev_c = np.stack([cell(vi, vj), cell(vi + 1, vj)], axis=1)
        hi_, hj = np.meshgrid(np.arange(nx), np.arange(ny - 1))
        hi_, hj = hi_.reshape(-1), hj.reshape(-1)
        eh_n = np.stack([node(hi_, hj + 1), node(hi_ + 1, hj + 1)], axis=1)
        eh_c = np.stack([cell(hi_, hj), cell(hi_, hj + 1)], axis=1)
        en = np.concatenate([ev_n, eh_n])
        ec = np.concatenate([ev_c, eh_c])
        nedge = en.shape[0]
        pe = _perm(rng, nedge, permute)
        pedge = np.zeros_like(en)
        pecell = np.zeros_like(ec)
        pedge[pe] = en
        pecell[pe] = ec
        # boundary edges: bottom, top, left, right (outward normals)
        ib = np.arange(nx)
        jb = np.arange(ny)
        b_n = np.concatenate([
            np.stack([node(ib + 1, 0 * ib), node(ib, 0 * ib)], axis=1),
            np.stack([node(ib, 0 * ib + ny), node(ib + 1, 0 * ib + ny)], axis=1),
            np.stack([node(0 * jb, jb), node(0 * jb, jb + 1)], axis=1),
            np.stack([node(0 * jb + nx, jb + 1), node(0 * jb + nx, jb)], axis=1)])
        b_c = np.concatenate([cell(ib, 0 * ib), cell(ib, 0 * ib + ny - 1),
                              cell(0 * jb, jb), cell(0 * jb + nx - 1, jb)])
        nb = b_n.shape[0]
        pb = _perm(rng, nb, permute)
        pbedge = np.zeros_like(b_n)
        pbecell = np.zeros(nb, dtype=np.int64)
        pbedge[pb] = b_n
        pbecell[pb] = b_c
        self.x = x
        self.pcell = pcell.astype(np.int32)
        self.pedge = pedge.astype(np.int32)
        self.pecell = pecell.astype(np.int32)
        self.pbedge = pbedge.astype(np.int32)
        self.pbecell = pbecell.reshape(-1, 1).astype(np.int32)
        self.bound = rng.integers(0, 2, size=nb).astype(np.int32).reshape(-1, 1)
        self.nnode, self.ncell, self.nedge, self.nbedge = nnode, ncell, nedge, nb


class TriMesh(object):
    """Unit square with n x n quads, each split into 2 triangles along one
    diagonal: 2n^2 elements, (n+1)^2 vertices, vertex and element numbering
    permuted.  ``elem_node`` (int32, counter-clockwise) and ``coords``."""

    def __init__(self, n, seed=0, permute=True):
        rng = np.random.default_rng(seed)
        nv = (n + 1) * (n + 1)
        pv = _perm(rng, nv, permute)
        xs = np.linspace(0.0, 1.0, n + 1)
        X, Y = np.meshgrid(xs, xs)
        coords = np.zeros((nv, 2))
        coords[pv, 0] = X.reshape(-1)
        coords[pv, 1] = Y.reshape(-1)
        I, J = np.meshgrid(np.arange(n), np.arange(n))
        I, J = I.reshape(-1), J.reshape(-1)
        v = lambda i, j: pv[j * (n + 1) + i]
        t1 = np.stack([v(I, J), v(I + 1, J), v(I + 1, J + 1)], axis=1)
        t2 = np.stack([v(I, J), v(I + 1, J + 1), v(I, J + 1)], axis=1)
        el = np.concatenate([t1, t2])
        pe = _perm(rng, el.shape[0], permute)
        elem_node = np.zeros_like(el)
        elem_node[pe] = el
        self.coords = coords
        self.elem_node = elem_node.astype(np.int32)
        self.nelem, self.nvert = el.shape[0], nv


class TetMesh(object):
    """Unit cube with n^3 hexes, each split into the 6 Kuhn tetrahedra that
    share the main diagonal: 6n^3 elements, (n+1)^3 vertices, numbering
    permuted.  ``elem_node`` (int32, positively oriented) and ``coords``."""

    KUHN = [(0, 1, 3, 7), (0, 1, 5, 7), (0, 2, 3, 7), (0, 2, 6, 7), (0, 4, 5, 7), (0, 4, 6, 7)]

    def __init__(self, n, seed=0, permute=True):
        rng = np.random.default_rng(seed)
        n1 = n + 1
        nv = n1 ** 3
        pv = _perm(rng, nv, permute)
        g = np.linspace(0.0, 1.0, n1)
        Z, Y, X = np.meshgrid(g, g, g, indexing='ij')
        coords = np.zeros((nv, 3))
        coords[pv, 0] = X.reshape(-1)
        coords[pv, 1] = Y.reshape(-1)
        coords[pv, 2] = Z.reshape(-1)
        K, J, I = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing='ij')
        I, J, K = I.reshape(-1), J.reshape(-1), K.reshape(-1)
        corner = [pv[(K + (c >> 2 & 1)) * n1 * n1 + (J + (c >> 1 & 1)) * n1 + (I + (c & 1))]
                  for c in range(8)]
        tets = []
        for t in self.KUHN:
            tets.append(np.stack([corner[c] for c in t], axis=1))
        el = np.concatenate(tets)
        # orient positively
        p = coords[el]
        d = np.einsum('ij,ij->i', np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0]), p[:, 3] - p[:, 0])
        neg = d < 0
        el[neg, 2], el[neg, 3] = el[neg, 3].copy(), el[neg, 2].copy()
        pe = _perm(rng, el.shape[0], permute)
        elem_node = np.zeros_like(el)
        elem_node[pe] = el
        self.coords = coords
        self.elem_node = elem_node.astype(np.int32)
        self.nelem, self.nvert = el.shape[0], nv


def _triangle_table(path):
    """Rows of a Triangle .node/.ele file after its header: comments ('#' to
    end of line) and blank lines dropped."""
    with open(path) as h:
        lines = [ln.split('#', 1)[0].split() for ln in h]
    lines = [ln for ln in lines if ln]
    return [int(t) for t in lines[0][:3]], lines[1:]


def read_triangle(op2, prefix):
    """Read the Triangle mesh ``prefix``.node / ``prefix``.ele into op2
    objects, as the reference's demos do (demo/triangle_reader.py:39-89):
    1-based ids, x and y of each node, the first three nodes of each
    triangle; attributes and boundary markers ignored.  Returns
    ``(nodes, vnodes, coords, elements, elem_node, elem_vnode)``."""
    head, rows = _triangle_table(prefix + '.node')
    num_nodes = head[0]
    if len(rows) != num_nodes:
        raise ValueError("%s.node: header says %d nodes, %d given" % (prefix, num_nodes, len(rows)))
    ids = np.array([int(r[0]) for r in rows]) - 1
    xy = np.array([[float(r[1]), float(r[2])] for r in rows])
    coord_values = np.zeros((num_nodes, 2))
    coord_values[ids] = xy
    nodes = op2.Set(num_nodes, 1, "nodes")
    vnodes = op2.Set(num_nodes, 2, "vnodes")
    coords = op2.Dat(vnodes, coord_values, np.float64, "coords")

    (num_tri, nodes_per_tri, _), rows = _triangle_table(prefix + '.ele')
    if nodes_per_tri != 3:
        raise ValueError("%s.ele: %d nodes per triangle; linear triangles (3) only"
                         % (prefix, nodes_per_tri))
    if len(rows) != num_tri:
        raise ValueError("%s.ele: header says %d triangles, %d given" % (prefix, num_tri, len(rows)))
    tri = np.array([int(r[0]) for r in rows]) - 1
    map_values = np.zeros((num_tri, 3), dtype=np.int32)
    map_values[tri] = np.array([[int(v) - 1 for v in r[1:4]] for r in rows])
    elements = op2.Set(num_tri, 1, "elements")
    elem_node = op2.Map(elements, nodes, 3, map_values, "elem_node")
    elem_vnode = op2.Map(elements, vnodes, 3, map_values, "elem_vnode")
    return nodes, vnodes, coords, elements, elem_node, elem_vnode
