"""Device data layout (no GPU needed): orderings are permutations, maps are
translated consistently, locality improves on permuted meshes, and a CSR
built in device numbering translates back to the user-numbered CSR of the
oracle bit-exactly."""

import numpy as np
import pytest

from oracle import sequential as seq
from pyop2_amd import hip, layout
from pyop2_amd.workloads import meshes


def airfoil_sets(mesh):
    nodes = hip.Set(mesh.nnode, 2, 'nodes')
    edges = hip.Set(mesh.nedge, 1, 'edges')
    cells = hip.Set(mesh.ncell, 1, 'cells')
    vcells = hip.Set(mesh.ncell, 4, 'vcells')
    pedge = hip.Map(edges, nodes, 2, mesh.pedge, 'pedge')
    pecell = hip.Map(edges, cells, 2, mesh.pecell, 'pecell')
    pevcell = hip.Map(edges, vcells, 2, mesh.pecell, 'pevcell')
    pcell = hip.Map(cells, nodes, 4, mesh.pcell, 'pcell')
    return nodes, edges, cells, vcells, pedge, pecell, pevcell, pcell


def same_block(vals, b):
    """Fraction of map rows whose entries all fall in one aligned range of
    ``b`` positions."""
    v = vals.astype(np.int64) // b
    return (v == v[:, :1]).all(axis=1).mean()


def test_layout_permutations_and_translation():
    mesh = meshes.QuadMesh(60, 30, seed=4)
    nodes, edges, cells, vcells, pedge, pecell, pevcell, pcell = airfoil_sets(mesh)
    layout.link(vcells, cells)
    for s in (cells, vcells, nodes, edges):
        lay = layout.of(s)
        assert not lay.identity
        assert sorted(lay.perm.tolist()) == list(range(s.total_size))
        assert (lay.iperm[lay.perm] == np.arange(s.total_size)).all()
    assert np.array_equal(layout.of(cells).perm, layout.of(vcells).perm)
    for m in (pedge, pecell, pevcell, pcell):
        dev = layout.map_values(m)
        src, dst = layout.of(m.iterset), layout.of(m.dataset)
        # device row i is user element perm[i]; its entries are device
        # positions of the user targets
        assert np.array_equal(dst.perm[dev], m.values[src.perm])
    # locality: the two cells of most edges share an aligned range of 32
    # device positions (a block that owns the range runs the edge once)
    dev = same_block(layout.map_values(pecell), 32)
    assert dev > 0.5 and dev > 10 * same_block(mesh.pecell, 32)


@pytest.mark.parametrize('how', ['bisect', 'bfs'])
def test_adjacency_order_is_a_permutation(how):
    mesh = meshes.TriMesh(20, seed=3)
    perm = layout.adjacency_order(mesh.nvert, mesh.elem_node, how)
    assert sorted(perm.tolist()) == list(range(mesh.nvert))
    # targets no element names still get positions (isolated components)
    extra = layout.adjacency_order(mesh.nvert + 7, mesh.elem_node, how)
    assert sorted(extra.tolist()) == list(range(mesh.nvert + 7))


def test_bisect_order_compact_ranges():
    """Aligned ranges of the bisection order are compact patches: on a
    1200 x 600 quad mesh an owner block of 128 cells runs far fewer
    edges twice than with a breadth-first order."""
    mesh = meshes.QuadMesh(400, 200, seed=5)

    def shared_edges(how, b):
        perm = layout.adjacency_order(mesh.ncell, mesh.pecell, how)
        iperm = np.empty_like(perm)
        iperm[perm] = np.arange(len(perm))
        return 1.0 - same_block(iperm[mesh.pecell], b)

    for b in (64, 128, 256):
        assert shared_edges('bisect', b) < 0.6 * shared_edges('bfs', b)


def test_bisect_order_independent_of_threads(monkeypatch):
    """The ordering is the same whatever the host thread count."""
    mesh = meshes.TetMesh(24, seed=2)
    out = []
    for nt in ('1', '3', '8'):
        monkeypatch.setenv('OP2HIP_HOST_THREADS', nt)
        out.append(layout.adjacency_order(mesh.nvert, mesh.elem_node, 'bisect'))
    assert all(np.array_equal(out[0], o) for o in out[1:])


@pytest.mark.parametrize('dims', [(1, 1), (2, 2)])
def test_csr_translation_bit_exact(dims):
    mesh = meshes.TriMesh(12, seed=8)
    rng = np.random.default_rng(1)
    n = mesh.nvert
    perm = rng.permutation(n)
    lay = layout.Layout(perm)
    # user-numbered pattern by the oracle's std::set restatement
    S1 = seq.Set(mesh.nelem)
    S2 = seq.Set(n, dims[0])
    m_user = seq.Map(S1, S2, 3, mesh.elem_node)
    sp_user = seq.Sparsity((m_user, m_user))
    # device numbering: targets renumbered by iperm
    m_dev = seq.Map(S1, S2, 3, lay.iperm[mesh.elem_node].astype(np.int32))
    sp_dev = seq.Sparsity((m_dev, m_dev))
    rp, ci, nnz, slot = layout.csr_to_user(sp_dev._rowptr, sp_dev._colidx, dims[0], dims[1],
                                           lay, lay)
    assert np.array_equal(rp, sp_user._rowptr)
    assert np.array_equal(ci, sp_user._colidx)
    assert np.array_equal(nnz, sp_user._d_nnz)
    assert sorted(slot.tolist()) == list(range(len(ci)))


@pytest.mark.parametrize('how', ['bisect', 'bfs'])
def test_components_ordered_as_alone(how):
    """Disconnected meshes in one Set (a batch of meshes) keep contiguous
    position ranges, each ordered exactly as that mesh alone: the batched
    solvers rely on it for alone-vs-batch bit identity."""
    ms = [meshes.TriMesh(n, seed=s) for n, s in ((9, 1), (23, 2), (4, 3), (40, 4))]
    off = np.cumsum([0] + [m.nvert for m in ms])
    union = np.concatenate([m.elem_node + o for m, o in zip(ms, off[:-1])])
    perm = layout.adjacency_order(int(off[-1]), union, how)
    for k, m in enumerate(ms):
        alone = layout.adjacency_order(m.nvert, m.elem_node, how)
        assert np.array_equal(perm[off[k]:off[k + 1]] - off[k], alone)


def test_identical_maps_share_target_layout():
    """Element-to-node maps with identical values onto scalar and vector
    node Sets put both Sets in one layout, so their device maps coincide."""
    mesh = meshes.TriMesh(15, seed=2)
    el = hip.Set(mesh.nelem, 1, 'el')
    nodes = hip.Set(mesh.nvert, 1, 'nodes')
    vnodes = hip.Set(mesh.nvert, 2, 'vnodes')
    en = hip.Map(el, nodes, 3, mesh.elem_node, 'en')
    ev = hip.Map(el, vnodes, 3, mesh.elem_node, 'ev')
    assert np.array_equal(layout.of(nodes).perm, layout.of(vnodes).perm)
    assert np.array_equal(layout.map_values(en), layout.map_values(ev))


@pytest.mark.parametrize('bad', [-1, 5, 1 << 20])
def test_map_values_out_of_range_raise(bad):
    """A map entry outside its target Set raises MapValueError at its first
    device use (layout decision or upload), before any host permutation
    could wrap a negative entry to the last target or a gather could leave
    the target's device array."""
    from pyop2_amd.exceptions import MapValueError
    src = hip.Set(4, 1, 'rc_src')
    dst = hip.Set(5, 1, 'rc_dst')
    vals = np.array([[0, 1], [1, 2], [2, 3], [3, 4]], dtype=np.int32)
    vals[2, 1] = bad
    m = hip.Map(src, dst, 2, vals, 'rc_map')
    with pytest.raises(MapValueError):
        layout.map_values(m)
    with pytest.raises(MapValueError):
        layout.of(dst)


def test_map_values_in_range_pass():
    src = hip.Set(3, 1, 'ok_src')
    dst = hip.Set(4, 1, 'ok_dst')
    m = hip.Map(src, dst, 2, np.array([[0, 3], [1, 2], [3, 0]], dtype=np.int32), 'ok_map')
    dev = layout.map_values(m)
    assert np.array_equal(layout.of(dst).perm[dev], m.values[layout.of(src).perm]
                          if not layout.of(src).identity else m.values)
