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


def spread(vals, npos):
    """Mean distance between the first and other columns of map rows."""
    v = vals.astype(np.int64)
    return np.abs(v[:, 1:] - v[:, :1]).mean()


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
    # locality: neighbouring cells of an edge are close in device numbering
    assert spread(layout.map_values(pecell), mesh.ncell) < 0.05 * spread(mesh.pecell, mesh.ncell)


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
