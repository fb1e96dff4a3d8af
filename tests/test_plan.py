"""Plan invariants (host-built, no GPU needed).

The reference pins its plan by the conflict-freedom invariant
(test/unit/test_coloring.py:104-112) and structural properties
(plan.pyx:107-110, 374-376): every element in exactly one block, ncolblk
sums to nblocks, blkmap is a permutation sorted by colour, no two elements
of one (block, thread colour) and no two blocks of one colour share an INC
target.  These are checked here with numpy, independently of the native
verifier, on random maps (hypothesis) and on permuted airfoil meshes."""

import numpy as np
import pytest
from hypothesis import given, settings, strategies as st

from pyop2_amd.hip import Plan
from pyop2_amd.workloads import meshes


def check_plan(p, maps, n):
    """maps: list of (values[n, dim], idx, space) race maps."""
    assert sorted(p.elems.tolist()) == list(range(n))
    assert p.nelems.sum() == n
    assert (p.offset == np.concatenate([[0], np.cumsum(p.nelems)[:-1]])).all()
    assert p.ncolblk.sum() == p.nblocks
    assert sorted(p.blkmap.tolist()) == list(range(p.nblocks))
    assert (np.diff(p.blkcol[p.blkmap]) >= 0).all()
    # per-element target lists
    def targets(e):
        out = set()
        for v, idx, space in maps:
            cols = [idx] if idx >= 0 else range(v.shape[1])
            for c in cols:
                out.add((space, int(v[e, c])))
        return out
    owner_blockcol = {}
    for b in range(p.nblocks):
        els = p.elems[p.offset[b]:p.offset[b] + p.nelems[b]]
        cols = p.thrcol[p.offset[b]:p.offset[b] + p.nelems[b]]
        assert cols.max() < p.nthrcol[b] if len(cols) else True
        seen = {}
        blk_targets = set()
        for e, c in zip(els, cols):
            for t in targets(e):
                assert seen.get((t, c)) in (None, e), "thread colour conflict"
                seen[(t, c)] = e
                blk_targets.add(t)
        for t in blk_targets:
            key = (t, p.blkcol[b])
            assert owner_blockcol.get(key) in (None, b), "block colour conflict"
            owner_blockcol[key] = b
    assert p.verify()


@settings(max_examples=25, deadline=None)
@given(n=st.integers(1, 600), ntarget=st.integers(1, 80), dim=st.integers(1, 4),
       ps=st.sampled_from([8, 32, 64, 256]), reorder=st.sampled_from([0, 1, 2]),
       seed=st.integers(0, 1000))
def test_random_maps(n, ntarget, dim, ps, reorder, seed):
    rng = np.random.default_rng(seed)
    v = rng.integers(0, ntarget, size=(n, dim)).astype(np.int32)
    race = [(v, dim, -1, 'a', ntarget)]
    p = Plan(n, ps, reorder, race)
    check_plan(p, [(v, -1, 'a')], n)


def test_two_spaces_and_index():
    rng = np.random.default_rng(3)
    n = 2000
    a = rng.integers(0, 300, size=(n, 2)).astype(np.int32)
    b = rng.integers(0, 50, size=(n, 3)).astype(np.int32)
    race = [(a, 2, 0, 's1', 300), (a, 2, 1, 's1', 300), (b, 3, -1, 's2', 50)]
    p = Plan(n, 128, True, race)
    check_plan(p, [(a, 0, 's1'), (a, 1, 's1'), (b, -1, 's2')], n)


def test_all_elements_one_target():
    """Every element increments one target: one element per thread colour."""
    n = 300
    v = np.zeros((n, 1), dtype=np.int32)
    p = Plan(n, 64, True, [(v, 1, 0, 'u', 1)])
    check_plan(p, [(v, 0, 'u')], n)
    assert p.ncolors == p.nblocks
    assert (p.nthrcol == p.nelems).all()


def test_element_colouring():
    """reorder = 2: elements coloured one by one, each colour's run cut into
    blocks; every block has one thread colour, blocks follow colour order,
    elements ascend within a colour, and the plan is conflict-free."""
    m = meshes.QuadMesh(120, 60, seed=5)
    race = [(m.pecell, 2, 0, 'cells', m.ncell), (m.pecell, 2, 1, 'cells', m.ncell)]
    p = Plan(m.nedge, 256, 2, race)
    check_plan(p, [(m.pecell, 0, 'cells'), (m.pecell, 1, 'cells')], m.nedge)
    assert (p.nthrcol == 1).all() and (p.thrcol == 0).all()
    assert p.ncolors <= 8
    assert np.array_equal(p.blkmap, np.arange(p.nblocks))
    assert (np.diff(p.blkcol) >= 0).all()
    for c in range(p.ncolors):
        blocks = np.nonzero(p.blkcol == c)[0]
        els = np.concatenate([p.elems[p.offset[b]:p.offset[b] + p.nelems[b]] for b in blocks])
        assert (np.diff(els) > 0).all()
    assert np.array_equal(np.sort(p.elems), np.arange(m.nedge))


@pytest.mark.parametrize('reorder', [True, False])
def test_airfoil_edges(reorder):
    m = meshes.QuadMesh(120, 60, seed=5)
    race = [(m.pecell, 2, 0, 'cells', m.ncell), (m.pecell, 2, 1, 'cells', m.ncell)]
    p = Plan(m.nedge, 256, reorder, race)
    check_plan(p, [(m.pecell, 0, 'cells'), (m.pecell, 1, 'cells')], m.nedge)
    if reorder:
        # compact regions: few block colours, few thread colours
        assert p.ncolors <= 8
        assert p.nthrcol.max() <= 8
    else:
        # contiguous blocks of a permuted numbering conflict with most others
        assert p.ncolors > 20


# -- owner-computes plan --------------------------------------------------------

from pyop2_amd.hip import OwnerPlan  # noqa: E402


def check_owner(p, cols, n, ntarget):
    """cols: list of (values[n, dim], idx).  Every (element, race column)
    increment appears once, in its target's list, in ascending element
    order; owned ranges are disjoint; each element is primary once."""
    colv = []
    for v, idx in cols:
        colv += [v[:, idx]] if idx >= 0 else [v[:, c] for c in range(v.shape[1])]
    nac = len(colv)
    assert p.nac == nac
    seen = set()
    primary = np.zeros(n, dtype=int)
    owned = np.zeros(ntarget, dtype=int)
    for b in range(p.nactive):
        t0, t1 = p.blk_t0[b], p.blk_t1[b]
        owned[t0:t1] += 1
        els, prim = p.block_elements(b)
        assert len(np.unique(els)) == len(els)
        rn = int(p.rng_n[b])
        # a run of consecutive primaries, then the listed elements ascending
        assert (np.diff(els[:rn]) == 1).all() and prim[:rn].all()
        assert (np.diff(els[rn:]) > 0).all()
        assert len(els) <= p.max_exec
        primary[els[prim]] += 1
        for t in range(t0, t1):
            prev = -1
            for c in p.contrib[p.tp[t]:p.tp[t + 1]]:
                slot, ac = divmod(int(c), nac)
                e = int(els[slot])
                assert colv[ac][e] == t
                assert e >= prev
                prev = e
                seen.add((e, ac))
    assert owned.max() <= 1
    assert (primary == 1).all()
    touched = np.concatenate([p.tgt[p.tg_off[b]:p.tg_off[b + 1]] for b in range(p.nactive)]) \
        if p.nactive else np.zeros(0, dtype=np.int32)
    assert np.array_equal(np.sort(touched), np.nonzero(np.diff(p.tp))[0])
    assert len(seen) == n * nac


@settings(max_examples=20, deadline=None)
@given(n=st.integers(1, 500), ntarget=st.integers(1, 120), dim=st.integers(1, 4),
       k=st.sampled_from([1, 4, 32, 256]), seed=st.integers(0, 1000))
def test_owner_random(n, ntarget, dim, k, seed):
    rng = np.random.default_rng(seed)
    v = rng.integers(0, ntarget, size=(n, dim)).astype(np.int32)
    cols = [(v, -1)] if seed % 2 else [(v, 0)] + ([(v, dim - 1)] if dim > 1 else [])
    p = OwnerPlan(n, ntarget, [(c[0], dim, c[1]) for c in cols], k)
    check_owner(p, cols, n, ntarget)


def test_owner_airfoil_edges():
    m = meshes.QuadMesh(60, 30, seed=5)
    cols = [(m.pecell, 0), (m.pecell, 1)]
    p = OwnerPlan(m.nedge, m.ncell, [(v, 2, i) for v, i in cols], 64)
    check_owner(p, cols, m.nedge, m.ncell)


def test_owner_primary_runs():
    """Edges sorted by their first cell: every block's primary edges are one
    run (no list entries for them), and the plan stays valid."""
    m = meshes.QuadMesh(60, 30, seed=5)
    order = np.argsort(m.pecell[:, 0], kind='stable')
    pec = np.ascontiguousarray(m.pecell[order])
    cols = [(pec, 0), (pec, 1)]
    p = OwnerPlan(m.nedge, m.ncell, [(v, 2, i) for v, i in cols], 64)
    check_owner(p, cols, m.nedge, m.ncell)
    assert (p.rng_n > 0).all()
    assert p.rng_n.sum() == m.nedge
    assert ((p.exec.view(np.uint32) >> 31) == 1).all()


def _node_csr(elem_node, nnodes, rd, cd):
    """Scalar CSR of the node pattern of one map pair, expanded rd x cd
    (sorted unique columns), by brute force."""
    nbrs = [set() for _ in range(nnodes)]
    for row in elem_node:
        for a in row:
            nbrs[a].update(int(b) for b in row)
    rows, cols = [0], []
    for n in range(nnodes):
        cs = sorted(nbrs[n])
        for r in range(rd):
            cols += [cd * m + c for m in cs for c in range(cd)]
            rows.append(len(cols))
    return np.array(rows, dtype=np.int32), np.array(cols, dtype=np.int32)


@pytest.mark.parametrize('rd,cd,k', [(1, 1, 5), (2, 2, 7), (1, 1, 1)])
def test_mat_scatter_program(rd, cd, k):
    """The owner Mat scatter program lists, for every node pair, exactly the
    (element, row position, column position) entries that add into it, in
    ascending user element order then column position, and every entry of
    every element once (brute force on a permuted triangle mesh)."""
    from pyop2_amd.hip import OwnerPlan, mat_program_host
    mesh = meshes.TriMesh(7, seed=11)
    en = mesh.elem_node.astype(np.int32)
    n, nn = en.shape[0], mesh.nvert
    order = np.random.default_rng(2).permutation(n).astype(np.int32)   # user order
    p = OwnerPlan(n, nn, [(en, 3, -1)], k, order)
    rp, ci = _node_csr(en, nn, rd, cd)
    poff, prog, wide = mat_program_host(p, en, 3, nn, rp, ci, rd, cd)
    assert not wide
    rank = np.empty(n, dtype=np.int64)
    rank[order] = np.arange(n)
    seen = set()
    npairs = len(poff) - 1
    assert npairs == rp[nn * rd] // (rd * cd)
    for b in range(p.nactive):
        cbase = p.tp[p.blk_t0[b]]
        for node in range(p.blk_t0[b], p.blk_t1[b]):
            s0 = rp[node * rd]
            cols = ci[s0:rp[node * rd + 1]:cd] // cd
            for kk, m in enumerate(cols):
                pp = s0 // (rd * cd) + kk
                got = []
                for code in prog[poff[pp]:poff[pp + 1]]:
                    c, j = divmod(int(code), 3)
                    slot, i0 = divmod(int(p.contrib[cbase + c]), p.nac)
                    e = int(p.block_elements(b)[0][slot])
                    assert en[e, i0] == node and en[e, j] == m
                    got.append((rank[e], j))
                    seen.add((e, i0, j))
                assert got == sorted(got)
    assert len(seen) == n * 9
