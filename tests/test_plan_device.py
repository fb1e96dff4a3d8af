"""The owner-computes plan built on the device against the host build
(gpu): every array entry for entry (tp, contrib, blk_t0/blk_t1, rng_lo/rng_n,
ex_off/exec, tg_off/tgt, tg_cb) and every scalar, on random maps, airfoil
edges, primary runs, element orders given as ranks, all-column (Mat) race
maps, a fan whose hub target has thousands of increments (the long-segment
sort) and blocks of more than 4096 incidences (the block pass in global
memory).  The host build is itself checked against the plan invariants in
test_plan.py (conflict-freedom, each increment once, sequential order)."""

import numpy as np
import pytest

from pyop2_amd.workloads import meshes

pytestmark = pytest.mark.gpu

ARRAYS = ('blk_t0', 'blk_t1', 'rng_lo', 'rng_n', 'ex_off', 'exec', 'tg_off', 'tgt', 'tp',
          'contrib', 'tg_cb')


def build_both(op2, vals, ntargets, colspec, k, order=None):
    """colspec: [(values, idx)]; order: user element order (elem_order of
    the host build) or None."""
    from pyop2_amd import hip
    n = vals[0].shape[0]
    src = op2.Set(n, 1)
    dst = op2.Set(ntargets, 1)
    from pyop2_amd import layout
    cols = []
    for v, idx in colspec:
        m = op2.Map(src, dst, v.shape[1], v)
        cols.append((m, idx))
    # identity device layouts: the device values are the user values
    lay_src, lay_dst = layout._group(src), layout._group(dst)
    lay_src.layout = layout.IDENTITY
    lay_dst.layout = layout.IDENTITY
    rank = None
    if order is not None:
        r = np.empty(n, dtype=np.int32)
        r[order] = np.arange(n, dtype=np.int32)
        rank = hip.torch.from_numpy(r).to(hip.device())
    d = hip.DeviceOwnerPlan(n, ntargets, cols, k, rank)
    h = hip.OwnerPlan(n, ntargets, [(v, v.shape[1], i) for v, i in colspec], k,
                      None if order is None else np.ascontiguousarray(order, dtype=np.int32))
    return d, h


def same(d, h):
    for name in ('nactive', 'max_exec', 'nac', 'max_codes', 'sum_exec'):
        assert getattr(d, name) == getattr(h, name), name
    for name in ARRAYS:
        a, b = np.asarray(getattr(d, name)), np.asarray(getattr(h, name))
        assert a.shape == b.shape and np.array_equal(a, b), name


@pytest.mark.parametrize('seed', range(4))
@pytest.mark.parametrize('k', [1, 7, 64, 300])
def test_random_maps(hip, seed, k):
    rng = np.random.default_rng(seed)
    n, nt, dim = 3000, 900, 3
    v = rng.integers(0, nt, size=(n, dim)).astype(np.int32)
    spec = [(v, -1)] if seed % 2 else [(v, 0), (v, dim - 1)]
    order = rng.permutation(n) if seed >= 2 else None
    same(*build_both(hip, [v], nt, spec, k, order))


def test_airfoil_edges_with_runs(hip):
    m = meshes.QuadMesh(60, 30, seed=5)
    order = np.argsort(m.pecell[:, 0], kind='stable')
    pec = np.ascontiguousarray(m.pecell[order])
    d, h = build_both(hip, [pec], m.ncell, [(pec, 0), (pec, 1)], 64)
    same(d, h)
    assert (d.rng_n > 0).all()


def test_mat_columns_user_order(hip):
    mesh = meshes.TetMesh(8, seed=3)
    en = mesh.elem_node.astype(np.int32)
    order = np.random.default_rng(1).permutation(en.shape[0])
    same(*build_both(hip, [en], mesh.nvert, [(en, -1)], 13, order))


def test_fan_and_large_blocks(hip):
    """A hub target named by 5000 elements (a segment beyond the wavefront
    sort) and blocks of 2000 targets (more incidences than the LDS pass
    holds)."""
    n = 5000
    k_ = np.arange(1, n + 1)
    fan = np.stack([np.zeros(n, dtype=np.int64), k_, k_ + 1], axis=1).astype(np.int32)
    order = np.random.default_rng(4).permutation(n)
    same(*build_both(hip, [fan], n + 2, [(fan, -1)], 5, order))
    same(*build_both(hip, [fan], n + 2, [(fan, -1)], 2000, order))


def test_empty_targets_and_elements(hip):
    v = np.array([[3, 3], [3, 9], [9, 3]], dtype=np.int32)
    same(*build_both(hip, [v], 40, [(v, -1)], 4))
    same(*build_both(hip, [v], 40, [(v, 1)], 64))


@pytest.mark.parametrize('dim,k', [(1, 9), (2, 5), (1, 1)])
def test_mat_program_device_matches_host(hip, dim, k):
    """The owner Mat scatter program built on the device equals the host
    build (poff and every code) for scalar and 2x2-block patterns."""
    from pyop2_amd import hip as H, layout
    mesh = meshes.TetMesh(6, seed=2)
    en = mesh.elem_node.astype(np.int32)
    n, nv = en.shape[0], mesh.nvert
    el = hip.Set(n, 1)
    nodes = hip.Set(nv, dim)
    m = hip.Map(el, nodes, 4, en)
    layout._group(el).layout = layout.IDENTITY
    layout._group(nodes).layout = layout.IDENTITY
    sp = hip.Sparsity((m, m))
    order = np.random.default_rng(k).permutation(n)
    r = np.empty(n, dtype=np.int32)
    r[order] = np.arange(n, dtype=np.int32)
    plan = H.DeviceOwnerPlan(n, nv, [(m, -1)], k, H.torch.from_numpy(r).to(H.device()))

    class _Arg(object):
        map = (m, m)

    poff_d, prog_d, wide = H._mat_program_device(plan, _Arg, 4, sp, dim, dim)
    rp = sp._dev_rowptr.cpu().numpy()
    ci = sp._dev_colidx[:sp.nz].cpu().numpy()
    poff_h, prog_h, wide_h = H.mat_program_host(plan, en, 4, nv, rp, ci, dim, dim)
    assert wide == wide_h
    assert np.array_equal(poff_d.cpu().numpy(), poff_h)
    pd = prog_d.cpu().numpy().view(np.int32 if wide else np.uint16)[:len(prog_h)]
    assert np.array_equal(pd, prog_h)


def _coloured(op2, vals_spec, ntargets_by_space, n, ps, mode):
    """A device coloured plan over hip Maps with identity layouts;
    vals_spec: [(values, idx, space)]."""
    from pyop2_amd import hip as H, layout
    src = op2.Set(n, 1)
    layout._group(src).layout = layout.IDENTITY
    dsts, race = {}, []
    for v, idx, sp in vals_spec:
        if sp not in dsts:
            dsts[sp] = op2.Set(ntargets_by_space[sp], 1)
            layout._group(dsts[sp]).layout = layout.IDENTITY
        m = op2.Map(src, dsts[sp], v.shape[1], v)
        race.append((m, v.shape[1], idx, sp, ntargets_by_space[sp]))
    return H.DevicePlan(n, ps, mode, race)


def _check_coloured(p, n):
    assert p.verify()
    assert sorted(p.elems.tolist()) == list(range(n))
    assert p.nelems.sum() == n and p.ncolblk.sum() == p.nblocks
    assert sorted(p.blkmap.tolist()) == list(range(p.nblocks))
    for b in range(p.nblocks):
        assert p.offset[b] + p.nelems[b] <= n
    # blocks in blkmap order run colour by colour
    assert (np.diff(p.blkcol[p.blkmap]) >= 0).all()


@pytest.mark.parametrize('mode', [0, 2])
@pytest.mark.parametrize('seed', range(3))
def test_device_coloured_plan_random(hip, mode, seed):
    rng = np.random.default_rng(seed)
    n, nt = 6000, 1500
    v = rng.integers(0, nt, size=(n, 3)).astype(np.int32)
    w = rng.integers(0, 700, size=(n, 2)).astype(np.int32)
    spec = [(v, -1, 'a')] if seed == 0 else [(v, 0, 'a'), (v, 2, 'a'), (w, 1, 'b')]
    p = _coloured(hip, spec, {'a': nt, 'b': 700}, n, 128 if seed != 2 else 512, mode)
    _check_coloured(p, n)
    if mode == 2:
        assert (p.nthrcol == 1).all() and (p.thrcol == 0).all()


@pytest.mark.parametrize('mode', [0, 2])
def test_device_coloured_plan_meshes(hip, mode):
    m = meshes.QuadMesh(60, 30, seed=5)
    p = _coloured(hip, [(m.pecell, 0, 'c'), (m.pecell, 1, 'c')], {'c': m.ncell}, m.nedge, 256,
                  mode)
    _check_coloured(p, m.nedge)
    t = meshes.TetMesh(7, seed=3)
    p = _coloured(hip, [(t.elem_node.astype(np.int32), -1, 'n')], {'n': t.nvert}, t.nelem, 256,
                  mode)
    _check_coloured(p, t.nelem)


def test_device_coloured_plan_fan_falls_back(hip):
    """A target shared by 3000 elements needs more element colours than the
    device colouring keeps: the plan is built on the host instead; the
    two-level form colours its blocks and threads on the device."""
    from pyop2_amd import hip as H, layout, configuration as cfg
    n = 3000
    k_ = np.arange(1, n + 1)
    fan = np.stack([np.zeros(n, dtype=np.int64), k_, k_ + 1], axis=1).astype(np.int32)
    p = _coloured(hip, [(fan, -1, 'f')], {'f': n + 2}, n, 256, 0)
    _check_coloured(p, n)
    src = hip.Set(n, 1)
    dst = hip.Set(n + 2, 1)
    layout._group(src).layout = layout.IDENTITY
    layout._group(dst).layout = layout.IDENTITY
    m = hip.Map(src, dst, 3, fan)
    d = hip.Dat(dst, np.zeros(n + 2), np.float64)
    args = [d(m[hip.i[0]], hip.INC)]
    with cfg.override(hip_colouring='elements'):
        q = H.get_plan(src, n, args, part_size=256)
    assert not isinstance(q, H.DevicePlan) and q.verify()
