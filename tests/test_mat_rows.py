"""Owner Mat assembly by row blocks (the "rows" template).

Plan: the row-block plan built on the device (op2hip_mat_rows_build) equals,
array for array, a plain construction from its definition: per row node its
records (element, row position) in ascending (user element, row position)
order, coded slot * e0 + i0 with slot the element's place in its block's
exec list; per block the elements naming one of its nodes in ascending user
order; per record the position of each column node in the row.  Checked on
scalar and 2x2-block patterns, identity and shuffled user orders, blocks
of 8 to 64 nodes.

Assembly: P1 mass and Laplace matrices on triangles and tetrahedra, a
vector mass, a Mat WRITE and an element naming one node twice are
bit-exact to the oracle's sequential loop (mat_utils.cxx:4-28 addto order)
and to the owner template; the rows template is the one that ran."""

import numpy as np
import pytest

from oracle import sequential as seq
from pyop2_amd.workloads import fem, meshes


def _expected(en_r, en_c, rank, nnodes, k, rp, ci, rd, cd):
    n, e0 = en_r.shape
    e1 = en_c.shape[1]
    user = np.argsort(rank, kind='stable')
    recs = [[] for _ in range(nnodes)]
    for e in user:
        for i0 in range(e0):
            recs[en_r[e, i0]].append((e, i0))
    nblk = (nnodes + k - 1) // k
    exe = []
    for b in range(nblk):
        els = {e for t in range(b * k, min(nnodes, (b + 1) * k)) for e, _ in recs[t]}
        exe.append(sorted(els, key=lambda e: rank[e]))
    tp = np.zeros(nnodes + 1, dtype=np.int64)
    tp[1:] = np.cumsum([len(r) for r in recs])
    codes, pos = [], []
    W = (e1 + 3) // 4
    for t in range(nnodes):
        slot = {e: s for s, e in enumerate(exe[t // k])}
        cols = list(ci[rp[t * rd]:rp[t * rd + 1]:cd] // cd)
        for e, i0 in recs[t]:
            codes.append(slot[e] * e0 + i0)
            words = [0] * W
            for j in range(e1):
                words[j // 4] |= cols.index(en_c[e, j]) << (8 * (j % 4))
            pos += words
    xoff = np.zeros(nblk + 1, dtype=np.int64)
    xoff[1:] = np.cumsum([len(x) for x in exe])
    return tp, xoff, [np.array(x, dtype=np.int64) for x in exe], np.array(codes), np.array(pos)


@pytest.mark.gpu
@pytest.mark.parametrize('dim,k,shuffle', [(1, 8, False), (1, 13, True), (2, 29, True),
                                           (1, 64, True)])
def test_rows_plan_matches_definition(hip, dim, k, shuffle):
    from pyop2_amd import hip as H, layout
    mesh = meshes.TetMesh(4, seed=2)
    en = mesh.elem_node.astype(np.int32)
    n, nv = en.shape[0], mesh.nvert
    el = hip.Set(n, 1)
    nodes = hip.Set(nv, dim)
    m = hip.Map(el, nodes, 4, en)
    layout._group(el).layout = layout.IDENTITY
    layout._group(nodes).layout = layout.IDENTITY
    sp = hip.Sparsity((m, m))
    A = hip.Mat(sp, np.float64)
    rank = (np.random.default_rng(k).permutation(n) if shuffle else np.arange(n)).astype(np.int32)
    arg = A((m[hip.i[0]], m[hip.i[1]]), hip.INC)
    plan = H.RowsMatPlan(n, arg, 4, 4, k, H.torch.from_numpy(rank).to(H.device()))
    rp = sp._dev_rowptr.cpu().numpy().astype(np.int64)
    ci = sp._dev_colidx[:sp.nz].cpu().numpy().astype(np.int64)
    tp, xoff, exe, codes, pos = _expected(en, en, rank, nv, k, rp, ci, dim, dim)
    assert plan.fits and plan.nblk == len(exe)
    assert np.array_equal(plan.host('tp'), tp)
    assert np.array_equal(plan.host('xoff'), xoff)
    ex = plan.host('exec')
    for b in range(plan.nblk):
        assert np.array_equal(ex[xoff[b]:xoff[b + 1]], exe[b])
    assert np.array_equal(plan.host('code'), codes)
    assert np.array_equal(plan.host('pos'), pos)
    assert plan.max_code == codes.max()
    assert plan.max_exec == max(len(x) for x in exe)
    assert plan.sum_exec == xoff[-1]
    slices = [rp[min(nv, (b + 1) * k) * dim] - rp[b * k * dim] for b in range(plan.nblk)]
    assert plan.max_slice == max(slices)


@pytest.mark.gpu
@pytest.mark.skip(reason="an earlier hand-written device build of the staging lists faulted with "
                         "a memory aperture violation on its first GPU runs, cause not established; "
                         "it was removed and this test of its layout is kept skipped. The lists are "
                         "now built by csrc/staging.hip (test_rows_stage_lists_definition, "
                         "test_rows_stage_plan, tests/test_staging.py)")
@pytest.mark.parametrize('k', [8, 29])
def test_rows_staging_lists(hip, k):
    """Per block the sorted distinct targets its elements name through a
    map (the reference's ind_map) and per exec slot their positions in that
    list (its loc_map)."""
    from pyop2_amd import hip as H, layout
    mesh = meshes.TetMesh(4, seed=5)
    en = mesh.elem_node.astype(np.int32)
    n, nv = en.shape[0], mesh.nvert
    el = hip.Set(n, 1)
    nodes = hip.Set(nv, 1)
    m = hip.Map(el, nodes, 4, en)
    layout._group(el).layout = layout.IDENTITY
    layout._group(nodes).layout = layout.IDENTITY
    A = hip.Mat(hip.Sparsity((m, m)), np.float64)
    rank = np.random.default_rng(k).permutation(n).astype(np.int32)
    plan = H.RowsMatPlan(n, A((m[hip.i[0]], m[hip.i[1]]), hip.INC), 4, 4, k,
                         H.torch.from_numpy(rank).to(H.device()))
    sn, sc, xl, smax = plan.stage(m)
    sn, sc = sn.cpu().numpy(), sc.cpu().numpy()
    xl = xl.cpu().numpy().view(np.uint16)
    xoff, ex = plan.host('xoff'), plan.host('exec')
    longest = 0
    for b in range(plan.nblk):
        els = ex[xoff[b]:xoff[b + 1]]
        want = np.unique(en[els].ravel())
        got = sn[xoff[b] * 4:xoff[b] * 4 + sc[b]]
        assert np.array_equal(got, want)
        loc = xl[xoff[b] * 4:xoff[b + 1] * 4].reshape(-1, 4)
        assert np.array_equal(want[loc], en[els])
        longest = max(longest, len(want))
    assert smax == max(2, (longest + 1) & ~1)
    # the padding past the last slot names position 0
    assert (xl[xoff[-1] * 4:] == 0).all()


def _mass_tri(W, mesh, mode, kernel=fem.MASS_TRI, name='mass_tri'):
    nodes = W.Set(mesh.nvert)
    vn = W.Set(mesh.nvert, 2)
    el = W.Set(mesh.nelem)
    en = W.Map(el, nodes, 3, mesh.elem_node)
    ev = W.Map(el, vn, 3, mesh.elem_node)
    x = W.Dat(vn, mesh.coords.copy(), np.float64)
    A = W.Mat(W.Sparsity((en, en)), np.float64)
    W.par_loop(W.Kernel(kernel, name), el(3, 3), A((en[W.i[0]], en[W.i[1]]), mode),
               x(ev, W.READ))
    return np.array(A.values)


def _loop_modes(hip):
    from pyop2_amd import hip as H
    return {e.loop.source.mode for e in H._call_cache.values() if e.loop is not None}


@pytest.mark.gpu
@pytest.mark.parametrize('opts', [{}, {'hip_mat_rows_k': 16}, {'hip_mat_rows_block_size': 128},
                                  {'hip_mat_rows_k': 8, 'hip_mat_rows_block_size': 64}])
def test_rows_mass_tri_bitexact(hip, opts):
    from pyop2_amd import configuration as cfg, hip as H
    mesh = meshes.TriMesh(40, seed=7)
    ref = _mass_tri(seq, mesh, seq.INC)
    H._call_cache.clear()
    with cfg.override(**opts):
        got = _mass_tri(hip, mesh, hip.INC)
        assert 'rows' in _loop_modes(hip)
    np.testing.assert_array_equal(got, ref)


@pytest.mark.gpu
def test_rows_laplace_and_vector_mass_bitexact(hip):
    """Tetrahedra: the Laplace matrix through the rows template, bit-exact
    to the oracle and to the owner template; the vector mass (whose local
    matrices do not fit the rows template's LDS) bit-exact too."""
    from pyop2_amd import configuration as cfg
    mesh = meshes.TetMesh(9, seed=3)
    kappa = np.random.default_rng(4).lognormal(0, 1, mesh.nvert)
    o = fem.Assembly3D(seq, mesh, kappa, name='o3r')
    o.laplace()
    o.vector_mass()
    out = {}
    for rows in (True, False):
        with cfg.override(hip_mat_rows=rows):
            a = fem.Assembly3D(hip, mesh, kappa, name='h3r%d' % rows)
            a.laplace()
            a.vector_mass()
            out[rows] = (np.array(a.A.array), np.array(a.M.array))
    for rows in (True, False):
        np.testing.assert_array_equal(out[rows][0], o.A.array)
        np.testing.assert_array_equal(out[rows][1], o.M.array)


@pytest.mark.gpu
def test_rows_insert_and_repeated_node(hip):
    """Mat WRITE keeps the last element's value (sequential order), and an
    element naming one node at two positions adds both terms in order."""
    mesh = meshes.TriMesh(11, seed=9)
    code = """void k(double A[1][1], double *x[2], int i, int j) {
      A[0][0] = x[i][0] * 10.0 + x[j][1] + (double)(i * 3 + j); }"""
    ref = _mass_tri(seq, mesh, seq.WRITE, code, 'k')
    np.testing.assert_array_equal(_mass_tri(hip, mesh, hip.WRITE, code, 'k'), ref)
    en = mesh.elem_node.copy()
    en[::7, 2] = en[::7, 0]
    inc = """void k2(double A[1][1], double *x[2], int i, int j) {
      A[0][0] += x[i][0] * 0.3 + x[j][1] * 0.7 + 0.1 * (double)(i + 2 * j); }"""
    out = []
    for W in (seq, hip):
        nodes = W.Set(mesh.nvert)
        vn = W.Set(mesh.nvert, 2)
        el = W.Set(mesh.nelem)
        m = W.Map(el, nodes, 3, en)
        mv = W.Map(el, vn, 3, en)
        x = W.Dat(vn, mesh.coords.copy(), np.float64)
        A = W.Mat(W.Sparsity((m, m)), np.float64)
        W.par_loop(W.Kernel(inc, 'k2'), el(3, 3), A((m[W.i[0]], m[W.i[1]]), W.INC),
                   x(mv, W.READ))
        out.append(np.array(A.values))
    np.testing.assert_array_equal(out[1], out[0])


@pytest.mark.gpu
def test_rows_stage_lists_definition(hip):
    """The staging lists on the device (csrc/staging.hip): per block the
    sorted distinct targets of its slots' map rows, per slot the position of
    its target in that list, every position below the block's list length,
    the padding zero; blocks of 0 to 2000 elements of arity 1 to 4, targets
    drawn from few or many values."""
    from pyop2_amd import hip as H
    torch = H.torch
    rng = np.random.default_rng(3)
    most = H.nat.lib.op2hip_stage_max_keys()
    assert most >= 4096
    for nb, dim, hi, cmax in ((1, 4, 500, 30), (7, 3, 500, 30), (40, 4, 500, 30),
                              (5, 1, 3, 2000), (3, 4, 10 ** 6, most // 4)):
        counts = rng.integers(0, cmax, nb)
        counts[nb // 2] = 0
        xoff = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
        S = int(xoff[-1])
        rows = rng.integers(0, hi, (S, dim)).astype(np.int32)
        dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(H.device())
        st = H.stage_lists(dev(xoff), dev(np.arange(max(S, 1), dtype=np.int32)),
                           dev(rows.reshape(-1) if S else np.zeros(1, np.int32)), dim, S * dim,
                           int(counts.max()) * dim, 16)
        sn, so, xl, smax = (t.cpu().numpy() if hasattr(t, 'cpu') else t for t in st[:4])
        xl = xl.view(np.uint16)
        longest = 0
        for b in range(nb):
            want = np.unique(rows[xoff[b]:xoff[b + 1]].ravel())
            assert np.array_equal(sn[so[b]:so[b + 1]], want)
            loc = xl[xoff[b] * dim:xoff[b + 1] * dim].reshape(-1, dim)
            assert (loc < max(len(want), 1)).all()
            assert np.array_equal(want[loc], rows[xoff[b]:xoff[b + 1]])
            longest = max(longest, len(want))
        assert smax == max(longest, 2)
        assert len(xl) >= (S + 16) * dim and (xl[S * dim:] == 0).all()


@pytest.mark.gpu
def test_rows_stage_plan(hip):
    """A row-block plan's staging lists on the device: per block the
    distinct targets of its exec slots through the map, per slot their
    positions (uint16) in that list."""
    from pyop2_amd import hip as H, layout
    mesh = meshes.TetMesh(5, seed=8)
    en = mesh.elem_node.astype(np.int32)
    n, nv = en.shape[0], mesh.nvert
    el = hip.Set(n, 1)
    nodes = hip.Set(nv, 1)
    m = hip.Map(el, nodes, 4, en)
    layout._group(el).layout = layout.IDENTITY
    layout._group(nodes).layout = layout.IDENTITY
    A = hip.Mat(hip.Sparsity((m, m)), np.float64)
    rank = np.random.default_rng(1).permutation(n).astype(np.int32)
    plan = H.RowsMatPlan(n, A((m[hip.i[0]], m[hip.i[1]]), hip.INC), 4, 4, 16,
                         H.torch.from_numpy(rank).to(H.device()))
    sn, so, xl, smax = plan.stage(m)
    sn, so = sn.cpu().numpy(), so.cpu().numpy()
    xl = xl.cpu().numpy().view(np.uint16)
    xoff, ex = plan.host('xoff'), plan.host('exec')
    for b in range(plan.nblk):
        els = ex[xoff[b]:xoff[b + 1]]
        want = np.unique(en[els].ravel())
        assert np.array_equal(sn[so[b]:so[b + 1]], want)
        loc = xl[xoff[b] * 4:xoff[b + 1] * 4].reshape(-1, 4)
        assert np.array_equal(want[loc], en[els])
        assert len(want) <= smax
    assert (xl[xoff[-1] * 4:] == 0).all()


@pytest.mark.gpu
def test_rows_staged_assembly_bitexact(hip):
    """The rows template with its elements' rows staged in LDS per block:
    P1 mass on triangles (two block shapes), the Laplace and vector mass on
    tetrahedra and an element naming one node twice, bit-exact to the
    oracle; the staged template is the one that ran."""
    from pyop2_amd import configuration as cfg, hip as H
    mesh = meshes.TriMesh(40, seed=7)
    ref = _mass_tri(seq, mesh, seq.INC)
    for opts in ({}, {'hip_mat_rows_k': 16}):
        H._call_cache.clear()
        with cfg.override(hip_mat_rows_stage=True, **opts):
            got = _mass_tri(hip, mesh, hip.INC)
            srcs = [e.loop.source for e in H._call_cache.values() if e.loop is not None]
            assert any(s.mode == 'rows' and s.rows_stage for s in srcs)
        np.testing.assert_array_equal(got, ref)
    mesh3 = meshes.TetMesh(9, seed=3)
    kappa = np.random.default_rng(4).lognormal(0, 1, mesh3.nvert)
    o = fem.Assembly3D(seq, mesh3, kappa, name='o3s')
    o.laplace()
    o.vector_mass()
    with cfg.override(hip_mat_rows_stage=True):
        a = fem.Assembly3D(hip, mesh3, kappa, name='h3s')
        a.laplace()
        a.vector_mass()
        np.testing.assert_array_equal(np.array(a.A.array), o.A.array)
        np.testing.assert_array_equal(np.array(a.M.array), o.M.array)
        en = mesh.elem_node.copy()
        en[::7, 2] = en[::7, 0]
        inc = """void k3(double A[1][1], double *x[2], int i, int j) {
          A[0][0] += x[i][0] * 0.3 + x[j][1] * 0.7 + 0.1 * (double)(i + 2 * j); }"""
        out = []
        for W in (seq, hip):
            nodes = W.Set(mesh.nvert)
            vn = W.Set(mesh.nvert, 2)
            el = W.Set(mesh.nelem)
            m = W.Map(el, nodes, 3, en)
            mv = W.Map(el, vn, 3, en)
            x = W.Dat(vn, mesh.coords.copy(), np.float64)
            A = W.Mat(W.Sparsity((m, m)), np.float64)
            W.par_loop(W.Kernel(inc, 'k3'), el(3, 3), A((m[W.i[0]], m[W.i[1]]), W.INC),
                       x(mv, W.READ))
            out.append(np.array(A.values))
        np.testing.assert_array_equal(out[1], out[0])


@pytest.mark.gpu
@pytest.mark.parametrize('stage', [False, True])
@pytest.mark.parametrize('nel', [1, 37])
def test_rows_edge_sets(hip, stage, nel):
    """Row blocks on edge-case sets: one element, and a node set with nodes
    no element names (empty rows, and blocks of them only); Mat values
    bit-exact to the oracle, with and without staged rows.  (A Sparsity
    needs populated maps, so an empty element set cannot be assembled, as
    in the reference: base.py:1164-1168.)"""
    from pyop2_amd import configuration as cfg
    mesh = meshes.TriMesh(6, seed=5)
    en = mesh.elem_node[:nel].astype(np.int32)
    nv = mesh.nvert + 70          # 70 nodes no element names
    coords = np.vstack([mesh.coords, np.random.default_rng(2).random((70, 2))])
    out = []
    for W in (seq, hip):
        nodes = W.Set(nv)
        vn = W.Set(nv, 2)
        el = W.Set(nel)
        m = W.Map(el, nodes, 3, en.reshape(-1, 3))
        mv = W.Map(el, vn, 3, en.reshape(-1, 3))
        x = W.Dat(vn, coords.copy(), np.float64)
        A = W.Mat(W.Sparsity((m, m)), np.float64)
        with cfg.override(hip_mat_rows_stage=stage):
            W.par_loop(W.Kernel(fem.MASS_TRI, 'mass_tri'), el(3, 3),
                       A((m[W.i[0]], m[W.i[1]]), W.INC), x(mv, W.READ))
        out.append(np.array(A.values))
    np.testing.assert_array_equal(out[1], out[0])
