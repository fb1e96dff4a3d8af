"""Device Sparsity build against the oracle on patterns that exercise each
path of the build (gpu): rows of up to 64, 128 and 256 candidates (the
wavefront paths), a fan vertex whose row has thousands of candidates (the
workgroup hash path), a fan with more distinct columns than the hash table
holds (the minimum-search path), several map pairs of different arities,
rectangular patterns, 2x2 blocks and rows no element names.  rowptr and
colidx must equal the oracle's restatement of the reference's
std::set construction (pyop2/sparsity_utils.cxx:5-51) bit for bit."""

import numpy as np
import pytest

from oracle import sequential as seq

pytestmark = pytest.mark.gpu


def fan(n, hub=0):
    """Triangles (hub, k, k+1), k = 1..n, around one hub vertex."""
    k = np.arange(1, n + 1)
    return np.stack([np.full(n, hub), k, k + 1], axis=1).astype(np.int32), n + 2


def star_rows(nvert, deg, arity, seed):
    """Elements whose first column cycles over a few hubs so each hub row
    gathers about ``deg`` elements, the rest random."""
    rng = np.random.default_rng(seed)
    nel = 4 * deg
    v = rng.integers(0, nvert, size=(nel, arity)).astype(np.int32)
    v[:, 0] = np.arange(nel) % 4
    return v


def build_both(W_list, nrow, ncol, pairs, rdim=1, cdim=1, tag='s'):
    out = []
    for W in W_list:
        rows = W.Set(nrow, rdim, tag + 'r')
        cols = rows if (ncol is None) else W.Set(ncol, cdim, tag + 'c')
        maps = []
        for k, (rv, cv) in enumerate(pairs):
            it = W.Set(len(rv), 1, tag + 'it%d' % k)
            rm = W.Map(it, rows, rv.shape[1], rv, tag + 'rm%d' % k)
            cm = rm if cv is None else W.Map(it, cols, cv.shape[1], cv, tag + 'cm%d' % k)
            maps.append((rm, cm))
        sp = W.Sparsity(tuple(maps) if len(maps) > 1 else maps[0], tag + 'sp')
        out.append((np.array(sp._rowptr), np.array(sp._colidx), np.array(sp._d_nnz), sp.nz))
    return out


def check(h, o):
    assert h[3] == o[3]
    assert np.array_equal(h[0], o[0])
    assert np.array_equal(h[1], o[1])
    assert np.array_equal(h[2], o[2])


@pytest.mark.parametrize('deg,arity', [(20, 3), (40, 3), (80, 3), (60, 4)])
def test_wavefront_rows(hip, deg, arity):
    v = star_rows(500, deg, arity, seed=deg)
    h, o = build_both([hip, seq], 500, None, [(v, None)], tag='w%d%d' % (deg, arity))
    check(h, o)


@pytest.mark.parametrize('n', [100, 2500, 7000])
def test_fan_rows(hip, n):
    """n = 100: 300 candidates (hash path, few values); 2500: 2502
    distinct columns; 7000: 7002 distinct, beyond the hash table's 6144
    (minimum-search path)."""
    t, nv = fan(n)
    h, o = build_both([hip, seq], nv, None, [(t, None)], tag='f%d' % n)
    check(h, o)
    assert h[2][0] == n + 2


def test_pairs_rectangular_blocks_and_empty_rows(hip):
    rng = np.random.default_rng(5)
    nrow, ncol = 300, 170
    r1 = rng.integers(0, nrow - 40, size=(400, 3)).astype(np.int32)   # rows >= 260 unused
    c1 = rng.integers(0, ncol, size=(400, 2)).astype(np.int32)
    r2 = rng.integers(0, nrow - 40, size=(250, 1)).astype(np.int32)
    c2 = rng.integers(0, ncol, size=(250, 4)).astype(np.int32)
    h, o = build_both([hip, seq], nrow, ncol, [(r1, c1), (r2, c2)], rdim=2, cdim=3, tag='pr')
    check(h, o)
    assert (h[2][2 * (nrow - 40):] == 0).all()


def test_rebuild_same_result(hip):
    """Two builds of one pattern from the resident maps agree (the
    incidence lists' order varies with the atomics, the result does not)."""
    from pyop2_amd.workloads import meshes
    m = meshes.TetMesh(9, seed=4)
    nodes = hip.Set(m.nvert, 1, 'rbn')
    tets = hip.Set(m.nelem, 1, 'rbt')
    tn = hip.Map(tets, nodes, 4, m.elem_node, 'rbtn')
    sp = hip.Sparsity((tn, tn), 'rbsp')
    a = sp._build_device()
    b = sp._build_device()
    assert a[3] == b[3] == sp.nz
    assert bool((a[0] == b[0]).all()) and bool((a[1][:a[3]] == b[1][:b[3]]).all())
