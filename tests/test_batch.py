"""Batched assembly + CG of many meshes (gpu): every mesh converges, its
solution is bit-identical solved alone and in the batch, and it matches the
CPU oracle's assembly + CG within the solver tolerance."""

import numpy as np
import pytest

from oracle import sequential as seq
from pyop2_amd.batch import MeshBatch
from pyop2_amd.workloads import fem, meshes

pytestmark = pytest.mark.gpu


def make_meshes(sizes, seed=0):
    rng = np.random.default_rng(seed)
    ms = [meshes.TriMesh(n, seed=int(rng.integers(1 << 30))) for n in sizes]
    fs = [rng.random(m.nvert) for m in ms]
    ks = [rng.lognormal(0.0, 1.0, m.nvert) for m in ms]
    return ms, fs, ks


def test_batch_alone_vs_mix_bit_identical(hip):
    ms, fs, ks = make_meshes([5, 17, 3, 30, 11, 8])
    mix = MeshBatch(hip, ms, fs, ks, name='mix')
    mix.assemble()
    out = mix.solve(rtol=1e-10)
    assert all(r > 0 for _, _, r in out), out
    for k in (0, 3, 5):
        alone = MeshBatch(hip, [ms[k]], [fs[k]], [ks[k]], name='alone%d' % k)
        alone.assemble()
        o1 = alone.solve(rtol=1e-10)
        assert o1[0][0] == out[k][0]          # same iteration count
        assert np.array_equal(alone.solution(0), mix.solution(k))


def test_batch_matches_oracle(hip):
    ms, fs, ks = make_meshes([12, 20], seed=4)
    mix = MeshBatch(hip, ms, fs, ks, name='mo')
    mix.assemble()
    mix.solve(rtol=1e-12)
    for k, m in enumerate(ms):
        W = seq
        nodes = W.Set(m.nvert, 1)
        vnodes = W.Set(m.nvert, 2)
        el = W.Set(m.nelem, 1)
        en = W.Map(el, nodes, 3, m.elem_node)
        ev = W.Map(el, vnodes, 3, m.elem_node)
        A = W.Mat(W.Sparsity((en, en)), np.float64)
        x = W.Dat(vnodes, m.coords.copy(), np.float64)
        kap = W.Dat(nodes, ks[k].copy(), np.float64)
        f = W.Dat(nodes, fs[k].copy(), np.float64)
        b = W.Dat(nodes, np.zeros(m.nvert), np.float64)
        u = W.Dat(nodes, np.zeros(m.nvert), np.float64)
        W.par_loop(W.Kernel(fem.MASS_TRI, 'mass_tri'), el(3, 3), A((en[W.i[0]], en[W.i[1]]), W.INC),
                   x(ev, W.READ))
        W.par_loop(W.Kernel(fem.LAPLACE_TRI, 'laplace_tri'), el(3, 3),
                   A((en[W.i[0]], en[W.i[1]]), W.INC), x(ev, W.READ), kap(en, W.READ))
        W.par_loop(W.Kernel(fem.RHS_TRI, 'rhs_tri'), el(3), b(en[W.i[0]], W.INC), x(ev, W.READ),
                   f(en, W.READ))
        W.Solver(relative_tolerance=1e-12, maximum_iterations=10000).solve(A, u, b)
        ref = u.data_ro.reshape(-1)
        got = mix.solution(k)
        assert np.abs(got - ref).max() <= 1e-9 * np.abs(ref).max()


def test_batch_graph_steps_match_eager(hip):
    """Replayed hipGraph steps (re-captured whenever finished systems leave
    the launches) give the eager steps' solutions bit for bit and the same
    iteration counts, also on a second solve that reuses the batch."""
    ms, fs, ks = make_meshes([9, 25, 4, 14], seed=7)
    mix = MeshBatch(hip, ms, fs, ks, name='gm')
    mix.assemble()
    eager = mix.solve(rtol=1e-10, chunk=8, graph=False)
    xe = [mix.solution(k).copy() for k in range(len(ms))]
    for _ in range(2):
        got = mix.solve(rtol=1e-10, chunk=8, graph=True)
        assert [o[0] for o in got] == [o[0] for o in eager]
        assert [o[2] for o in got] == [o[2] for o in eager]
        for k in range(len(ms)):
            assert np.array_equal(mix.solution(k), xe[k])
