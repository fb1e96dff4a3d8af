"""Serving loop (gpu): every request converges, and each result is the
same when served in the continuous batch as when served alone through the
plain op2 path (bit-identical: assembly sums in element order, CG
reductions are per system)."""

import numpy as np
import pytest

from pyop2_amd import serve
from pyop2_amd.workloads import meshes

pytestmark = pytest.mark.gpu


def test_serving_mix_equals_alone(hip):
    pool = [serve.ResidentMesh(hip, meshes.TriMesh(20, seed=1), 'p0'),
            serve.ResidentMesh(hip, meshes.TriMesh(9, seed=2), 'p1'),
            serve.ResidentMesh(hip, meshes.TetMesh(5, seed=3), 'p2'),
            serve.ResidentMesh(hip, meshes.TetMesh(7, seed=4), 'p3')]
    reqs = serve.poisson_requests(16, rate=5000.0, pool_size=len(pool), seed=3)
    budget = 6 * max(m.slot_bytes for m in pool)
    server = serve.Server(pool, budget, max_active=6, iters=8, rtol=1e-10, keep_solutions=True)
    server.run(reqs)
    assert all(r.reason is not None and r.reason > 0 for r in reqs)
    assert all(r.done >= r.arrival for r in reqs)
    alone = serve.poisson_requests(16, rate=5000.0, pool_size=len(pool), seed=3)
    serve.serve_sequential(pool, alone, rtol=1e-10, keep_solutions=True)
    for a, b in zip(reqs, alone):
        assert a.rid == b.rid and a.its == b.its
        assert np.array_equal(a.solution, b.solution)
    st = serve.latency_stats(reqs)
    assert st['p99_ms'] >= st['p50_ms'] > 0


def test_budget_admission(hip):
    """A budget of one slot serialises requests but serves them all."""
    pool = [serve.ResidentMesh(hip, meshes.TriMesh(12, seed=5), 'q0'),
            serve.ResidentMesh(hip, meshes.TriMesh(14, seed=6), 'q1')]
    reqs = serve.poisson_requests(6, rate=1e5, pool_size=2, seed=8)
    server = serve.Server(pool, max(m.slot_bytes for m in pool), max_active=4, iters=8)
    server.run(reqs)
    assert all(r.reason > 0 for r in reqs)
    assert server.used <= max(m.slot_bytes for m in pool)
