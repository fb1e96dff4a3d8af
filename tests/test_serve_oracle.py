"""Served requests against the oracle (gpu).

A mixed pool of 2-D triangle and 3-D tetrahedral meshes is served with
continuous batching; each request (a mesh and a seeded lognormal(0, 1)
coefficient field) is also assembled and solved by the CPU oracle, the
reference's sequential assembly with its restated PETSc KSPCG + PCJACOBI
(pyop2/petsc_base.py:182-220; rtol 1e-10, atol 1e-50, dtol 1e4).  The
served solve must stop for the same reason, within one iteration of the
oracle's count (reductions are summed in another order, so the last
residual may cross the tolerance a step apart), with a solution within
1e-9 of the oracle's, normwise.
"""

import numpy as np
import pytest

from oracle import sequential as seq
from pyop2_amd import serve
from pyop2_amd.workloads import meshes

pytestmark = pytest.mark.gpu

RTOL = 1e-10


def oracle_solve(rm, kappa):
    """The request's assembly (A = M + K(kappa), b = M f with f = 1) and
    CG solve on the oracle, on oracle objects of the same mesh ``rm``."""
    W = seq
    nv = rm.dim + 1
    k = W.Dat(rm.nodes, kappa.copy(), np.float64)
    A = W.Mat(rm.sparsity, np.float64)
    b = W.Dat(rm.nodes, np.zeros(rm.n), np.float64)
    x = W.Dat(rm.nodes, np.zeros(rm.n), np.float64)
    W.par_loop(rm.k_mass, rm.elems(nv, nv), A((rm.en[W.i[0]], rm.en[W.i[1]]), W.INC),
               rm.coords(rm.ex, W.READ))
    W.par_loop(rm.k_lap, rm.elems(nv, nv), A((rm.en[W.i[0]], rm.en[W.i[1]]), W.INC),
               rm.coords(rm.ex, W.READ), k(rm.en, W.READ))
    W.par_loop(rm.k_rhs, rm.elems(nv), b(rm.en[W.i[0]], W.INC),
               rm.coords(rm.ex, W.READ), rm.f(rm.en, W.READ))
    s = W.Solver(relative_tolerance=RTOL, absolute_tolerance=1e-50, divergence_tolerance=1e4,
                 maximum_iterations=20000, error_on_nonconvergence=False)
    s.solve(A, x, b)
    return np.array(x.data_ro).reshape(-1), s.getIterationNumber(), s.getConvergedReason()


def test_served_requests_match_oracle(hip):
    ms = [meshes.TriMesh(24, seed=11), meshes.TriMesh(41, seed=12), meshes.TriMesh(57, seed=13),
          meshes.TetMesh(7, seed=14), meshes.TetMesh(11, seed=15)]
    pool = [serve.ResidentMesh(hip, m, 'so%d' % k) for k, m in enumerate(ms)]
    opool = [serve.ResidentMesh(seq, m, 'oo%d' % k) for k, m in enumerate(ms)]
    reqs = serve.poisson_requests(15, rate=2000.0, pool_size=len(pool), seed=21)
    assert len(set(r.mesh_id for r in reqs)) == len(pool)
    budget = 4 * max(m.slot_bytes for m in pool)
    server = serve.Server(pool, budget, max_active=8, iters=8, rtol=RTOL, keep_solutions=True)
    server.run(reqs)
    for r in reqs:
        rm = opool[r.mesh_id]
        xo, its, reason = oracle_solve(rm, serve.kappa_of(r, rm.n))
        assert reason > 0 and r.reason == reason, (r.rid, r.reason, reason)
        assert abs(r.its - its) <= 1, (r.rid, r.its, its)
        err = np.linalg.norm(r.solution - xo) / np.linalg.norm(xo)
        assert err <= 1e-9, (r.rid, err)
