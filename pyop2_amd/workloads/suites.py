"""Batching and serving measurements (used by bench.py).

``batch_suite``: meshes with element counts log-uniform in a range, each
assembled (A = M + K(kappa), b = M f) and solved with CG to ``rtol``; the
batched path (one MeshBatch, one launch per loop, one batched CG) against
the plain path (one mesh at a time through op2 par_loops and
Solver.solve).  ``serve_suite``: Poisson request streams on a resident
pool, the continuous-batching server against the sequential plain path.
"""

import time

import numpy as np

from . import meshes


def tri_mesh_of(nelem, seed):
    n = max(2, int(round(np.sqrt(nelem / 2.0))))
    return meshes.TriMesh(n, seed=seed)


def tet_mesh_of(nelem, seed):
    n = max(2, int(round((nelem / 6.0) ** (1.0 / 3.0))))
    return meshes.TetMesh(n, seed=seed)


def batch_suite(op2, nmesh=64, emin=2000, emax=200000, rtol=1e-10, seed=0, rounds=2,
                progress=None):
    """``progress(msg)``, when given, is called as each phase (and every 32
    meshes of the one-at-a-time path) completes."""
    import torch
    from ..batch import MeshBatch
    say = progress or (lambda msg: None)
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(np.log(emin), np.log(emax), nmesh))
    ms = [tri_mesh_of(s, int(rng.integers(1 << 30))) for s in sizes]
    fs = [rng.random(m.nvert) for m in ms]
    ks = [rng.lognormal(0.0, 1.0, m.nvert) for m in ms]
    t = time.perf_counter()
    say('meshes generated: %d elements' % sum(m.nelem for m in ms))
    mix = MeshBatch(op2, ms, fs, ks, name='bsuite')
    mix.assemble()
    mix.solve(rtol=rtol)
    torch.cuda.synchronize()
    setup_b = time.perf_counter() - t
    say('batched warm-up done (%.1f s)' % setup_b)
    t = time.perf_counter()
    for _ in range(rounds):
        mix.assemble()
        out = mix.solve(rtol=rtol)
    torch.cuda.synchronize()
    tb = (time.perf_counter() - t) / rounds
    say('batched timed: %.3f s per round' % tb)
    assert all(r > 0 for _, _, r in out)
    singles = [MeshBatch(op2, [m], [f], [k], name='bs%d' % i)
               for i, (m, f, k) in enumerate(zip(ms, fs, ks))]
    t = time.perf_counter()
    for i, s1 in enumerate(singles):
        s1.assemble()
        s1.solve(rtol=rtol, chunk=16, graph=False)
        if (i + 1) % 32 == 0:
            say('one-at-a-time warm-up: %d meshes' % (i + 1))
    torch.cuda.synchronize()
    setup_s = time.perf_counter() - t
    t = time.perf_counter()
    for _ in range(rounds):
        for i, s1 in enumerate(singles):
            s1.assemble()
            s1.solve(rtol=rtol, chunk=16, graph=False)
            if (i + 1) % 32 == 0:
                say('one-at-a-time timed: %d meshes' % (i + 1))
    torch.cuda.synchronize()
    ts = (time.perf_counter() - t) / rounds
    same = all(np.array_equal(s1.solution(0), mix.solution(k)) for k, s1 in enumerate(singles))
    return {'meshes': nmesh, 'elements': int(sum(m.nelem for m in ms)),
            'element_range': [int(emin), int(emax)], 'rtol': rtol,
            'max_iterations': int(max(o[0] for o in out)),
            'batched_solves_per_s': round(nmesh / tb, 2),
            'one_at_a_time_solves_per_s': round(nmesh / ts, 2),
            'speedup': round(ts / tb, 2), 'alone_equals_mix_bitwise': bool(same),
            'setup_s': round(setup_b, 2)}


def make_pool(op2, n2d=8, n3d=8, tri_range=(8000, 500000), tet_range=(50000, 2000000), seed=0):
    from ..serve import ResidentMesh
    rng = np.random.default_rng(seed)
    pool = []
    for k in range(n2d):
        e = np.exp(rng.uniform(np.log(tri_range[0]), np.log(tri_range[1])))
        pool.append(ResidentMesh(op2, tri_mesh_of(e, int(rng.integers(1 << 30))), 'pt%d' % k))
    for k in range(n3d):
        e = np.exp(rng.uniform(np.log(tet_range[0]), np.log(tet_range[1])))
        pool.append(ResidentMesh(op2, tet_mesh_of(e, int(rng.integers(1 << 30))), 'pk%d' % k))
    return pool


def serve_suite(op2, pool, rates=(100, 1000, 5000), nreq=64, budget_gb=64.0, max_active=64,
                iters=16, rtol=1e-10, seed=0, progress=None):
    """``progress(msg)``, when given, is called as each phase completes."""
    import torch
    from .. import serve
    say = progress or (lambda msg: None)
    out = {'pool': len(pool), 'pool_elements': int(sum(m.mesh.nelem for m in pool)),
           'rtol': rtol, 'budget_GB': budget_gb, 'rates': {}}
    # warm every mesh (compile, plans, slots)
    warm = [serve.Request(k, 0.0, k, 12345 + k) for k in range(len(pool))]
    serve.Server(pool, budget_gb * 1e9, max_active, iters, rtol).run(warm)
    say('served warm-up done')
    serve.serve_sequential(pool, [serve.Request(k, 0.0, k, 1 + k) for k in range(len(pool))],
                           rtol=rtol)
    torch.cuda.synchronize()
    say('sequential warm-up done')
    for lam in rates:
        reqs = serve.poisson_requests(nreq, lam, len(pool), seed=seed + lam)
        server = serve.Server(pool, budget_gb * 1e9, max_active, iters, rtol)
        server.run(reqs)
        a = serve.latency_stats(reqs)
        say('rate %g: served %s' % (lam, a))
        base = serve.poisson_requests(nreq, lam, len(pool), seed=seed + lam)
        serve.serve_sequential(pool, base, rtol=rtol)
        b = serve.latency_stats(base)
        say('rate %g: sequential %s' % (lam, b))
        out['rates'][str(lam)] = {
            'served': {k: round(v, 3) for k, v in a.items()},
            'sequential': {k: round(v, 3) for k, v in b.items()},
            'converged': all(r.reason > 0 for r in reqs)}
    return out
