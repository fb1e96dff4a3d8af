"""Time the device Krylov solvers on one large 2-D system.

    python3 -m pyop2_amd.workloads.cgbench [elements] [iterations]

Builds a seeded triangle mesh of about ``elements`` elements (default 2M),
assembles A = M + K and b = M f with the batching layer's loops, then runs
``iterations`` (default 300) iterations (tolerance 1e-30, so none stop
early; each timed solve after an untimed one of the same size) of
Jacobi-PCG, GMRES(30) and BiCGStab, and prints one JSON line per
method: time per iteration and the algorithmic bytes of an iteration over
that time.  Bytes per iteration (n rows, nnz entries, 8-byte values, 4-byte
indices):
  CG        12 nnz + 108 n  (as fused: SpMV with the direction formed where
            gathered, and one update pass)
  GMRES(m)  12 nnz + 28 n for the SpMV with D^-1, then 8 n (j + 2) for the
            Gram-Schmidt dots and 8 n (j + 3) for the update of step j;
            averaged over j = 0 .. m-1
  BiCGStab  24 nnz + 168 n  (direction 32 n, two SpMV 12 nnz + 36 n and
            12 nnz + 44 n, update 56 n)
"""

import json
import sys
import time


def run(op2, ne=2000000, its=300):
    """Time the three solvers on one assembled system (``op2`` initialised
    with the hip backend); returns one result dict per method."""
    import torch
    from pyop2_amd import batch
    from pyop2_amd.workloads import suites
    mesh = suites.tri_mesh_of(ne, 1)
    B = batch.MeshBatch(op2, [mesh], name='cgbench')
    B.assemble()
    B.solve(rtol=1e-10, maxit=20)
    torch.cuda.synchronize()
    t = time.perf_counter()
    r = B.solve(rtol=1e-30, maxit=its)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t
    n, nnz, done = mesh.nvert, B.sparsity.nz, r[0][0]
    out = [report('cg', mesh, n, nnz, done, dt, 12 * nnz + 108 * n)]
    m = 30
    gm = 12 * nnz + 28 * n + sum(8 * n * (j + 2) + 8 * n * (j + 3) for j in range(m)) / m
    for method, per_it in (('gmres', gm), ('bcgs', 24 * nnz + 168 * n)):
        solver = op2.Solver(linear_solver=method, relative_tolerance=1e-30,
                            maximum_iterations=its, error_on_nonconvergence=False,
                            gmres_restart=m)
        # a first solve with the same arguments: the timed one replays its
        # captured round graphs
        solver.solve(B.A, B.x, B.b)
        torch.cuda.synchronize()
        t = time.perf_counter()
        solver.solve(B.A, B.x, B.b)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t
        out.append(report(method, mesh, n, nnz, solver.getIterationNumber(), dt, per_it))
    return out


def report(method, mesh, n, nnz, done, dt, per_it):
    return {'method': method, 'elements': mesh.nelem, 'rows': n, 'nnz': nnz,
            'iterations': done, 'us_per_iteration': round(dt / done * 1e6, 2),
            'bytes_per_iteration': int(per_it),
            'GBps': round(per_it * done / dt / 1e9, 1)}


def main(argv):
    from pyop2_amd import op2
    ne = int(float(argv[0])) if argv else 2000000
    its = int(argv[1]) if len(argv) > 1 else 300
    op2.init(backend='hip')
    for r in run(op2, ne, its):
        print(json.dumps(r), flush=True)


if __name__ == '__main__':
    main(sys.argv[1:])
