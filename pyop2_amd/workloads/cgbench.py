"""Time the device CG on one large 2-D system.

    python3 -m pyop2_amd.workloads.cgbench [elements] [iterations]

Builds a seeded triangle mesh of about ``elements`` elements (default 2M),
assembles A = M + K and b = M f with the batching layer's loops, then runs
``iterations`` (default 300) Jacobi-PCG iterations (tolerance 1e-30, so
none stop early) and prints one JSON line: time per iteration and the
bytes one iteration moves (12 nnz + 108 n) over that time.
"""

import json
import sys
import time


def main(argv):
    import torch
    from pyop2_amd import op2, batch
    from pyop2_amd.workloads import suites
    ne = int(float(argv[0])) if argv else 2000000
    its = int(argv[1]) if len(argv) > 1 else 300
    op2.init(backend='hip')
    mesh = suites.tri_mesh_of(ne, 1)
    B = batch.MeshBatch(op2, [mesh])
    B.assemble()
    B.solve(rtol=1e-10, maxit=20)
    torch.cuda.synchronize()
    t = time.perf_counter()
    r = B.solve(rtol=1e-30, maxit=its)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t
    n, nnz, done = mesh.nvert, B.sparsity.nz, r[0][0]
    per_it = 12 * nnz + 108 * n
    print(json.dumps({'elements': mesh.nelem, 'rows': n, 'nnz': nnz, 'iterations': done,
                      'us_per_iteration': round(dt / done * 1e6, 2),
                      'bytes_per_iteration': per_it,
                      'GBps': round(per_it * done / dt / 1e9, 1)}), flush=True)


if __name__ == '__main__':
    main(sys.argv[1:])
