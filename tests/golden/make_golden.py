"""Write the golden fixtures of tests/test_golden.py.

``oracle_small.npz``: small seeded cases run through the CPU oracle (the
restated sequential backend, registered as 'sequential'):
  * airfoil on a 16 x 8 quad mesh (seed 3, q0 seed 2), two time steps:
    q, qold, adt, res after them and rms of each step;
  * the mass2d pattern on an 8 x 8 triangle mesh (seed 1, f seed 7): the
    Sparsity CSR, the mass matrix values (``Mat.array``: CSR order, user
    numbering), b = M f and the CG
    solution of M x = b (rtol 1e-12);
  * the 3-D assembly on a 3^3 Kuhn tet mesh (seed 4, kappa seed 5): CSR of
    the scalar and vector patterns and the Laplace and vector-mass values.
``reference_fixtures.json`` holds values the reference's own tests state
(test/unit/test_matrices.py:54-62, 579-608, 765-778); it is written by
hand from those lines, not by this script.

Run from the repository root: ``python tests/golden/make_golden.py``.
"""

import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

OUT = os.path.join(ROOT, 'tests', 'golden', 'oracle_small.npz')


def cases(W):
    """The golden cases run in world ``W`` (an op2-like backend module);
    returns a dict of arrays."""
    from pyop2_amd.workloads import airfoil, fem, meshes
    out = {}
    mesh = meshes.QuadMesh(16, 8, seed=3)
    q0 = airfoil.initial_q(mesh.ncell, seed=2)
    A = airfoil.Airfoil(W, mesh, q0=q0)
    try:
        rms = []
        for _ in range(2):
            A.rms.data = 0.0
            A.iteration()
            rms.append(float(A.rms.data[0]))
        out['airfoil_q'] = np.array(A.p_q.data_ro)
        out['airfoil_qold'] = np.array(A.p_qold.data_ro)
        out['airfoil_adt'] = np.array(A.p_adt.data_ro)
        out['airfoil_res'] = np.array(A.p_res.data_ro)
        out['airfoil_rms'] = np.array(rms)
    finally:
        A.remove_consts()
    tri = meshes.TriMesh(8, seed=1)
    f = np.random.default_rng(7).random(tri.nvert)
    m = fem.mass2d(W, tri, f, solve_rtol=1e-12, name='g2d')
    out['mass2d_rowptr'] = m['rowptr']
    out['mass2d_colidx'] = m['colidx']
    out['mass2d_values'] = m['mat']
    out['mass2d_b'] = m['b']
    out['mass2d_x'] = m['x']
    tet = meshes.TetMesh(3, seed=4)
    kappa = np.random.default_rng(5).lognormal(0.0, 1.0, tet.nvert)
    a = fem.Assembly3D(W, tet, kappa, name='g3d')
    a.laplace()
    a.vector_mass()
    for tag, sp, mat in (('laplace', a.sp_scalar, a.A), ('vmass', a.sp_vector, a.M)):
        out[tag + '_rowptr'] = np.array(sp._rowptr)
        out[tag + '_colidx'] = np.array(sp._colidx)
        out[tag + '_values'] = np.array(mat.array)
    return out


def main():
    from oracle import sequential
    from pyop2_amd import op2
    sequential.register()
    op2.init(backend='sequential')
    np.savez_compressed(OUT, **cases(sequential))
    print('wrote', OUT, os.path.getsize(OUT), 'bytes')


if __name__ == '__main__':
    main()
