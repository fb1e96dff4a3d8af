"""Parity and properties at the benchmark's full sizes (gpu).

* Airfoil on the 1200 x 600 mesh (720,000 cells, 1,438,200 edges), the
  benchmark's seed: one full time step (save_soln, then twice adt_calc,
  res_calc, bres_calc, update) on the hip backend and on the oracle from the
  same inputs: q, adt and res bit-exact, the rms Global within 1e-12.
* 128^3 Kuhn tetrahedra (12,582,912 tets, 2,146,689 vertices): nonzero
  counts equal the closed form for the scalar and the 2x2-block pattern,
  every device row holds strictly increasing columns, the P1 Laplace rows
  sum to zero (K 1 = 0) and the vector-P1 mass entries sum to twice the
  cube's volume.  These checks run on the device arrays with torch.
"""

import numpy as np
import pytest

from oracle import sequential as seq
from pyop2_amd.workloads import airfoil, fem, meshes

pytestmark = pytest.mark.gpu


def test_airfoil_1200x600_step_bit_exact(hip):
    mesh = meshes.QuadMesh(1200, 600, seed=0)
    q0 = airfoil.initial_q(mesh.ncell)
    A = airfoil.Airfoil(hip, mesh, q0=q0)
    B = airfoil.Airfoil(seq, mesh, q0=q0, consts=False)
    try:
        A.iteration()
        B.iteration()
        assert np.array_equal(A.p_q.data_ro, B.p_q.data_ro)
        assert np.array_equal(A.p_adt.data_ro, B.p_adt.data_ro)
        assert np.array_equal(A.p_res.data_ro, B.p_res.data_ro)
        assert np.array_equal(A.p_qold.data_ro, B.p_qold.data_ro)
        ra, rb = A.rms.data[0], B.rms.data[0]
        assert abs(ra - rb) <= 1e-12 * abs(rb)
    finally:
        A.remove_consts()


def _rows_of(rowptr):
    import torch
    counts = (rowptr[1:] - rowptr[:-1]).long()
    return torch.repeat_interleave(torch.arange(len(counts), device=rowptr.device), counts)


def test_kuhn128_properties(hip):
    import torch
    n = 128
    mesh = meshes.TetMesh(n, seed=0)
    kappa = np.random.default_rng(1).lognormal(0.0, 1.0, mesh.nvert)
    a = fem.Assembly3D(hip, mesh, kappa, name='k128')
    assert a.sp_scalar.nz == fem.kuhn_nnz(n) == 31802497
    assert a.sp_vector.nz == 4 * fem.kuhn_nnz(n)
    for sp, nrows in ((a.sp_scalar, mesh.nvert), (a.sp_vector, 2 * mesh.nvert)):
        rp = sp._dev_rowptr[:nrows + 1]
        ci = sp._dev_colidx[:sp.nz]
        assert int(rp[0]) == 0 and int(rp[-1]) == sp.nz
        assert bool((rp[1:] >= rp[:-1]).all())
        rows = _rows_of(rp)
        same = rows[1:] == rows[:-1]
        assert bool((ci[1:][same] > ci[:-1][same]).all())       # sorted, unique
        assert int(ci.min()) >= 0 and int(ci.max()) < nrows
    a.laplace()
    a.vector_mass()
    torch.cuda.synchronize()
    a.A._check_err()
    a.M._check_err()
    vals = a.A._dev_vals[:a.sp_scalar.nz]
    rows = _rows_of(a.sp_scalar._dev_rowptr[:mesh.nvert + 1])
    rowsum = torch.zeros(mesh.nvert, dtype=torch.float64, device=vals.device)
    rowsum.index_add_(0, rows, vals)
    assert float(rowsum.abs().max()) <= 1e-12 * float(vals.abs().max())
    msum = float(a.M._dev_vals[:a.sp_vector.nz].sum())
    assert abs(msum - 2.0) <= 1e-10
