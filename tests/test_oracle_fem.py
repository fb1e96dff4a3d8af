"""Pins of the CPU oracle's assembly and solve against closed forms and
brute force (no GPU): sparsity = numpy unique of map-pair keys, the Kuhn
mesh nonzero count, P1 element closed forms, the Laplace patch test, and
CG against scipy's direct solve."""

import numpy as np
import pytest
import scipy.sparse as sps
import scipy.sparse.linalg as spla
from hypothesis import given, settings, strategies as st

from oracle import sequential as seq
from pyop2_amd.workloads import fem, meshes


def csr(mat, sp, n):
    return sps.csr_matrix((mat.array, sp._colidx, sp._rowptr), shape=(n, n))


@settings(max_examples=30, deadline=None)
@given(nel=st.integers(1, 200), nnode=st.integers(1, 60), dim=st.integers(1, 4),
       rmult=st.integers(1, 3), seed=st.integers(0, 10 ** 6))
def test_sparsity_brute_force(nel, nnode, dim, rmult, seed):
    rng = np.random.default_rng(seed)
    vals = rng.integers(0, nnode, size=(nel, dim)).astype(np.int32)
    E = seq.Set(nel)
    N = seq.Set(nnode, rmult)
    m = seq.Map(E, N, dim, vals)
    sp = seq.Sparsity((m, m))
    rows = (rmult * vals[:, :, None, None, None] + np.arange(rmult)[None, None, :, None, None])
    cols = (rmult * vals[:, None, None, :, None] + np.arange(rmult)[None, None, None, None, :])
    r, c = np.broadcast_arrays(rows, cols)
    keys = np.unique(r.astype(np.int64).ravel() * (nnode * rmult) + c.ravel())
    kr, kc = keys // (nnode * rmult), keys % (nnode * rmult)
    rp = np.zeros(nnode * rmult + 1, dtype=np.int64)
    np.add.at(rp, kr + 1, 1)
    assert np.array_equal(sp._rowptr, np.cumsum(rp))
    assert np.array_equal(sp._colidx, kc)


@pytest.mark.parametrize('n', [2, 5, 9])
def test_kuhn_nnz_closed_form(n):
    mesh = meshes.TetMesh(n, seed=n)
    a = fem.Assembly3D(seq, mesh, np.ones(mesh.nvert), name='kn%d' % n)
    assert a.sp_scalar.nz == fem.kuhn_nnz(n)
    assert a.sp_vector.nz == 4 * fem.kuhn_nnz(n)


def test_laplace_and_mass_closed_forms():
    mesh = meshes.TetMesh(6, seed=2)
    n = mesh.nvert
    kappa = np.random.default_rng(0).lognormal(0, 1, n)
    a = fem.Assembly3D(seq, mesh, kappa, name='cf')
    a.laplace()
    a.vector_mass()
    A = csr(a.A, a.sp_scalar, n)
    assert np.abs(A @ np.ones(n)).max() < 1e-12 * np.abs(A).max()   # constants in the kernel
    assert abs(A - A.T).max() == 0.0
    M = sps.csr_matrix((a.M.array, a.sp_vector._colidx, a.sp_vector._rowptr), shape=(2 * n, 2 * n))
    assert abs(M.sum() - 2.0) < 1e-12                      # 2 components x volume 1
    # patch test: constant coefficient, K x = 0 at interior vertices for linear x
    b = fem.Assembly3D(seq, mesh, np.ones(n), name='patch')
    b.laplace()
    K = csr(b.A, b.sp_scalar, n)
    c = mesh.coords
    inner = np.all((c > 1e-12) & (c < 1 - 1e-12), axis=1)
    for d in range(3):
        assert np.abs((K @ c[:, d])[inner]).max() < 1e-13


def test_tri_mass_closed_form_and_solve():
    mesh = meshes.TriMesh(8, seed=1)
    f = np.random.default_rng(4).random(mesh.nvert)
    r = fem.mass2d(seq, mesh, f, solve_rtol=1e-12, name='cfm')
    n = mesh.nvert
    M = sps.csr_matrix((r['mat'], r['colidx'], r['rowptr']), shape=(n, n))
    # assemble the closed form |T|/12 (1 + d_jk) directly
    ref = sps.lil_matrix((n, n))
    for el in mesh.elem_node:
        p = mesh.coords[el]
        u, v = p[1] - p[0], p[2] - p[0]
        area = 0.5 * abs(u[0] * v[1] - u[1] * v[0])
        for j in range(3):
            for k in range(3):
                ref[el[j], el[k]] += area / 12.0 * (2.0 if j == k else 1.0)
    assert abs(M - ref.tocsr()).max() < 1e-15
    assert abs(M.sum() - 1.0) < 1e-13
    np.testing.assert_allclose(r['b'].reshape(-1), M @ f, rtol=1e-13, atol=1e-16)
    x_direct = spla.spsolve(M.tocsc(), r['b'].reshape(-1))
    np.testing.assert_allclose(r['x'].reshape(-1), x_direct, rtol=1e-9)
    np.testing.assert_allclose(r['x'].reshape(-1), f, rtol=1e-9)
