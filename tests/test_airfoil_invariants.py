"""Invariants of the airfoil problem that pin the synthetic mesh generator
(workloads.meshes.QuadMesh) independently of the hip backend.

A uniform state q = qinf is a steady solution of the airfoil scheme
(demo/airfoil_kernels.py:99-165 of the reference): res_calc's flux through
an edge is then F(q) . n with n = (dy, -dx) and the artificial dissipation
term mu (q1 - q2) vanishes, so a cell's residual is F(q) . (sum of its
oriented edge vectors) = 0 for every closed cell; bres_calc's far-field
branch (bound 0) gives the same flux with qinf, so with every boundary edge
far-field every cell's residual vanishes, and with wall edges (bound 1)
every cell away from the boundary still does.  An interior edge whose node
order does not orient its normal from its first cell to its second, or a
boundary edge whose normal does not point out of the domain, breaks the
invariant at O(1): the planted-mistake tests below show that it is caught.
"""

import numpy as np
import pytest

from oracle import sequential as seq
from pyop2_amd.workloads import airfoil, meshes


def free_stream_res(W, mesh):
    A = airfoil.Airfoil(W, mesh, q0=airfoil.initial_q(mesh.ncell, noise=0.0),
                        consts=(W is not seq))
    try:
        A.adt_calc()
        A.res_calc()
        A.bres_calc()
        return np.array(A.p_res.data_ro)
    finally:
        if W is not seq:
            A.remove_consts()


def scale(mesh):
    """Size of one edge flux: |q| times the edge length."""
    h = max(2.0 / mesh.nx, 1.0 / mesh.ny)
    return np.abs(airfoil.free_stream()).max() * h


def boundary_cells(mesh):
    return np.unique(mesh.pbecell)


@pytest.fixture
def consts():
    cs = [seq.Const(1, v, k, np.float64) for k, v in airfoil.CONSTS.items()]
    cs.append(seq.Const(4, airfoil.free_stream(), 'qinf', np.float64))
    yield
    for c in cs:
        c.remove_from_namespace()


def test_free_stream_far_field_is_steady(consts):
    mesh = meshes.QuadMesh(40, 20, seed=3)
    mesh.bound[:] = 0
    res = free_stream_res(seq, mesh)
    assert np.abs(res).max() <= 1e-13 * scale(mesh)


def test_free_stream_interior_is_steady_with_walls(consts):
    mesh = meshes.QuadMesh(40, 20, seed=4)
    assert mesh.bound.min() == 0 and mesh.bound.max() == 1
    res = free_stream_res(seq, mesh)
    interior = np.setdiff1d(np.arange(mesh.ncell), boundary_cells(mesh))
    assert np.abs(res[interior]).max() <= 1e-13 * scale(mesh)
    # wall edges (pressure only) leave boundary cells unbalanced
    assert np.abs(res[boundary_cells(mesh)]).max() > 1e-3 * scale(mesh)


@pytest.mark.parametrize('which', ['interior', 'boundary'])
def test_planted_orientation_error_is_caught(consts, which):
    """A generator that orients one edge the wrong way fails the invariant."""
    mesh = meshes.QuadMesh(40, 20, seed=5)
    mesh.bound[:] = 0
    if which == 'interior':
        mesh.pedge[7] = mesh.pedge[7, ::-1].copy()
    else:
        mesh.pbedge[3] = mesh.pbedge[3, ::-1].copy()
    res = free_stream_res(seq, mesh)
    assert np.abs(res).max() > 1e-3 * scale(mesh)


@pytest.mark.gpu
def test_free_stream_full_size_hip(hip):
    """The invariant on the benchmark's 1200 x 600 mesh, computed by the hip
    backend."""
    mesh = meshes.QuadMesh(1200, 600, seed=0)
    bound = mesh.bound.copy()
    res = free_stream_res(hip, mesh)
    interior = np.setdiff1d(np.arange(mesh.ncell), boundary_cells(mesh))
    assert np.abs(res[interior]).max() <= 1e-13 * scale(mesh)
    mesh.bound[:] = 0
    res = free_stream_res(hip, mesh)
    assert np.abs(res).max() <= 1e-13 * scale(mesh)
    mesh.bound[:] = bound
