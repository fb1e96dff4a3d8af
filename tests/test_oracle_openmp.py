"""The OpenMP CPU baseline (oracle/openmp.py) computes what the sequential
oracle computes: airfoil steps and a Mat assembly within 1e-12 (block
colours reorder the INC sums), and its plan never gives two blocks of one
colour a common INC target."""

import numpy as np

from oracle import openmp as omp
from oracle import sequential as seq
from pyop2_amd.workloads import airfoil, fem, meshes


def test_openmp_airfoil_matches_sequential():
    mesh = meshes.QuadMesh(120, 60, seed=2)

    class _W(object):
        pass
    A = airfoil.Airfoil(seq, mesh)
    # a second problem on the same Consts, looped by the OpenMP baseline
    B = airfoil.Airfoil(seq, mesh, consts=False)
    W = _W()
    for k in dir(seq):
        setattr(W, k, getattr(seq, k))
    W.par_loop = omp.par_loop
    B.W = W
    try:
        for _ in range(2):
            A.iteration()
            B.iteration()
    finally:
        A.remove_consts()
    qa, qb = A.p_q.data_ro, B.p_q.data_ro
    assert np.abs(qa - qb).max() <= 1e-12 * np.abs(qa).max()
    ra, rb = A.rms.data[0], B.rms.data[0]
    assert abs(ra - rb) <= 1e-12 * abs(ra)


def test_openmp_plan_conflict_free():
    mesh = meshes.QuadMesh(90, 40, seed=1)
    edges = seq.Set(mesh.nedge)
    cells = seq.Set(mesh.ncell, 4)
    pecell = seq.Map(edges, cells, 2, mesh.pecell)
    res = seq.Dat(cells, np.zeros((mesh.ncell, 4)), np.float64)
    ncol, off, blk = omp.plan(edges, mesh.nedge, [res(pecell, seq.INC)], part=256)
    assert sorted(blk.tolist()) == list(range(len(blk)))
    for c in range(ncol):
        seen = set()
        for b in blk[off[c]:off[c + 1]]:
            t = set(mesh.pecell[b * 256:(b + 1) * 256].reshape(-1).tolist())
            assert not (t & seen)
            seen |= t


def test_openmp_mass_matches_sequential():
    mesh = meshes.TriMesh(20, seed=5)
    nodes = seq.Set(mesh.nvert)
    vnodes = seq.Set(mesh.nvert, 2)
    el = seq.Set(mesh.nelem)
    en = seq.Map(el, nodes, 3, mesh.elem_node)
    ev = seq.Map(el, vnodes, 3, mesh.elem_node)
    x = seq.Dat(vnodes, mesh.coords.copy(), np.float64)
    out = []
    for loop in (seq.par_loop, omp.par_loop):
        A = seq.Mat(seq.Sparsity((en, en)), np.float64)
        loop(seq.Kernel(fem.MASS_TRI, 'mass_tri'), el(3, 3),
             A((en[seq.i[0]], en[seq.i[1]]), seq.INC), x(ev, seq.READ))
        out.append(A.values.copy())
    assert np.abs(out[0] - out[1]).max() <= 1e-12 * np.abs(out[0]).max()
