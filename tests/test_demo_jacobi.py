"""The reference's Jacobi demo (demo/jacobi.py, a port of OP2-Common's):
a 5 x 5 grid of unknowns with a 5-point stencil stored per edge (self
edges carry the diagonal), iterated as
    du[n] += beta * A[e] * u[m]     for every edge e = (n, m)   (res)
    u += du + alpha * r;  du = 0;  u_sum += u*u;  u_max = max(u_max, u)
with a Const alpha, a READ Global beta, an INC into nodes through the edge
map's first column, a direct INC Dat and INC/MAX Globals.  The mesh is
built as the demo builds it (:80-112).  Pinned against the same iteration
written with numpy (np.add.at applies increments in edge order, the
sequential order), in float64 and float32."""

import numpy as np
import pytest

from pyop2_amd import op2

NN = 6


def jacobi_mesh():
    """(pp edge->node pairs, A per edge, r per node) as demo/jacobi.py:80-112
    lays them out."""
    nnode = (NN - 1) ** 2
    nedge = nnode + 4 * (NN - 1) * (NN - 2)
    pp = np.zeros(2 * nedge, dtype=np.int64)
    A = np.zeros(nedge)
    r = np.zeros(nnode)
    e = 0
    for i in range(1, NN):
        for j in range(1, NN):
            n = i - 1 + (j - 1) * (NN - 1)
            pp[2 * e] = pp[2 * e + 1] = n
            A[e] = -1
            e += 1
            for di, dj in ((-1, 0), (1, 0), (0, -1), (0, 1)):
                i2, j2 = i + di, j + dj
                if i2 == 0 or i2 == NN or j2 == 0 or j2 == NN:
                    r[n] += 0.25
                else:
                    pp[2 * e] = n
                    pp[2 * e + 1] = i2 - 1 + (j2 - 1) * (NN - 1)
                    A[e] = 0.25
                    e += 1
    assert e == nedge
    return pp, A, r


def numpy_jacobi(pp, A, r, niter, dtype):
    """The demo's two loops in numpy, in the same order of operations."""
    A = A.astype(dtype)
    r = r.astype(dtype)
    u = np.zeros_like(r)
    du = np.zeros_like(r)
    alpha, beta = dtype(1.0), dtype(1.0)
    e0, e1 = pp[0::2], pp[1::2]
    hist = []
    for _ in range(niter):
        np.add.at(du, e0, beta * A * u[e1])
        u = u + (du + alpha * r)
        du[:] = 0
        hist.append((float((u * u).sum()), float(u.max())))
    return u, hist


def run_demo(niter, dtype):
    pp, A, r = jacobi_mesh()
    nnode, nedge = len(r), len(A)
    t = 'double' if dtype == np.float64 else 'float'
    nodes = op2.Set(nnode, 1, "nodes")
    edges = op2.Set(nedge, 1, "edges")
    ppedge = op2.Map(edges, nodes, 2, pp, "ppedge")
    p_A = op2.Dat(edges, data=A.astype(dtype), name="p_A")
    p_r = op2.Dat(nodes, data=r.astype(dtype), name="p_r")
    p_u = op2.Dat(nodes, data=np.zeros(nnode, dtype), name="p_u")
    p_du = op2.Dat(nodes, data=np.zeros(nnode, dtype), name="p_du")
    alpha = op2.Const(1, data=1.0, name="alpha", dtype=dtype)
    beta = op2.Global(1, data=1.0, name="beta", dtype=dtype)
    res = op2.Kernel("void res(%(t)s *A, %(t)s *u, %(t)s *du, const %(t)s *beta) {"
                     " *du += (*beta)*(*A)*(*u); }" % {'t': t}, "res")
    update = op2.Kernel("""void update(%(t)s *r, %(t)s *du, %(t)s *u, %(t)s *u_sum,
                                       %(t)s *u_max) {
  *u += *du + alpha * (*r);
  *du = %(z)s;
  *u_sum += (*u)*(*u);
  *u_max = *u_max > *u ? *u_max : *u;
}""" % {'t': t, 'z': '0.0' if dtype == np.float64 else '0.0f'}, "update")
    hist = []
    try:
        for _ in range(niter):
            op2.par_loop(res, edges, p_A(op2.IdentityMap, op2.READ),
                         p_u(ppedge[1], op2.READ), p_du(ppedge[0], op2.INC), beta(op2.READ))
            u_sum = op2.Global(1, data=0.0, name="u_sum", dtype=dtype)
            u_max = op2.Global(1, data=0.0, name="u_max", dtype=dtype)
            op2.par_loop(update, nodes, p_r(op2.IdentityMap, op2.READ),
                         p_du(op2.IdentityMap, op2.RW), p_u(op2.IdentityMap, op2.INC),
                         u_sum(op2.INC), u_max(op2.MAX))
            hist.append((float(u_sum.data[0]), float(u_max.data[0])))
    finally:
        alpha.remove_from_namespace()
    return p_u.data_ro.ravel().copy(), hist


@pytest.mark.parametrize('dtype,niter', [(np.float64, 2), (np.float64, 40), (np.float32, 2),
                                         (np.float32, 40)])
def test_jacobi_demo(backend, dtype, niter):
    """u bit-exact against the numpy iteration (increments applied in edge
    order); the u_sum Global within rounding of its summation order; u_max
    exact."""
    pp, A, r = jacobi_mesh()
    u_ref, hist_ref = numpy_jacobi(pp, A, r, niter, dtype)
    u, hist = run_demo(niter, dtype)
    assert np.array_equal(u, u_ref)
    tol = 1e-12 if dtype == np.float64 else 1e-5
    for (s, m), (sr, mr) in zip(hist, hist_ref):
        assert abs(s - sr) <= tol * abs(sr)
        assert m == mr


def test_jacobi_converges(backend):
    """The iteration converges towards the discrete solution of the
    5-point Laplacian with unit boundary values: every u tends to 1."""
    u, hist = run_demo(400, np.float64)
    assert np.allclose(u, 1.0, atol=1e-6)
