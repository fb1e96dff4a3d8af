"""P1 finite-element assembly workloads (mass, right-hand side, Laplace with a
coefficient, vector mass) on triangles and tetrahedra.

Kernels follow the op2 kernel conventions of the reference's matrix tests
(test/unit/test_matrices.py: ``A[1][1]`` local entry plus ``int j, int k``
iteration-space indices for scalar Mats, ``A[rdim][cdim]`` for vector
Mats, ``double *x[2]`` vector-map coordinates); their bodies are written
here from the closed forms of P1 elements:

* triangle mass with the 3-point rule at the edge-interior points
  (2/3, 1/6, 1/6): exact, M_jk = |T|/12 (1 + d_jk);
* Laplace: K_jk = kappa_T |T| grad(phi_j).grad(phi_k) with kappa_T the
  mean of the vertex coefficients (exact for the P1 interpolant of kappa);
* tetrahedron mass M_jk = |T|/20 (1 + d_jk), vector mass its 2x2 block
  identity expansion.
"""

import numpy as np

MASS_TRI = """
void mass_tri(double A[1][1], double *x[2], int j, int k)
{
  const double phi[3][3] = {{2.0/3.0, 1.0/6.0, 1.0/6.0},
                            {1.0/6.0, 2.0/3.0, 1.0/6.0},
                            {1.0/6.0, 1.0/6.0, 2.0/3.0}};
  double det = (x[1][0] - x[0][0]) * (x[2][1] - x[0][1])
             - (x[2][0] - x[0][0]) * (x[1][1] - x[0][1]);
  double w = fabs(det) / 6.0;
  for (int q = 0; q < 3; ++q)
    A[0][0] += phi[q][j] * phi[q][k] * w;
}
"""

RHS_TRI = """
void rhs_tri(double A[1], double *x[2], double **f, int j)
{
  const double phi[3][3] = {{2.0/3.0, 1.0/6.0, 1.0/6.0},
                            {1.0/6.0, 2.0/3.0, 1.0/6.0},
                            {1.0/6.0, 1.0/6.0, 2.0/3.0}};
  double det = (x[1][0] - x[0][0]) * (x[2][1] - x[0][1])
             - (x[2][0] - x[0][0]) * (x[1][1] - x[0][1]);
  double w = fabs(det) / 6.0;
  for (int q = 0; q < 3; ++q) {
    double fq = phi[q][0] * f[0][0] + phi[q][1] * f[1][0] + phi[q][2] * f[2][0];
    A[0] += phi[q][j] * fq * w;
  }
}
"""

LAPLACE_TRI = """
void laplace_tri(double A[1][1], double *x[2], double **kappa, int j, int k)
{
  const double g[3][2] = {{-1.0, -1.0}, {1.0, 0.0}, {0.0, 1.0}};
  double J00 = x[1][0] - x[0][0], J01 = x[2][0] - x[0][0];
  double J10 = x[1][1] - x[0][1], J11 = x[2][1] - x[0][1];
  double det = J00 * J11 - J01 * J10;
  double gj0 = ( J11 * g[j][0] - J10 * g[j][1]) / det;
  double gj1 = (-J01 * g[j][0] + J00 * g[j][1]) / det;
  double gk0 = ( J11 * g[k][0] - J10 * g[k][1]) / det;
  double gk1 = (-J01 * g[k][0] + J00 * g[k][1]) / det;
  double kap = (kappa[0][0] + kappa[1][0] + kappa[2][0]) / 3.0;
  A[0][0] += kap * (gj0 * gk0 + gj1 * gk1) * 0.5 * fabs(det);
}
"""

LAPLACE_TET = """
void laplace_tet(double A[1][1], double *x[3], double **kappa, int j, int k)
{
  double a[3][3];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) a[r][c] = x[c + 1][r] - x[0][r];
  double c00 = a[1][1] * a[2][2] - a[1][2] * a[2][1];
  double c01 = a[1][2] * a[2][0] - a[1][0] * a[2][2];
  double c02 = a[1][0] * a[2][1] - a[1][1] * a[2][0];
  double c10 = a[0][2] * a[2][1] - a[0][1] * a[2][2];
  double c11 = a[0][0] * a[2][2] - a[0][2] * a[2][0];
  double c12 = a[0][1] * a[2][0] - a[0][0] * a[2][1];
  double c20 = a[0][1] * a[1][2] - a[0][2] * a[1][1];
  double c21 = a[0][2] * a[1][0] - a[0][0] * a[1][2];
  double c22 = a[0][0] * a[1][1] - a[0][1] * a[1][0];
  double det = a[0][0] * c00 + a[0][1] * c01 + a[0][2] * c02;
  double idet = 1.0 / det;
  /* rows of J^-T as cofactors / det; reference gradients of the 4 basis
     functions: (-1,-1,-1), e0, e1, e2 */
  const double g[4][3] = {{-1.0, -1.0, -1.0}, {1.0, 0.0, 0.0}, {0.0, 1.0, 0.0}, {0.0, 0.0, 1.0}};
  double gj[3], gk[3];
  for (int r = 0; r < 3; ++r) {
    double cr0 = r == 0 ? c00 : r == 1 ? c01 : c02;
    double cr1 = r == 0 ? c10 : r == 1 ? c11 : c12;
    double cr2 = r == 0 ? c20 : r == 1 ? c21 : c22;
    gj[r] = (cr0 * g[j][0] + cr1 * g[j][1] + cr2 * g[j][2]) * idet;
    gk[r] = (cr0 * g[k][0] + cr1 * g[k][1] + cr2 * g[k][2]) * idet;
  }
  double kap = 0.25 * (kappa[0][0] + kappa[1][0] + kappa[2][0] + kappa[3][0]);
  A[0][0] += kap * (gj[0] * gk[0] + gj[1] * gk[1] + gj[2] * gk[2]) * fabs(det) / 6.0;
}
"""

MASS_VEC_TET = """
void mass_vec_tet(double A[2][2], double *x[3], int j, int k)
{
  double a[3][3];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) a[r][c] = x[c + 1][r] - x[0][r];
  double det = a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1])
             - a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0])
             + a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]);
  double m = fabs(det) / 6.0 / 20.0 * (j == k ? 2.0 : 1.0);
  A[0][0] += m;
  A[1][1] += m;
}
"""

MASS_TET = """
void mass_tet(double A[1][1], double *x[3], int j, int k)
{
  double a[3][3];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) a[r][c] = x[c + 1][r] - x[0][r];
  double det = a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1])
             - a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0])
             + a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]);
  A[0][0] += fabs(det) / 6.0 / 20.0 * (j == k ? 2.0 : 1.0);
}
"""


RHS_TET = """
void rhs_tet(double A[1], double *x[3], double **f, int j)
{
  double a[3][3];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) a[r][c] = x[c + 1][r] - x[0][r];
  double det = a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1])
             - a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0])
             + a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]);
  /* row j of the exact P1 mass matrix |T|/20 (1 + d_jk) applied to f */
  double s = f[0][0] + f[1][0] + f[2][0] + f[3][0] + f[j][0];
  A[0] += fabs(det) / 6.0 / 20.0 * s;
}
"""


def mass2d(W, mesh, f, solve_rtol=1e-12, name='m2d'):
    """P1 mass matrix and right-hand side M f on a triangle mesh, then the
    CG solve of M x = b; returns host copies (the mass2d demo pattern,
    reference demo/mass2d_triangle.py:91-116)."""
    nodes = W.Set(mesh.nvert, 1, name + '_nodes')
    vnodes = W.Set(mesh.nvert, 2, name + '_vnodes')
    elements = W.Set(mesh.nelem, 1, name + '_elements')
    elem_node = W.Map(elements, nodes, 3, mesh.elem_node, name + '_elem_node')
    elem_vnode = W.Map(elements, vnodes, 3, mesh.elem_node, name + '_elem_vnode')
    sparsity = W.Sparsity((elem_node, elem_node), name + '_sparsity')
    mat = W.Mat(sparsity, np.float64, name + '_mat')
    coords = W.Dat(vnodes, mesh.coords.copy(), np.float64, name + '_coords')
    fd = W.Dat(nodes, np.asarray(f, dtype=np.float64).copy(), np.float64, name + '_f')
    b = W.Dat(nodes, np.zeros(mesh.nvert), np.float64, name + '_b')
    x = W.Dat(nodes, np.zeros(mesh.nvert), np.float64, name + '_x')
    W.par_loop(W.Kernel(MASS_TRI, 'mass_tri'), elements(3, 3),
               mat((elem_node[W.i[0]], elem_node[W.i[1]]), W.INC),
               coords(elem_vnode, W.READ))
    W.par_loop(W.Kernel(RHS_TRI, 'rhs_tri'), elements(3),
               b(elem_node[W.i[0]], W.INC), coords(elem_vnode, W.READ),
               fd(elem_node, W.READ))
    solver = W.Solver(linear_solver='cg', preconditioner='jacobi',
                      relative_tolerance=solve_rtol, maximum_iterations=10000)
    solver.solve(mat, x, b)
    return {'rowptr': np.array(sparsity._rowptr), 'colidx': np.array(sparsity._colidx),
            'mat': np.array(mat.array), 'b': np.array(b.data_ro), 'x': np.array(x.data_ro),
            'its': solver.getIterationNumber()}


class Assembly3D(object):
    """P1 Laplace (coefficient per vertex) and vector-P1 mass (2 dofs per
    vertex, block CSR) on a tetrahedral mesh in world ``W``: the 3-D
    assembly workload.  ``laplace()`` / ``vector_mass()`` run the assembly
    par_loops into ``A`` / ``M``."""

    def __init__(self, W, mesh, kappa, name='a3d'):
        self.W = W
        nodes = W.Set(mesh.nvert, 1, name + '_nodes')
        vnodes = W.Set(mesh.nvert, 2, name + '_vnodes')
        xnodes = W.Set(mesh.nvert, 3, name + '_xnodes')
        tets = W.Set(mesh.nelem, 1, name + '_tets')
        self.nodes, self.vnodes, self.xnodes, self.tets = nodes, vnodes, xnodes, tets
        self.tet_node = W.Map(tets, nodes, 4, mesh.elem_node, name + '_tet_node')
        self.tet_vnode = W.Map(tets, vnodes, 4, mesh.elem_node, name + '_tet_vnode')
        self.tet_xnode = W.Map(tets, xnodes, 4, mesh.elem_node, name + '_tet_xnode')
        self.coords = W.Dat(xnodes, mesh.coords.copy(), np.float64, name + '_coords')
        self.kappa = W.Dat(nodes, np.asarray(kappa, dtype=np.float64).copy(), np.float64,
                           name + '_kappa')
        self.sp_scalar = W.Sparsity((self.tet_node, self.tet_node), name + '_sp')
        self.sp_vector = W.Sparsity((self.tet_vnode, self.tet_vnode), name + '_vsp')
        self.A = W.Mat(self.sp_scalar, np.float64, name + '_A')
        self.M = W.Mat(self.sp_vector, np.float64, name + '_M')
        self.k_laplace = W.Kernel(LAPLACE_TET, 'laplace_tet')
        self.k_vmass = W.Kernel(MASS_VEC_TET, 'mass_vec_tet')

    def laplace(self):
        W = self.W
        W.par_loop(self.k_laplace, self.tets(4, 4),
                   self.A((self.tet_node[W.i[0]], self.tet_node[W.i[1]]), W.INC),
                   self.coords(self.tet_xnode, W.READ), self.kappa(self.tet_node, W.READ))

    def vector_mass(self):
        W = self.W
        W.par_loop(self.k_vmass, self.tets(4, 4),
                   self.M((self.tet_vnode[W.i[0]], self.tet_vnode[W.i[1]]), W.INC),
                   self.coords(self.tet_xnode, W.READ))


def kuhn_nnz(n):
    """Closed-form nonzero count of the P1 pattern of the n^3 Kuhn mesh
    (each vertex couples to itself and its Kuhn neighbours: offsets in
    {0,1}^3 and {0,-1}^3 and the 6 axis/face-diagonal offsets of the
    shared main diagonal), counted by brute force over vertex offsets."""
    offs = set()
    for t in [(0, 1, 3, 7), (0, 1, 5, 7), (0, 2, 3, 7), (0, 2, 6, 7), (0, 4, 5, 7), (0, 4, 6, 7)]:
        for a in t:
            for b in t:
                pa = (a & 1, a >> 1 & 1, a >> 2 & 1)
                pb = (b & 1, b >> 1 & 1, b >> 2 & 1)
                offs.add(tuple(pb[d] - pa[d] for d in range(3)))
    total = 0
    for o in offs:
        cnt = 1
        for d in range(3):
            cnt *= (n + 1) - abs(o[d])
        total += cnt
    return total
