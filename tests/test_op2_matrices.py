"""Sparsity, Mat assembly, zeroing and solves on every backend.

Golden values are the reference's (test/unit/test_matrices.py, kept as data
in tests/golden/reference_fixtures.json): the CSR of
its 4-element fixture (:54-62), the expected P1 mass matrix, vector mass
matrix and right-hand sides on its 2-triangle mesh (:579-608) with the
reference's tolerances, the zero_rows results (:765-778) and x = f for
M x = M f (:647-650, :780-783).  Kernels are the P1 kernels of
pyop2_amd.workloads.fem (exact quadrature)."""

import json
import os

import numpy as np
import pytest
from numpy.testing import assert_allclose

from pyop2_amd import op2
from pyop2_amd.exceptions import MapValueError
from pyop2_amd.workloads import fem

_FX = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden',
                                   'reference_fixtures.json')))
ELEM_NODE = np.asarray(_FX['two_triangles']['elem_node'], dtype=np.uint32)
COORDS = np.asarray(_FX['two_triangles']['coords'])
EXPECTED_MATRIX = np.asarray(_FX['expected_matrix']['value'])
EXPECTED_VECTOR_MATRIX = np.asarray(_FX['expected_vector_matrix']['value'])
EXPECTED_RHS = np.asarray(_FX['expected_rhs']['value'])
EXPECTED_VEC_RHS = np.asarray(_FX['expected_vec_rhs']['value'])
SPARSITY_FIXTURE = _FX['sparsity_fixture']
ZERO_ROWS = _FX['zero_rows']

MASS_VEC_TRI = """
void mass_vec_tri(double A[2][2], double *x[2], int j, int k)
{
  double det = (x[1][0] - x[0][0]) * (x[2][1] - x[0][1])
             - (x[2][0] - x[0][0]) * (x[1][1] - x[0][1]);
  double m = fabs(det) / 24.0 * (j == k ? 2.0 : 1.0);
  A[0][0] += m;
  A[1][1] += m;
}
"""

RHS_VEC_TRI = """
void rhs_vec_tri(double **A, double *x[2], double **f)
{
  double det = (x[1][0] - x[0][0]) * (x[2][1] - x[0][1])
             - (x[2][0] - x[0][0]) * (x[1][1] - x[0][1]);
  double a = fabs(det) / 24.0;
  for (int j = 0; j < 3; ++j)
    for (int r = 0; r < 2; ++r)
      A[j][r] += a * (2.0 * f[j][r] + f[(j + 1) % 3][r] + f[(j + 2) % 3][r]);
}
"""

RHS_TRI_ELEM = """
void rhs_tri_elem(double **A, double *x[2], double **f)
{
  const double phi[3][3] = {{2.0/3.0, 1.0/6.0, 1.0/6.0},
                            {1.0/6.0, 2.0/3.0, 1.0/6.0},
                            {1.0/6.0, 1.0/6.0, 2.0/3.0}};
  double det = (x[1][0] - x[0][0]) * (x[2][1] - x[0][1])
             - (x[2][0] - x[0][0]) * (x[1][1] - x[0][1]);
  double w = fabs(det) / 6.0;
  for (int j = 0; j < 3; ++j)
    for (int q = 0; q < 3; ++q) {
      double fq = phi[q][0] * f[0][0] + phi[q][1] * f[1][0] + phi[q][2] * f[2][0];
      A[j][0] += phi[q][j] * fq * w;
    }
}
"""

RHS_VEC_TRI_ITSPACE = """
void rhs_vec_tri_its(double *A, double *x[2], double **f, int j)
{
  double det = (x[1][0] - x[0][0]) * (x[2][1] - x[0][1])
             - (x[2][0] - x[0][0]) * (x[1][1] - x[0][1]);
  double a = fabs(det) / 24.0;
  for (int r = 0; r < 2; ++r)
    A[r] += a * (2.0 * f[j][r] + f[(j + 1) % 3][r] + f[(j + 2) % 3][r]);
}
"""

ZERO_DAT = "void zero_dat(double *d) { *d = 0.0; }"
ZERO_VEC_DAT = "void zero_vec_dat(double *d) { d[0] = 0.0; d[1] = 0.0; }"


@pytest.fixture
def mesh(backend):
    nodes = op2.Set(4, 1, "nodes")
    vnodes = op2.Set(4, 2, "vnodes")
    elements = op2.Set(2, 1, "elements")
    return {
        'nodes': nodes, 'vnodes': vnodes, 'elements': elements,
        'elem_node': op2.Map(elements, nodes, 3, ELEM_NODE, "elem_node"),
        'elem_vnode': op2.Map(elements, vnodes, 3, ELEM_NODE, "elem_vnode"),
        'coords': op2.Dat(vnodes, COORDS.copy(), np.float64, "coords"),
        'f': op2.Dat(nodes, np.asarray([1.0, 2.0, 3.0, 4.0]), np.float64, "f"),
        'f_vec': op2.Dat(vnodes, np.asarray([(1.0, 2.0)] * 4), np.float64, "f"),
    }


def mass_mat(m):
    sp = op2.Sparsity((m['elem_node'], m['elem_node']), "sparsity")
    mat = op2.Mat(sp, np.float64, "mat")
    op2.par_loop(op2.Kernel(fem.MASS_TRI, "mass_tri"), m['elements'](3, 3),
                 mat((m['elem_node'][op2.i[0]], m['elem_node'][op2.i[1]]), op2.INC),
                 m['coords'](m['elem_vnode'], op2.READ))
    return mat


def vec_mass_mat(m):
    sp = op2.Sparsity((m['elem_vnode'], m['elem_vnode']), "sparsity")
    mat = op2.Mat(sp, np.float64, "vecmat")
    op2.par_loop(op2.Kernel(MASS_VEC_TRI, "mass_vec_tri"), m['elements'](3, 3),
                 mat((m['elem_vnode'][op2.i[0]], m['elem_vnode'][op2.i[1]]), op2.INC),
                 m['coords'](m['elem_vnode'], op2.READ))
    return mat


def rhs(m):
    b = op2.Dat(m['nodes'], np.zeros(4), np.float64, "b")
    op2.par_loop(op2.Kernel(fem.RHS_TRI, "rhs_tri"), m['elements'](3),
                 b(m['elem_node'][op2.i[0]], op2.INC), m['coords'](m['elem_vnode'], op2.READ),
                 m['f'](m['elem_node'], op2.READ))
    return b


class TestSparsity:

    def test_build_sparsity(self, backend):
        elements = op2.Set(4)
        nodes = op2.Set(5)
        elem_node = op2.Map(elements, nodes, 3, SPARSITY_FIXTURE['elem_node'])
        sparsity = op2.Sparsity((elem_node, elem_node))
        assert list(sparsity._rowptr) == SPARSITY_FIXTURE['rowptr']
        assert list(sparsity._colidx) == SPARSITY_FIXTURE['colidx']

    def test_sparsity_null_maps(self, backend):
        s = op2.Set(5)
        with pytest.raises(MapValueError):
            m = op2.Map(s, s, 1)
            op2.Sparsity((m, m))

    def test_vector_sparsity_expands_blocks(self, mesh):
        sp = op2.Sparsity((mesh['elem_vnode'], mesh['elem_vnode']))
        scalar = op2.Sparsity((mesh['elem_node'], mesh['elem_node']))
        assert sp.nz == 4 * scalar.nz
        for r in range(8):
            cols = sp._colidx[sp._rowptr[r]:sp._rowptr[r + 1]]
            n = r // 2
            base = scalar._colidx[scalar._rowptr[n]:scalar._rowptr[n + 1]]
            assert list(cols) == [2 * c + k for c in base for k in range(2)]

    def test_sparsity_cache(self, mesh):
        en = mesh['elem_node']
        assert op2.Sparsity((en, en)) is op2.Sparsity((en, en))
        assert op2.Sparsity(en) is op2.Sparsity((en, en))


class TestMatrices:

    def test_minimal_zero_mat(self, backend):
        n = 128
        s = op2.Set(n)
        m = op2.Map(s, s, 1, np.arange(n, dtype=np.uint32))
        mat = op2.Mat(op2.Sparsity((m, m)), np.float64)
        k = op2.Kernel("void zero_mat(double local_mat[1][1], int i, int j) "
                       "{ local_mat[i][j] = 0.0; }", "zero_mat")
        op2.par_loop(k, s(1, 1), mat((m[op2.i[0]], m[op2.i[1]]), op2.WRITE))
        assert_allclose(mat.values, np.zeros((n, n)), 1e-12)

    def test_assemble_mat(self, mesh):
        assert_allclose(mass_mat(mesh).values, EXPECTED_MATRIX, 1e-5)

    def test_assemble_rhs(self, mesh):
        assert_allclose(rhs(mesh).data, EXPECTED_RHS, 1e-6)

    def test_solve(self, mesh):
        mat = mass_mat(mesh)
        b = rhs(mesh)
        x = op2.Dat(mesh['nodes'], np.zeros(4), np.float64, "x")
        op2.solve(mat, x, b)
        assert_allclose(x.data, mesh['f'].data, 1e-8)

    @pytest.mark.parametrize('method', ['gmres', 'bcgs', 'bicgstab'])
    def test_solve_krylov(self, mesh, method):
        """The same solve through GMRES and BiCGStab (the reference's demos
        ask for GMRES, demo/laplace_ffc.py:151)."""
        mat = mass_mat(mesh)
        b = rhs(mesh)
        x = op2.Dat(mesh['nodes'], np.zeros(4), np.float64, "x")
        solver = op2.Solver(linear_solver=method, relative_tolerance=1e-12)
        solver.solve(mat, x, b)
        assert solver.getConvergedReason() > 0
        assert_allclose(x.data, mesh['f'].data, 1e-8)

    def test_zero_matrix(self, mesh):
        mat = mass_mat(mesh)
        mat.zero()
        assert_allclose(mat.values, np.zeros((4, 4)), 1e-14)

    def test_set_matrix(self, mesh):
        mat = mass_mat(mesh)
        mat.zero()
        g = op2.Global(1, 1.0, np.float64, "g")
        en, el = mesh['elem_node'], mesh['elements']
        kinc = op2.Kernel("void kernel_inc(double entry[1][1], double* g, int i, int j) "
                          "{ entry[0][0] += *g; }", "kernel_inc")
        kset = op2.Kernel("void kernel_set(double entry[1][1], double* g, int i, int j) "
                          "{ entry[0][0] = *g; }", "kernel_set")
        op2.par_loop(kinc, el(3, 3), mat((en[op2.i[0]], en[op2.i[1]]), op2.INC), g(op2.READ))
        assert mat.array.sum() == 3 * 3 * el.size
        op2.par_loop(kset, el(3, 3), mat((en[op2.i[0]], en[op2.i[1]]), op2.WRITE), g(op2.READ))
        assert_allclose(mat.array, np.ones_like(mat.array))

    def test_set_matrix_vec(self, mesh):
        vm = vec_mass_mat(mesh)
        vm.zero()
        g = op2.Global(1, 1.0, np.float64, "g")
        ev, el = mesh['elem_vnode'], mesh['elements']
        kinc = op2.Kernel("""void kernel_inc_vec(double entry[2][2], double* g, int i, int j) {
            entry[0][0] += *g; entry[0][1] += *g; entry[1][0] += *g; entry[1][1] += *g; }""",
                          "kernel_inc_vec")
        kset = op2.Kernel("""void kernel_set_vec(double entry[2][2], double* g, int i, int j) {
            entry[0][0] = *g; entry[0][1] = *g; entry[1][0] = *g; entry[1][1] = *g; }""",
                          "kernel_set_vec")
        op2.par_loop(kinc, el(3, 3), vm((ev[op2.i[0]], ev[op2.i[1]]), op2.INC), g(op2.READ))
        assert vm.array.sum() == 2 * 2 * 3 * 3 * el.size
        op2.par_loop(kset, el(3, 3), vm((ev[op2.i[0]], ev[op2.i[1]]), op2.WRITE), g(op2.READ))
        assert_allclose(vm.array, np.ones_like(vm.array))

    def test_zero_rows(self, mesh):
        mat = mass_mat(mesh)
        expected = EXPECTED_MATRIX.copy()
        first, last = ZERO_ROWS['first'], ZERO_ROWS['last']
        expected[first['row']] = first['value']
        mat.zero_rows([first['row']], first['diag'])
        assert_allclose(mat.values, expected, 1e-5)
        expected[last['row']] = last['value']
        mat.zero_rows([last['row']], last['diag'])
        assert_allclose(mat.values, expected, 1e-5)

    def test_assemble_vec_mass(self, mesh):
        assert_allclose(vec_mass_mat(mesh).values, EXPECTED_VECTOR_MATRIX, 1e-6)

    def test_rhs_vector(self, mesh):
        b = op2.Dat(mesh['vnodes'], np.zeros((4, 2)), np.float64, "b")
        op2.par_loop(op2.Kernel(RHS_VEC_TRI, "rhs_vec_tri"), mesh['elements'],
                     b(mesh['elem_vnode'], op2.INC), mesh['coords'](mesh['elem_vnode'], op2.READ),
                     mesh['f_vec'](mesh['elem_vnode'], op2.READ))
        assert_allclose(b.data, EXPECTED_VEC_RHS, 1e-6)

    def test_vector_solve(self, mesh):
        vm = vec_mass_mat(mesh)
        b = op2.Dat(mesh['vnodes'], np.zeros((4, 2)), np.float64, "b")
        op2.par_loop(op2.Kernel(RHS_VEC_TRI, "rhs_vec_tri"), mesh['elements'],
                     b(mesh['elem_vnode'], op2.INC), mesh['coords'](mesh['elem_vnode'], op2.READ),
                     mesh['f_vec'](mesh['elem_vnode'], op2.READ))
        x = op2.Dat(mesh['vnodes'], np.zeros((4, 2)), np.float64, "x")
        op2.solve(vm, x, b)
        assert_allclose(x.data, mesh['f_vec'].data, 1e-12)

    @pytest.mark.parametrize('method', ['gmres', 'bicgstab'])
    def test_vector_solve_krylov(self, mesh, method):
        """The vector mass system (2 x 2 blocks expanded into the scalar
        CSR) through GMRES and BiCGStab."""
        vm = vec_mass_mat(mesh)
        b = op2.Dat(mesh['vnodes'], np.zeros((4, 2)), np.float64, "b")
        op2.par_loop(op2.Kernel(RHS_VEC_TRI, "rhs_vec_tri"), mesh['elements'],
                     b(mesh['elem_vnode'], op2.INC), mesh['coords'](mesh['elem_vnode'], op2.READ),
                     mesh['f_vec'](mesh['elem_vnode'], op2.READ))
        x = op2.Dat(mesh['vnodes'], np.zeros((4, 2)), np.float64, "x")
        op2.Solver(linear_solver=method, relative_tolerance=1e-14).solve(vm, x, b)
        assert_allclose(x.data, mesh['f_vec'].data, 1e-11)

    def test_zero_rhs(self, mesh):
        """A WRITE loop zeroes the right-hand side (reference
        test_matrices.py:693-697)."""
        b = rhs(mesh)
        op2.par_loop(op2.Kernel(ZERO_DAT, "zero_dat"), mesh['nodes'],
                     b(op2.IdentityMap, op2.WRITE))
        assert all(b.data_ro.ravel() == 0.0)

    def test_rhs_whole_element(self, mesh):
        """The right-hand side from a kernel that fills all three rows of an
        element through a vector map (reference rhs_ffc, :718-726)."""
        b = op2.Dat(mesh['nodes'], np.zeros(4), np.float64, "b")
        op2.par_loop(op2.Kernel(RHS_TRI_ELEM, "rhs_tri_elem"), mesh['elements'],
                     b(mesh['elem_node'], op2.INC), mesh['coords'](mesh['elem_vnode'], op2.READ),
                     mesh['f'](mesh['elem_node'], op2.READ))
        assert_allclose(b.data, EXPECTED_RHS, 1e-6)

    def test_rhs_vector_itspace(self, mesh):
        """The vector right-hand side one row per iteration-space index,
        after zeroing it with a WRITE loop (reference :751-763)."""
        b = op2.Dat(mesh['vnodes'], np.ones((4, 2)), np.float64, "b")
        op2.par_loop(op2.Kernel(ZERO_VEC_DAT, "zero_vec_dat"), mesh['vnodes'],
                     b(op2.IdentityMap, op2.WRITE))
        op2.par_loop(op2.Kernel(RHS_VEC_TRI_ITSPACE, "rhs_vec_tri_its"), mesh['elements'](3),
                     b(mesh['elem_vnode'][op2.i[0]], op2.INC),
                     mesh['coords'](mesh['elem_vnode'], op2.READ),
                     mesh['f_vec'](mesh['elem_vnode'], op2.READ))
        assert_allclose(b.data, EXPECTED_VEC_RHS, 1e-6)

    def test_zero_vector_matrix(self, mesh):
        vm = vec_mass_mat(mesh)
        vm.zero()
        assert_allclose(vm.values, np.zeros((8, 8)), 1e-14)

    def test_solver_nonconvergence_raises(self, mesh):
        mat = mass_mat(mesh)
        b = rhs(mesh)
        x = op2.Dat(mesh['nodes'], np.zeros(4), np.float64, "x")
        solver = op2.Solver(maximum_iterations=1, relative_tolerance=1e-15)
        with pytest.raises(RuntimeError):
            solver.solve(mat, x, b)
        solver = op2.Solver(maximum_iterations=1, relative_tolerance=1e-15,
                            error_on_nonconvergence=False)
        solver.solve(mat, x, b)

    def test_union_sparsity_assembly(self, mesh):
        """A Sparsity built from two map pairs holds the union of their
        patterns; assembling through one pair fills its entries only."""
        en, el = mesh['elem_node'], mesh['elements']
        other = op2.Map(el, mesh['nodes'], 3, np.asarray([0, 0, 0, 2, 2, 2], dtype=np.uint32))
        sp = op2.Sparsity(((en, en), (other, other)))
        mat = op2.Mat(sp, np.float64)
        k = op2.Kernel("void kone(double e[1][1], int i, int j) { e[0][0] = 1.0; }", "kone")
        op2.par_loop(k, el(3, 3), mat((en[op2.i[0]], en[op2.i[1]]), op2.INC))
        assert mat.array.sum() == 18
