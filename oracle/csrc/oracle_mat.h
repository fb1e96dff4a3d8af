/* CSR matrix insertion of the CPU oracle (C, header-only).
 *
 * TEST INFRASTRUCTURE ONLY: included by the oracle's generated loop wrappers.
 * Restates addto_scalar / addto_vector (pyop2/mat_utils.cxx:4-28) applied as
 * PETSc MatSetValuesLocal on an AIJ matrix with IGNORE_ZERO_ENTRIES and
 * NEW_NONZERO_ALLOCATION_ERR (pyop2/petsc_base.py:110-119): an added zero is
 * skipped, INSERT overwrites, an entry outside the pattern sets `err`. */
#ifndef ORACLE_MAT_H
#define ORACLE_MAT_H

typedef struct {
  const int* rowptr;
  const int* colidx;
  double* vals;
  int nrows;
  int err;
} oracle_csr;

static long long oracle_csr_find(const oracle_csr* m, int row, int col) {
  int lo, hi;
  if (row < 0 || row >= m->nrows) return -1;
  lo = m->rowptr[row];
  hi = m->rowptr[row + 1] - 1;
  while (lo <= hi) {
    int mid = (lo + hi) / 2;
    int c = m->colidx[mid];
    if (c == col) return mid;
    if (c < col) lo = mid + 1;
    else hi = mid - 1;
  }
  return -1;
}

static void oracle_set_value(oracle_csr* m, int row, int col, double v, int insert) {
  long long k;
  if (!insert && v == 0.0) return;
  k = oracle_csr_find(m, row, col);
  if (k < 0) {
    m->err = 1;
    return;
  }
  if (insert) m->vals[k] = v;
  else m->vals[k] += v;
}

static void addto_scalar(oracle_csr* mat, const void* value, int row, int col, int insert) {
  const double* v = (const double*)value;
  if (v[0] == 0.0 && !insert) return;
  oracle_set_value(mat, row, col, v[0], insert);
}

static void addto_vector(oracle_csr* mat, const void* values, int nrows, const int* irows,
                         int ncols, const int* icols, int insert) {
  const double* v = (const double*)values;
  int i, j;
  for (i = 0; i < nrows; ++i)
    for (j = 0; j < ncols; ++j) oracle_set_value(mat, irows[i], icols[j], v[i * ncols + j], insert);
}

#endif
