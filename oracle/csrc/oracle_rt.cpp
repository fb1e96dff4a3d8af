// CPU oracle runtime: restatements of the reference's native helpers.
//
// TEST INFRASTRUCTURE ONLY.  This library is the checker the hip backend is
// compared against; nothing in the product links or calls it.
//
//   oracle_sparsity_seq   build_sparsity_pattern_seq, pyop2/sparsity_utils.cxx:5-51
//                         (std::set per row, sorted unique colidx)
//   oracle_addto_*        addto_scalar / addto_vector, pyop2/mat_utils.cxx:4-28,
//                         applied to a CSR matrix as PETSc MatSetValuesLocal
//                         with IGNORE_ZERO_ENTRIES and NEW_NONZERO_ALLOCATION_ERR
//                         (pyop2/petsc_base.py:110-119)
//   oracle_pcg            PETSc 3.3 KSPSolve_CG + PCJACOBI + KSPConvergedDefault
//                         (parameters from pyop2/base.py:1591-1601)
//   oracle_plan_colour    the block colouring of pyop2/plan.pyx:214-376
//                         (greedy 32-colour bit-mask rounds), used by the
//                         OpenMP baseline
// Built with -O2 -ffp-contract=off, so FP64 arithmetic is plain IEEE.
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <set>
#include <vector>

extern "C" {

// --- sparsity ---------------------------------------------------------------

struct oracle_map {
  const int* values;
  int dim;
  int from_size;
};

// Two calls: with colidx == NULL fills rowptr (lsize+1) and nnz (lsize) and
// returns the total; with colidx != NULL also fills colidx.
long long oracle_sparsity_seq(int rmult, int cmult, int nrows, int nmaps,
                              const oracle_map* rowmaps, const oracle_map* colmaps,
                              int* nnz, int* rowptr, int* colidx) {
  int lsize = nrows * rmult;
  std::vector<std::set<int> > s_diag(lsize);
  for (int m = 0; m < nmaps; m++) {
    const oracle_map& rowmap = rowmaps[m];
    const oracle_map& colmap = colmaps[m];
    int rsize = rowmap.from_size;
    for (int e = 0; e < rsize; ++e) {
      for (int i = 0; i < rowmap.dim; ++i) {
        for (int r = 0; r < rmult; r++) {
          int row = rmult * rowmap.values[i + e * rowmap.dim] + r;
          for (int d = 0; d < colmap.dim; d++) {
            for (int c = 0; c < cmult; c++) {
              s_diag[row].insert(cmult * colmap.values[d + e * colmap.dim] + c);
            }
          }
        }
      }
    }
  }
  rowptr[0] = 0;
  for (int row = 0; row < lsize; ++row) {
    nnz[row] = static_cast<int>(s_diag[row].size());
    rowptr[row + 1] = rowptr[row] + nnz[row];
  }
  if (colidx) {
    for (int row = 0; row < lsize; ++row) {
      int k = rowptr[row];
      for (std::set<int>::const_iterator it = s_diag[row].begin(); it != s_diag[row].end(); ++it)
        colidx[k++] = *it;
    }
  }
  return rowptr[lsize];
}

// --- matrix utilities ----------------------------------------------------------

}  // extern "C"
#include "oracle_mat.h"
extern "C" {

void oracle_zero_rows(oracle_csr* mat, const int* rows, int n, double diag) {
  for (int i = 0; i < n; ++i) {
    int r = rows[i];
    for (int k = mat->rowptr[r]; k < mat->rowptr[r + 1]; ++k)
      mat->vals[k] = mat->colidx[k] == r ? diag : 0.0;
  }
}

// --- CG -----------------------------------------------------------------------

static void spmv(int n, const int* rp, const int* ci, const double* v, const double* x,
                 double* y) {
  for (int r = 0; r < n; ++r) {
    double s = 0.0;
    for (int k = rp[r]; k < rp[r + 1]; ++k) s += v[k] * x[ci[k]];
    y[r] = s;
  }
}

static double dot(int n, const double* a, const double* b) {
  double s = 0.0;
  for (int i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}

// Returns the PETSc converged reason; *its and *rnorm as KSPGetIterationNumber
// and KSPGetResidualNorm; reshist (max_it+1, may be NULL) the monitor values.
int oracle_pcg(int n, const int* rp, const int* ci, const double* vals, const double* b, double* x,
               double rtol, double atol, double dtol, int max_it, int* its, double* rnorm,
               double* reshist) {
  std::vector<double> inv(n), r(n), z(n), p(n), w(n);
  for (int i = 0; i < n; ++i) {
    double d = 0.0;
    for (int k = rp[i]; k < rp[i + 1]; ++k)
      if (ci[k] == i) d = vals[k];
    inv[i] = d != 0.0 ? 1.0 / d : 1.0;
  }
  for (int i = 0; i < n; ++i) {
    x[i] = 0.0;
    r[i] = b[i];
    z[i] = inv[i] * r[i];
  }
  double dp = std::sqrt(dot(n, z.data(), z.data()));
  if (reshist) reshist[0] = dp;
  *its = 0;
  *rnorm = dp;
  double rnorm0 = dp;
  double ttol = std::fmax(rtol * rnorm0, atol);
  auto test = [&](double d) -> int {
    if (d != d) return -9;
    if (d <= ttol) return d < atol ? 3 : 2;
    if (d >= dtol * rnorm0) return -4;
    return 0;
  };
  int reason = test(dp);
  if (reason) return reason;
  double beta = dot(n, z.data(), r.data()), betaold = 0.0, dpi = 0.0, dpiold;
  int i = 0;
  do {
    *its = i + 1;
    if (beta == 0.0) return 3;
    if (i == 0) {
      for (int k = 0; k < n; ++k) p[k] = z[k];
    } else {
      double bb = beta / betaold;
      for (int k = 0; k < n; ++k) p[k] = z[k] + bb * p[k];
    }
    dpiold = dpi;
    spmv(n, rp, ci, vals, p.data(), w.data());
    dpi = dot(n, p.data(), w.data());
    if (dpi == 0.0 || (i > 0 && dpi * dpiold <= 0.0)) return -8;
    double a = beta / dpi;
    for (int k = 0; k < n; ++k) {
      x[k] = x[k] + a * p[k];
      r[k] = r[k] - a * w[k];
    }
    for (int k = 0; k < n; ++k) z[k] = inv[k] * r[k];
    dp = std::sqrt(dot(n, z.data(), z.data()));
    *rnorm = dp;
    if (reshist) reshist[i + 1] = dp;
    reason = test(dp);
    if (reason) return reason;
    betaold = beta;
    beta = dot(n, z.data(), r.data());
    i++;
  } while (i < max_it);
  return -3;
}

// --- reference-style block colouring (OpenMP baseline) -------------------------

// Contiguous blocks of part_size elements; greedy colouring in rounds of 32
// colours over the race targets (map columns, all of them).  Outputs blkcol
// (nblocks) and returns the number of colours.
int oracle_plan_colour(int nelems, int part_size, int nmaps, const int* const* maps,
                       const int* dims, const int* target_sizes, int* blkcol) {
  int nblocks = (nelems + part_size - 1) / part_size;
  std::vector<std::vector<unsigned>> masks(nmaps);
  for (int m = 0; m < nmaps; ++m) masks[m].assign(target_sizes[m], 0u);
  for (int b = 0; b < nblocks; ++b) blkcol[b] = -1;
  int base = 0, left = nblocks, ncol = 0;
  while (left > 0) {
    for (int m = 0; m < nmaps; ++m) std::fill(masks[m].begin(), masks[m].end(), 0u);
    for (int b = 0; b < nblocks; ++b) {
      if (blkcol[b] >= 0) continue;
      unsigned mask = 0;
      int e0 = b * part_size, e1 = std::min(nelems, e0 + part_size);
      for (int m = 0; m < nmaps; ++m)
        for (int e = e0; e < e1; ++e)
          for (int k = 0; k < dims[m]; ++k) mask |= masks[m][maps[m][e * dims[m] + k]];
      if (mask == 0xffffffffu) continue;
      int c = __builtin_ctz(~mask);
      blkcol[b] = base + c;
      if (base + c + 1 > ncol) ncol = base + c + 1;
      --left;
      for (int m = 0; m < nmaps; ++m)
        for (int e = e0; e < e1; ++e)
          for (int k = 0; k < dims[m]; ++k) masks[m][maps[m][e * dims[m] + k]] |= (1u << c);
    }
    base += 32;
  }
  return ncol;
}

}  // extern "C"
