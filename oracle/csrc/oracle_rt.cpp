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
//   oracle_gmres          PETSc 3.3 KSPGMRES (classical Gram-Schmidt, restart)
//   oracle_bcgs           PETSc 3.3 KSPBCGS, all three with left PCJACOBI
//                         (jacobi 1) or PCNONE (jacobi 0), the last argument
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
               double* reshist, int jacobi) {
  std::vector<double> inv(n), r(n), z(n), p(n), w(n);
  for (int i = 0; i < n; ++i) {
    double d = 0.0;
    for (int k = rp[i]; k < rp[i + 1]; ++k)
      if (ci[k] == i) d = vals[k];
    inv[i] = (jacobi && d != 0.0) ? 1.0 / d : 1.0;
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
    if (d != d || std::isinf(d)) return -9;
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

// --- GMRES and BiCGStab -----------------------------------------------------------
//
// The reference hands linear_solver to PETSc's KSPSetType (pyop2/petsc_base.py:172-176):
// 'gmres' is PETSc 3.3 KSPGMRES (restart 30, the reference's gmres_restart
// default, pyop2/base.py:1591-1601; classical Gram-Schmidt without
// refinement; happy-breakdown tolerance 1e-30), 'bcgs' KSPBCGS, both with
// PCJACOBI on the left and the preconditioned residual norm.  KSPConvergedDefault as in
// oracle_pcg, with rnorm0 the first (iteration 0) norm.

static int ksp_test(int k, double rn, double rtol, double atol, double dtol, double& rnorm0,
                    double& ttol) {
  if (k == 0) {
    rnorm0 = rn;
    ttol = std::fmax(rtol * rnorm0, atol);
  }
  if (rn != rn || std::isinf(rn)) return -9;
  if (rn <= ttol) return rn < atol ? 3 : 2;
  if (rn >= dtol * rnorm0) return -4;
  return 0;
}

// PCJACOBI's inverse diagonal (1 where it is zero), or 1 everywhere (PCNONE)
static void jacobi_inverse(int n, const int* rp, const int* ci, const double* vals, double* inv,
                           int jacobi) {
  for (int i = 0; i < n; ++i) {
    double d = 0.0;
    for (int k = rp[i]; k < rp[i + 1]; ++k)
      if (ci[k] == i) d = vals[k];
    inv[i] = (jacobi && d != 0.0) ? 1.0 / d : 1.0;
  }
}

// KSPSolve_GMRES / KSPGMRESCycle / KSPGMRESClassicalGramSchmidtOrthogonalization /
// KSPGMRESUpdateHessenberg / KSPGMRESBuildSoln.  reshist as in oracle_pcg: one
// entry per iteration count, the restart's true residual replacing the last
// estimate of the cycle before.
int oracle_gmres(int n, const int* rp, const int* ci, const double* vals, const double* b,
                 double* x, int restart, double rtol, double atol, double dtol, int max_it,
                 int* its, double* rnorm, double* reshist, int jacobi) {
  const int m = restart;
  const double haptol = 1e-30;
  std::vector<double> inv(n), tmp(n), V((size_t)(m + 1) * n);
  std::vector<double> H((size_t)(m + 1) * (m > 0 ? m : 1)), cc(m + 1), ss(m + 1), grs(m + 2),
      h(m + 1);
  auto HH = [&](int i, int j) -> double& { return H[(size_t)j * (m + 1) + i]; };
  auto col = [&](int j) { return V.data() + (size_t)j * n; };
  jacobi_inverse(n, rp, ci, vals, inv.data(), jacobi);
  for (int i = 0; i < n; ++i) x[i] = 0.0;
  int reason = 0;
  double rnorm0 = 0.0, ttol = 0.0, res = 0.0;
  *its = 0;
  while (!reason) {
    // v0 = D^-1 (b - A x), normalised
    spmv(n, rp, ci, vals, x, tmp.data());
    double* v0 = col(0);
    for (int i = 0; i < n; ++i) v0[i] = inv[i] * (b[i] - tmp[i]);
    res = std::sqrt(dot(n, v0, v0));
    if (res != 0.0)
      for (int i = 0; i < n; ++i) v0[i] *= 1.0 / res;
    grs[0] = res;
    *rnorm = res;
    if (reshist) reshist[*its] = res;
    if (res == 0.0) {
      reason = 3;
      break;
    }
    reason = ksp_test(*its, res, rtol, atol, dtol, rnorm0, ttol);
    int it = 0;
    bool hapend = false;
    while (!reason && it < m && *its < max_it) {
      double* w = col(it + 1);
      spmv(n, rp, ci, vals, col(it), tmp.data());
      for (int i = 0; i < n; ++i) w[i] = inv[i] * tmp[i];
      // classical Gram-Schmidt: h = V^T w; w -= V h
      for (int j = 0; j <= it; ++j) h[j] = dot(n, w, col(j));
      for (int j = 0; j <= it; ++j) h[j] = -h[j];
      for (int j = 0; j <= it; ++j) {
        const double* vj = col(j);
        for (int i = 0; i < n; ++i) w[i] = w[i] + h[j] * vj[i];
      }
      for (int j = 0; j <= it; ++j) HH(j, it) = 0.0 - h[j];
      double tt = std::sqrt(dot(n, w, w));
      if (tt != 0.0)
        for (int i = 0; i < n; ++i) w[i] *= 1.0 / tt;
      HH(it + 1, it) = tt;
      double hapbnd = std::fabs(tt / grs[it]);
      if (hapbnd > haptol) hapbnd = haptol;
      if (tt < hapbnd) hapend = true;
      // previous rotations on the new column, then a new one
      for (int j = 1; j <= it; ++j) {
        double t0 = HH(j - 1, it);
        HH(j - 1, it) = cc[j - 1] * t0 + ss[j - 1] * HH(j, it);
        HH(j, it) = cc[j - 1] * HH(j, it) - ss[j - 1] * t0;
      }
      if (!hapend) {
        double t = std::sqrt(HH(it, it) * HH(it, it) + HH(it + 1, it) * HH(it + 1, it));
        if (t == 0.0) {
          reason = -2;
        } else {
          cc[it] = HH(it, it) / t;
          ss[it] = HH(it + 1, it) / t;
          grs[it + 1] = -(ss[it] * grs[it]);
          grs[it] = cc[it] * grs[it];
          HH(it, it) = cc[it] * HH(it, it) + ss[it] * HH(it + 1, it);
          res = std::fabs(grs[it + 1]);
        }
      } else {
        res = 0.0;
      }
      ++it;
      ++*its;
      *rnorm = res;
      if (reshist) reshist[*its] = res;
      if (reason) break;
      reason = ksp_test(*its, res, rtol, atol, dtol, rnorm0, ttol);
      if (hapend && !reason) {
        reason = -5;
        break;
      }
    }
    // x += V y, H y = g by back substitution
    int top = it - 1;
    if (top >= 0) {
      bool ok = true;
      if (HH(top, top) != 0.0) {
        grs[top] = grs[top] / HH(top, top);
      } else {
        reason = -5;
        ok = false;
      }
      for (int k = top - 1; ok && k >= 0; --k) {
        double t = grs[k];
        for (int j = k + 1; j <= top; ++j) t = t - HH(k, j) * grs[j];
        if (HH(k, k) == 0.0) {
          reason = -5;
          ok = false;
          break;
        }
        grs[k] = t / HH(k, k);
      }
      if (ok) {
        for (int i = 0; i < n; ++i) tmp[i] = 0.0;
        for (int j = 0; j <= top; ++j) {
          const double* vj = col(j);
          for (int i = 0; i < n; ++i) tmp[i] = tmp[i] + grs[j] * vj[i];
        }
        for (int i = 0; i < n; ++i) x[i] = x[i] + tmp[i];
      }
    }
    if (*its >= max_it) {
      if (!reason) reason = -3;
      break;
    }
  }
  return reason;
}

// KSPSolve_BCGS (left PCJACOBI, preconditioned norm).
int oracle_bcgs(int n, const int* rp, const int* ci, const double* vals, const double* b,
                double* x, double rtol, double atol, double dtol, int max_it, int* its,
                double* rnorm, double* reshist, int jacobi) {
  std::vector<double> inv(n), R(n), RP(n), V(n, 0.0), T(n), S(n), P(n, 0.0), tmp(n);
  jacobi_inverse(n, rp, ci, vals, inv.data(), jacobi);
  for (int i = 0; i < n; ++i) {
    x[i] = 0.0;
    R[i] = inv[i] * b[i];
  }
  double dp = std::sqrt(dot(n, R.data(), R.data()));
  double rnorm0 = 0.0, ttol = 0.0;
  *its = 0;
  *rnorm = dp;
  if (reshist) reshist[0] = dp;
  int reason = ksp_test(0, dp, rtol, atol, dtol, rnorm0, ttol);
  if (reason) return reason;
  for (int i = 0; i < n; ++i) RP[i] = R[i];
  double rhoold = 1.0, alpha = 1.0, omegaold = 1.0, rho, beta, omega, d1, d2;
  int i = 0;
  do {
    rho = dot(n, R.data(), RP.data());
    beta = (rho / rhoold) * (alpha / omegaold);
    const double wb = -omegaold * beta;
    for (int k = 0; k < n; ++k) P[k] = R[k] + wb * V[k] + beta * P[k];
    spmv(n, rp, ci, vals, P.data(), tmp.data());
    for (int k = 0; k < n; ++k) V[k] = inv[k] * tmp[k];
    d1 = dot(n, V.data(), RP.data());
    if (d1 == 0.0) {
      reason = -5;
      break;
    }
    alpha = rho / d1;
    for (int k = 0; k < n; ++k) S[k] = -alpha * V[k] + R[k];
    spmv(n, rp, ci, vals, S.data(), tmp.data());
    for (int k = 0; k < n; ++k) T[k] = inv[k] * tmp[k];
    d1 = dot(n, S.data(), T.data());
    d2 = dot(n, T.data(), T.data());
    if (d2 == 0.0) {
      // t = 0: if s = 0 too, alpha p finishes the solve
      if (dot(n, S.data(), S.data()) != 0.0) {
        reason = -5;
        break;
      }
      for (int k = 0; k < n; ++k) x[k] = x[k] + alpha * P[k];
      ++*its;
      *rnorm = 0.0;
      if (reshist) reshist[*its] = 0.0;
      reason = 2;
      break;
    }
    omega = d1 / d2;
    for (int k = 0; k < n; ++k) x[k] = alpha * P[k] + omega * S[k] + x[k];
    for (int k = 0; k < n; ++k) R[k] = -omega * T[k] + S[k];
    dp = std::sqrt(dot(n, R.data(), R.data()));
    rhoold = rho;
    omegaold = omega;
    ++*its;
    *rnorm = dp;
    if (reshist) reshist[*its] = dp;
    reason = ksp_test(i + 1, dp, rtol, atol, dtol, rnorm0, ttol);
    if (reason) break;
    if (rho == 0.0) {
      reason = -5;
      break;
    }
    i++;
  } while (i < max_it);
  if (!reason && i >= max_it) reason = -3;
  return reason;
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
