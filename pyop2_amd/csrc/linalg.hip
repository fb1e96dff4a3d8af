// FP64 CSR linear algebra on the device: SpMV, Jacobi-preconditioned CG with
// PETSc KSPCG semantics, and the matrix utilities of the Mat API.
//
// Replaces the reference's PETSc calls: Mat.zero / zero_rows
// (pyop2/petsc_base.py:127-135) and Solver.solve (petsc_base.py:182-220,
// KSP 'cg' + PC 'jacobi', parameters from pyop2/base.py:1591-1601).
//
// The CG follows PETSc's KSPSolve_CG with the preconditioned norm:
//   r = b, z = D^-1 r, dp = |z|, converged test at it 0, beta = z.r
//   loop: p = z (+ beta/betaold p); w = A p; dpi = p.w; a = beta/dpi;
//         x += a p; r -= a w; z = D^-1 r; dp = |z|; test; beta = z.r
// with ttol = max(rtol*|z0|, atol), converged when dp <= ttol, diverged
// when dp >= dtol*|z0|, and DIVERGED_ITS after max_it iterations.
// Every reduction is a per-block partial followed by a single-block pass in
// fixed order, so a solve is bit-reproducible run to run.  Iterations are
// enqueued in chunks without a host round trip: the scalar state lives on
// the device and every kernel returns early once it is marked finished.
#include <cmath>
#include <vector>

#include "common.h"

namespace {

constexpr int kBlock = 256;
constexpr int kMaxPartials = 2048;

struct CgState {
  double beta, betaold, dpi, dpiold, dp, rnorm0, ttol, alpha;
  double rtol, atol, dtol;
  int its, reason, done, maxit;
};

// PETSc KSPConvergedReason values used here.
enum {
  KSP_CONVERGED_RTOL = 2,
  KSP_CONVERGED_ATOL = 3,
  KSP_CONVERGED_ITERATING = 0,
  KSP_DIVERGED_ITS = -3,
  KSP_DIVERGED_DTOL = -4,
  KSP_DIVERGED_INDEFINITE_MAT = -8,
  KSP_DIVERGED_NANORINF = -9,
};

__device__ inline double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Block-wide sum in fixed order; result valid in thread 0.
__device__ inline double block_sum(double v, double* sh) {
  v = wave_sum(v);
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0) {
    for (int k = 0; k < (int)(blockDim.x >> 6); ++k) t += sh[k];
  }
  __syncthreads();
  return t;
}

// Grid size for the streaming kernels: bounded so partials fit one block's
// final pass, and every row is covered by the grid-stride loop.
inline int stream_grid(long long n) {
  long long g = (n + kBlock - 1) / kBlock;
  if (g > kMaxPartials) g = kMaxPartials;
  if (g < 1) g = 1;
  return static_cast<int>(g);
}

// y = A x with G lanes per row (G a power of two <= 64).  When dot != null,
// also writes the block partial of sum_r w[r]*y[r].
template <int G>
__global__ void __launch_bounds__(kBlock) spmv_kernel(int n, const int* __restrict__ rowptr,
                                                      const int* __restrict__ colidx,
                                                      const double* __restrict__ vals,
                                                      const double* __restrict__ x,
                                                      double* __restrict__ y,
                                                      const double* __restrict__ w,
                                                      double* __restrict__ partial,
                                                      const CgState* __restrict__ st) {
  if (st && st->done) return;
  __shared__ double sh[kBlock / 64];
  const int sub = threadIdx.x & (G - 1);
  const long long groups_per_grid = static_cast<long long>(gridDim.x) * (kBlock / G);
  double acc_dot = 0.0;
  for (long long row = (static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x) / G; row < n;
       row += groups_per_grid) {
    int b = rowptr[row], e = rowptr[row + 1];
    double s = 0.0;
    for (int k = b + sub; k < e; k += G) s += vals[k] * x[colidx[k]];
    for (int off = G / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, G);
    if (sub == 0) {
      y[row] = s;
      if (partial) acc_dot += w[row] * s;
    }
  }
  if (partial) {
    double t = block_sum(acc_dot, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
  }
}

template <int G>
void launch_spmv_g(int grid, hipStream_t s, int n, const int* rp, const int* ci, const double* v,
                   const double* x, double* y, const double* w, double* partial,
                   const CgState* st) {
  hipLaunchKernelGGL(spmv_kernel<G>, dim3(grid), dim3(kBlock), 0, s, n, rp, ci, v, x, y, w,
                     partial, st);
}

int pick_lanes(long long n, long long nnz) {
  double avg = n > 0 ? static_cast<double>(nnz) / n : 1.0;
  int g = 1;
  while (g < 64 && g * 2 <= avg) g *= 2;
  return g;
}

void launch_spmv(int lanes, hipStream_t s, int n, const int* rp, const int* ci, const double* v,
                 const double* x, double* y, const double* w, double* partial,
                 const CgState* st, int grid_override = 0) {
  long long threads = static_cast<long long>(n) * lanes;
  long long g = (threads + kBlock - 1) / kBlock;
  if (partial && g > kMaxPartials) g = kMaxPartials;
  if (g < 1) g = 1;
  if (grid_override) g = grid_override;
  int grid = static_cast<int>(g);
  switch (lanes) {
    case 1: launch_spmv_g<1>(grid, s, n, rp, ci, v, x, y, w, partial, st); break;
    case 2: launch_spmv_g<2>(grid, s, n, rp, ci, v, x, y, w, partial, st); break;
    case 4: launch_spmv_g<4>(grid, s, n, rp, ci, v, x, y, w, partial, st); break;
    case 8: launch_spmv_g<8>(grid, s, n, rp, ci, v, x, y, w, partial, st); break;
    case 16: launch_spmv_g<16>(grid, s, n, rp, ci, v, x, y, w, partial, st); break;
    case 32: launch_spmv_g<32>(grid, s, n, rp, ci, v, x, y, w, partial, st); break;
    default: launch_spmv_g<64>(grid, s, n, rp, ci, v, x, y, w, partial, st); break;
  }
}

// inv_diag[r] = 1/A[r][r] (1 where the diagonal is missing or zero, as
// PETSc's PCJACOBI does for zero diagonal entries).
__global__ void jacobi_setup(int n, const int* __restrict__ rowptr, const int* __restrict__ colidx,
                             const double* __restrict__ vals, double* __restrict__ inv_diag) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  double d = 0.0;
  for (int k = rowptr[r]; k < rowptr[r + 1]; ++k)
    if (colidx[k] == r) d = vals[k];
  inv_diag[r] = d != 0.0 ? 1.0 / d : 1.0;
}

// Setup: x = 0, r = b, z = D^-1 r, partials of z.z and z.r.
__global__ void __launch_bounds__(kBlock) cg_init(int n, const double* __restrict__ b,
                                                  const double* __restrict__ inv_diag,
                                                  double* __restrict__ x, double* __restrict__ r,
                                                  double* __restrict__ z,
                                                  double* __restrict__ part_zz,
                                                  double* __restrict__ part_zr) {
  __shared__ double sh[kBlock / 64];
  double zz = 0.0, zr = 0.0;
  for (long long i = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x; i < n;
       i += static_cast<long long>(gridDim.x) * kBlock) {
    double ri = b[i];
    double zi = inv_diag[i] * ri;
    x[i] = 0.0;
    r[i] = ri;
    z[i] = zi;
    zz += zi * zi;
    zr += zi * ri;
  }
  double t = block_sum(zz, sh);
  if (threadIdx.x == 0) part_zz[blockIdx.x] = t;
  t = block_sum(zr, sh);
  if (threadIdx.x == 0) part_zr[blockIdx.x] = t;
}

__device__ inline double fixed_sum(const double* __restrict__ part, int np, double* sh) {
  double v = 0.0;
  for (int k = threadIdx.x; k < np; k += blockDim.x) v += part[k];
  return block_sum(v, sh);
}

__device__ inline void cg_check(CgState* st, int it) {
  double dp = st->dp;
  if (!(dp == dp) || isinf(dp)) {
    st->reason = KSP_DIVERGED_NANORINF;
    st->done = 1;
  } else if (dp <= st->ttol) {
    st->reason = dp < st->atol ? KSP_CONVERGED_ATOL : KSP_CONVERGED_RTOL;
    st->done = 1;
  } else if (dp >= st->dtol * st->rnorm0) {
    st->reason = KSP_DIVERGED_DTOL;
    st->done = 1;
  }
  (void)it;
}

// After cg_init: |z0|, ttol, the iteration-0 test, beta = z.r.
__global__ void __launch_bounds__(kBlock) cg_start(int np, const double* __restrict__ part_zz,
                                                   const double* __restrict__ part_zr,
                                                   CgState* st, double* reshist) {
  __shared__ double sh[kBlock / 64];
  double zz = fixed_sum(part_zz, np, sh);
  double zr = fixed_sum(part_zr, np, sh);
  if (threadIdx.x == 0) {
    st->dp = sqrt(zz);
    st->rnorm0 = st->dp;
    st->ttol = fmax(st->rtol * st->rnorm0, st->atol);
    st->its = 0;
    st->reason = KSP_CONVERGED_ITERATING;
    st->done = 0;
    if (reshist) reshist[0] = st->dp;
    cg_check(st, 0);
    st->beta = zr;
    st->dpi = 0.0;
  }
}

// p = z (first iteration) or p = z + (beta/betaold) p.
__global__ void __launch_bounds__(kBlock) cg_update_p(int n, const double* __restrict__ z,
                                                      double* __restrict__ p, CgState* st) {
  if (st->done) return;
  bool first = st->its == 0;
  double bb = first ? 0.0 : st->beta / st->betaold;
  for (long long i = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x; i < n;
       i += static_cast<long long>(gridDim.x) * kBlock) {
    p[i] = first ? z[i] : z[i] + bb * p[i];
  }
}

// dpi = p.w, indefinite-matrix check, alpha = beta/dpi.
__global__ void __launch_bounds__(kBlock) cg_alpha(int np, const double* __restrict__ part,
                                                   CgState* st) {
  if (st->done) return;
  __shared__ double sh[kBlock / 64];
  double pw = fixed_sum(part, np, sh);
  if (threadIdx.x == 0) {
    st->its += 1;
    if (st->beta == 0.0) {
      st->reason = KSP_CONVERGED_ATOL;
      st->done = 1;
      return;
    }
    st->dpiold = st->dpi;
    st->dpi = pw;
    if (pw == 0.0 || (st->its > 1 && st->dpi * st->dpiold <= 0.0)) {
      st->reason = KSP_DIVERGED_INDEFINITE_MAT;
      st->done = 1;
      return;
    }
    st->alpha = st->beta / pw;
  }
}

// x += a p, r -= a w, z = D^-1 r, partials of z.z and z.r.
__global__ void __launch_bounds__(kBlock) cg_axpy(int n, const double* __restrict__ p,
                                                  const double* __restrict__ w,
                                                  const double* __restrict__ inv_diag,
                                                  double* __restrict__ x, double* __restrict__ r,
                                                  double* __restrict__ z,
                                                  double* __restrict__ part_zz,
                                                  double* __restrict__ part_zr,
                                                  const CgState* st) {
  if (st->done) return;
  __shared__ double sh[kBlock / 64];
  double a = st->alpha;
  double zz = 0.0, zr = 0.0;
  for (long long i = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x; i < n;
       i += static_cast<long long>(gridDim.x) * kBlock) {
    x[i] = x[i] + a * p[i];
    double ri = r[i] - a * w[i];
    r[i] = ri;
    double zi = inv_diag[i] * ri;
    z[i] = zi;
    zz += zi * zi;
    zr += zi * ri;
  }
  double t = block_sum(zz, sh);
  if (threadIdx.x == 0) part_zz[blockIdx.x] = t;
  t = block_sum(zr, sh);
  if (threadIdx.x == 0) part_zr[blockIdx.x] = t;
}

// dp = |z|, convergence test, betaold = beta, beta = z.r.
__global__ void __launch_bounds__(kBlock) cg_finish(int np, const double* __restrict__ part_zz,
                                                    const double* __restrict__ part_zr,
                                                    CgState* st, double* reshist) {
  if (st->done) return;
  __shared__ double sh[kBlock / 64];
  double zz = fixed_sum(part_zz, np, sh);
  double zr = fixed_sum(part_zr, np, sh);
  if (threadIdx.x == 0) {
    st->dp = sqrt(zz);
    if (reshist) reshist[st->its] = st->dp;
    cg_check(st, st->its);
    if (!st->done && st->its >= st->maxit) {
      st->reason = KSP_DIVERGED_ITS;
      st->done = 1;
    }
    if (!st->done) {
      st->betaold = st->beta;
      st->beta = zr;
    }
  }
}

__global__ void zero_rows_kernel(int nrows_list, const int* __restrict__ rows,
                                 const int* __restrict__ rowptr, const int* __restrict__ colidx,
                                 double* __restrict__ vals, double diag) {
  int i = blockIdx.x;
  if (i >= nrows_list) return;
  int r = rows[i];
  for (int k = rowptr[r] + threadIdx.x; k < rowptr[r + 1]; k += blockDim.x)
    vals[k] = colidx[k] == r ? diag : 0.0;
}

}  // namespace

extern "C" {

// y = A x for an FP64 CSR matrix with n rows.
int op2hip_spmv(int n, long long nnz, const int* rowptr, const int* colidx, const double* vals,
                const double* x, double* y, void* stream) {
  OP2HIP_REQUIRE(n >= 0, "bad spmv size");
  if (n == 0) return OP2HIP_OK;
  hipStream_t s = static_cast<hipStream_t>(stream);
  launch_spmv(pick_lanes(n, nnz), s, n, rowptr, colidx, vals, x, y, nullptr, nullptr, nullptr);
  OP2HIP_CHECK(hipGetLastError());
  return OP2HIP_OK;
}

// Zero the listed rows (device int32 row ids) except their diagonal, which
// is set to `diag` (PETSc MatZeroRowsLocal with KEEP_NONZERO_PATTERN).
int op2hip_csr_zero_rows(int nlist, const int* rows, const int* rowptr, const int* colidx,
                         double* vals, double diag, void* stream) {
  if (nlist <= 0) return OP2HIP_OK;
  hipLaunchKernelGGL(zero_rows_kernel, dim3(nlist), dim3(64), 0,
                     static_cast<hipStream_t>(stream), nlist, rows, rowptr, colidx, vals, diag);
  OP2HIP_CHECK(hipGetLastError());
  return OP2HIP_OK;
}

// Jacobi-preconditioned CG solve of A x = b from x = 0.
//   work: device doubles, at least 5*n + 2*2048 + 16 (owned by the caller)
//   reshist: optional host array of max_it + 1 doubles (residual history)
//   chunk: iterations enqueued between host checks of the device state
// Outputs the iteration count, final preconditioned residual norm and the
// PETSc converged reason (>0 converged, <0 diverged).
int op2hip_pcg(int n, long long nnz, const int* rowptr, const int* colidx, const double* vals,
               const double* b, double* x, double rtol, double atol, double dtol, int maxit,
               double* work, double* reshist_host, int chunk, int* its_out, double* rnorm_out,
               int* reason_out, void* stream) {
  OP2HIP_REQUIRE(n >= 0 && maxit >= 0 && work && its_out && rnorm_out && reason_out,
                 "bad pcg arguments");
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (chunk <= 0) chunk = 16;
  double* inv_diag = work;
  double* r = inv_diag + n;
  double* z = r + n;
  double* p = z + n;
  double* w = p + n;
  double* part_a = w + n;
  double* part_b = part_a + kMaxPartials;
  CgState* st = reinterpret_cast<CgState*>(part_b + kMaxPartials);
  double* reshist_dev = nullptr;
  if (reshist_host) OP2HIP_CHECK(hipMallocAsync(&reshist_dev, sizeof(double) * (maxit + 1), s));
  CgState init = {};
  init.rtol = rtol;
  init.atol = atol;
  init.dtol = dtol;
  init.maxit = maxit;
  OP2HIP_CHECK(hipMemcpyAsync(st, &init, sizeof(CgState), hipMemcpyHostToDevice, s));
  int lanes = pick_lanes(n, nnz);
  int g = stream_grid(n);
  long long sp_threads = static_cast<long long>(n) * lanes;
  long long sp_grid = (sp_threads + kBlock - 1) / kBlock;
  if (sp_grid > kMaxPartials) sp_grid = kMaxPartials;
  if (sp_grid < 1) sp_grid = 1;
  if (n > 0) {
    hipLaunchKernelGGL(jacobi_setup, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, n,
                       rowptr, colidx, vals, inv_diag);
  }
  hipLaunchKernelGGL(cg_init, dim3(g), dim3(kBlock), 0, s, n, b, inv_diag, x, r, z, part_a,
                     part_b);
  hipLaunchKernelGGL(cg_start, dim3(1), dim3(kBlock), 0, s, g, part_a, part_b, st, reshist_dev);
  OP2HIP_CHECK(hipGetLastError());
  CgState h = {};
  int enqueued = 0;
  while (true) {
    OP2HIP_CHECK(hipMemcpyAsync(&h, st, sizeof(CgState), hipMemcpyDeviceToHost, s));
    OP2HIP_CHECK(hipStreamSynchronize(s));
    if (h.done) break;
    if (enqueued >= maxit) {
      // should have been caught on the device; keep PETSc's reason
      h.reason = KSP_DIVERGED_ITS;
      break;
    }
    int k = chunk;
    if (enqueued + k > maxit) k = maxit - enqueued;
    for (int it = 0; it < k; ++it) {
      hipLaunchKernelGGL(cg_update_p, dim3(g), dim3(kBlock), 0, s, n, z, p, st);
      launch_spmv(lanes, s, n, rowptr, colidx, vals, p, w, p, part_a, st,
                  static_cast<int>(sp_grid));
      hipLaunchKernelGGL(cg_alpha, dim3(1), dim3(kBlock), 0, s, static_cast<int>(sp_grid), part_a,
                         st);
      hipLaunchKernelGGL(cg_axpy, dim3(g), dim3(kBlock), 0, s, n, p, w, inv_diag, x, r, z, part_a,
                         part_b, st);
      hipLaunchKernelGGL(cg_finish, dim3(1), dim3(kBlock), 0, s, g, part_a, part_b, st,
                         reshist_dev);
    }
    OP2HIP_CHECK(hipGetLastError());
    enqueued += k;
    if (maxit == 0) break;
  }
  if (maxit == 0 && !h.done) h.reason = KSP_DIVERGED_ITS;
  if (reshist_host && reshist_dev) {
    int nh = h.its + 1;
    if (nh > maxit + 1) nh = maxit + 1;
    OP2HIP_CHECK(hipMemcpyAsync(reshist_host, reshist_dev, sizeof(double) * nh,
                                hipMemcpyDeviceToHost, s));
    OP2HIP_CHECK(hipStreamSynchronize(s));
    OP2HIP_CHECK(hipFreeAsync(reshist_dev, s));
  }
  *its_out = h.its;
  *rnorm_out = h.dp;
  *reason_out = h.reason;
  return OP2HIP_OK;
}

// Size in doubles of the work array op2hip_pcg needs for n rows.
long long op2hip_pcg_work_size(int n) {
  return 5LL * n + 2LL * kMaxPartials + (sizeof(CgState) + sizeof(double) - 1) / sizeof(double) +
         1;
}

}  // extern "C"
