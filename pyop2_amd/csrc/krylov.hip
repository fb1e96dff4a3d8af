// Krylov solvers for one FP64 CSR system on the device: PETSc 3.3's KSPGMRES
// (restarted, classical Gram-Schmidt without refinement) and KSPBCGS, both
// with PCJACOBI applied on the left and the preconditioned residual norm in
// KSPConvergedDefault.  The reference hands linear_solver to KSPSetType
// (pyop2/petsc_base.py:172-176; parameters pyop2/base.py:1591-1601, restart
// 30); its cusp backend offers the same methods (pyop2/cuda.py:504-508).
//
// The solver's scalars and flags live in a device state block, so
// iterations are enqueued without host round trips.  Row kernels (SpMV with
// the Jacobi scaling, dot partials, vector updates) only read the state;
// one-workgroup kernels sum the partials in workgroup order and advance it
// (Hessenberg column, Givens rotations, convergence test, phase).  Every
// kernel returns at once outside its phase, so the host enqueues iterations
// in chunks and reads the state between chunks.  Sums run in a fixed order:
// a solve is deterministic for a given n.
//
// GMRES keeps the Krylov basis as (restart + 1) columns of n doubles.  A new
// column is normalised where it is next read: the SpMV of the next step
// scales what it gathers by 1/|w| (the product PETSc's VecNormalize stores),
// and the Gram-Schmidt dot kernel stores the scaled column.
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "common.h"
#include "reduce.h"

namespace {

using op2hip::block_sum;

constexpr int kBlock = 256;
constexpr int kRBlock = 1024;
constexpr int kRows = 4;         // rows per thread of the row kernels
constexpr int kTile = kBlock * kRows;  // rows per workgroup (and per partial)
constexpr int kWaves = kBlock / 64;
constexpr int kLanes = 16;       // most lanes per SpMV row
constexpr int kMaxSpmvBlocks = 2048;
constexpr int kChunk = 32;       // most GMRES steps per round (a cycle, up to 32)
constexpr int kBcgsChunk = 16;   // BiCGStab iterations per round
constexpr int kMaxRestart = 1024;  // LDS of k_g_dots and k_g_rot
constexpr size_t kSolveLds = 60 * 1024;  // k_g_solve stages H in LDS up to this
constexpr double kHapTol = 1e-30;

enum { METHOD_GMRES = 1, METHOD_BCGS = 2 };

// PH_STEP: iterations run; PH_END: a GMRES cycle is over (solution update
// next); PH_RESTART: a residual is to be formed; PH_FINAL_X: BiCGStab's
// t = 0 exit (x += alpha p, then done); PH_DONE: finished.
enum { PH_STEP = 0, PH_END = 1, PH_RESTART = 2, PH_DONE = 3, PH_FINAL_X = 4 };

enum {
  KSP_CONVERGED_RTOL = 2,
  KSP_CONVERGED_ATOL = 3,
  KSP_DIVERGED_NULL = -2,
  KSP_DIVERGED_ITS = -3,
  KSP_DIVERGED_DTOL = -4,
  KSP_DIVERGED_BREAKDOWN = -5,
  KSP_DIVERGED_NANORINF = -9,
};

struct State {
  int phase, its, reason;
  int it;   // GMRES: Arnoldi steps in the current cycle
  int upd;  // GMRES: the cycle's correction V y is to be added to x
  int pad_;
  double rtol, atol, dtol, rnorm0, ttol, res;
  double scale;  // GMRES: factor normalising basis column `it`
  double rho, rhoold, alpha, omega, omegaold, beta;  // BiCGStab
};
static_assert(sizeof(State) <= 32 * sizeof(double), "state block");

struct Prob {
  const int* rowptr;
  const int* colidx;
  const double* vals;
  const double* b;
  double* x;
  double* inv;
  State* st;
  double* part;  // partials by quantity: quantity k of workgroup b at k * ps + b
  double* hist;  // residual history (hist_len) or null
  int n, lanes, ps, m, maxit, hist_len;  // ps = max(nb, nbs)
  int jacobi;  // 1: D^-1 is the inverse diagonal (PCJACOBI); 0: I (PCNONE)
  int nb;   // workgroups of the tile and element kernels (kTile rows each)
  int nbs;  // workgroups of the SpMV kernels (more, for gather latency)
  // GMRES
  double* V;    // m + 1 columns of n
  double* H;    // (m + 1) x m, column j at H + j * (m + 1)
  double* cc;
  double* ss;
  double* grs;  // rotated right-hand side; the coefficients y after the solve
  double* lhh;  // minus the Gram-Schmidt coefficients of the current step
  // BiCGStab
  double *R, *RP, *Vb, *T, *S, *P;
};

// KSPConvergedDefault's test of rn once rnorm0 and ttol are set: 0 to go
// on, else the reason
__device__ inline int ksp_test_rn(double rn, double ttol, double atol, double dtol,
                                  double rnorm0) {
  if (!(rn == rn) || isinf(rn)) return KSP_DIVERGED_NANORINF;
  if (rn <= ttol) return rn < atol ? KSP_CONVERGED_ATOL : KSP_CONVERGED_RTOL;
  if (rn >= dtol * rnorm0) return KSP_DIVERGED_DTOL;
  return 0;
}

__device__ inline int ksp_test(State& s, double rn) {
  if (s.its == 0) {
    s.rnorm0 = rn;
    s.ttol = fmax(s.rtol * rn, s.atol);
  }
  return ksp_test_rn(rn, s.ttol, s.atol, s.dtol, s.rnorm0);
}

__device__ inline void log_res(const Prob& p, int k, double v) {
  if (p.hist && k < p.hist_len) p.hist[k] = v;
}

// sums of partials k0 .. k0+K-1 over the nparts workgroups of a row
// kernel, in one pass (one workgroup of kRBlock; sh holds K * kRBlock / 64
// doubles); the results are valid in thread 0.  A thread adds its
// partials b = tid, tid + kRBlock, ... in that order, loading four of each
// quantity before the first wait: a partial past the end loads the last
// one (b < nparts, so nparts >= 1) and adds +0.0 instead, which leaves a
// sum that starts at +0.0 unchanged bit for bit (such a sum is never -0.0).
template <int K>
__device__ inline void reduce_parts(const Prob& p, int nparts, int k0, double* out, double* sh) {
  constexpr int kBatch = 4;
  double v[K];
#pragma unroll
  for (int q = 0; q < K; ++q) v[q] = 0.0;
  const double* __restrict__ base = p.part + (size_t)k0 * p.ps;
  for (int b0 = threadIdx.x; b0 < nparts; b0 += kBatch * kRBlock) {
    double t[kBatch][K];
#pragma unroll
    for (int r = 0; r < kBatch; ++r) {
      const int b = b0 + r * kRBlock;
      const int bc = b < nparts ? b : nparts - 1;
#pragma unroll
      for (int q = 0; q < K; ++q) t[r][q] = base[(size_t)q * p.ps + bc];
    }
#pragma unroll
    for (int r = 0; r < kBatch; ++r) {
      const bool in = b0 + r * kRBlock < nparts;
#pragma unroll
      for (int q = 0; q < K; ++q) v[q] += in ? t[r][q] : 0.0;
    }
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < K; ++q) {
    v[q] = op2hip::wave_sum(v[q]);
    if (lane == 0) sh[q * (kRBlock / 64) + w] = v[q];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < K; ++q) {
      double t = 0.0;
      for (int k = 0; k < kRBlock / 64; ++k) t += sh[q * (kRBlock / 64) + k];
      out[q] = t;
    }
  }
  __syncthreads();
}

__device__ inline double reduce_part(const Prob& p, int nparts, int k, double* sh) {
  double out[1];
  reduce_parts<1>(p, nparts, k, out, sh);
  return out[0];
}

__device__ inline void store_part(const Prob& p, int k, double v, double* sh) {
  double t = block_sum<kBlock>(v, sh);
  if (threadIdx.x == 0) p.part[(size_t)k * p.ps + blockIdx.x] = t;
}

#define ROWS(i) for (int i = blockIdx.x * kBlock + threadIdx.x; i < p.n; i += p.nb * kBlock)

// SpMV rows: `lanes` threads per row, grid-stride over row groups of the
// nbs workgroups.  finish(row, sum) runs on the row's first lane;
// gather(col) is the gathered operand.
template <class G, class F>
__device__ inline void spmv_rows(const Prob& p, G gather, F finish) {
  const int lanes = p.lanes;
  const int rpb = kBlock / lanes;
  const int sub = threadIdx.x & (lanes - 1);
  const int* __restrict__ rowptr = p.rowptr;
  const int* __restrict__ colidx = p.colidx;
  const double* __restrict__ vals = p.vals;
  for (int row = blockIdx.x * rpb + (int)threadIdx.x / lanes; row < p.n; row += p.nbs * rpb) {
    const int kb = rowptr[row], ke = rowptr[row + 1];
    double v = 0.0;
    for (int k = kb + sub; k < ke; k += lanes) v += vals[k] * gather(colidx[k]);
    for (int off = lanes / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, lanes);
    if (sub == 0) finish(row, v);
  }
}

// --- shared --------------------------------------------------------------------

// Jacobi inverse diagonal (1 where the diagonal is zero or absent; 1
// everywhere without a preconditioner) and x = 0
__global__ void __launch_bounds__(kBlock) k_init(Prob p) {
  const int* __restrict__ rowptr = p.rowptr;
  const int* __restrict__ colidx = p.colidx;
  ROWS(i) {
    double d = 0.0;
    for (int k = rowptr[i]; k < rowptr[i + 1]; ++k)
      if (colidx[k] == i) d = p.vals[k];
    p.inv[i] = (p.jacobi && d != 0.0) ? 1.0 / d : 1.0;
    p.x[i] = 0.0;
  }
}

// out = D^-1 (b - A x) (and a copy in out2), partial |out|^2; phase RESTART
__global__ void __launch_bounds__(kBlock) k_resid(Prob p, double* out, double* out2) {
  __shared__ double sh[kBlock / 64];
  if (p.st->phase != PH_RESTART) return;
  const double* __restrict__ x = p.x;
  double acc = 0.0;
  spmv_rows(p, [&](int c) { return x[c]; }, [&](int row, double ax) {
    const double r = p.inv[row] * (p.b[row] - ax);
    out[row] = r;
    if (out2) out2[row] = r;
    acc += r * r;
  });
  store_part(p, 0, acc, sh);
}

// --- GMRES -------------------------------------------------------------------------

// cycle start: |v0|, the iteration-`its` test (sets rnorm0 and ttol at 0)
__global__ void __launch_bounds__(kRBlock) k_g_start(Prob p) {
  __shared__ double sh[kRBlock / 64];
  State& s = *p.st;
  if (s.phase != PH_RESTART) return;
  const double rr = reduce_part(p, p.nbs, 0, sh);
  if (threadIdx.x != 0) return;
  const double res = sqrt(rr);
  s.res = res;
  p.grs[0] = res;
  s.scale = res != 0.0 ? 1.0 / res : 1.0;
  s.it = 0;
  s.upd = 0;
  log_res(p, s.its, res);
  if (res == 0.0) {
    s.reason = KSP_CONVERGED_ATOL;
    s.phase = PH_END;
    return;
  }
  s.reason = ksp_test(s, res);
  s.phase = (s.reason || p.m <= 0 || s.its >= p.maxit) ? PH_END : PH_STEP;
}

// w = v_{j+1} = D^-1 A v_j, with v_j scaled as it is gathered
__global__ void __launch_bounds__(kBlock) k_g_spmv(Prob p) {
  const State& s = *p.st;
  if (s.phase != PH_STEP) return;
  const int j = s.it;
  const double sc = s.scale;
  const double* __restrict__ vj = p.V + (size_t)j * p.n;
  double* __restrict__ w = p.V + (size_t)(j + 1) * p.n;
  spmv_rows(p, [&](int c) { return vj[c] * sc; },
            [&](int row, double av) { w[row] = p.inv[row] * av; });
}

// partials of w . v_k for k <= j in one pass over the workgroup's tile
// (w held in registers, one wavefront sum per column into LDS); stores the
// scaled v_j.  Dynamic LDS: (m + 1) * kWaves doubles.
__global__ void __launch_bounds__(kBlock) k_g_dots(Prob p) {
  extern __shared__ double dsh[];
  const State& s = *p.st;
  if (s.phase != PH_STEP) return;
  const int j = s.it;
  const double sc = s.scale;
  const size_t n = p.n;
  const double* __restrict__ w = p.V + (size_t)(j + 1) * n;
  double* __restrict__ V = p.V;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int base = blockIdx.x * kTile + threadIdx.x;
  double wv[kRows];
#pragma unroll
  for (int r = 0; r < kRows; ++r) {
    const int i = base + r * kBlock;
    wv[r] = i < p.n ? w[i] : 0.0;
  }
  // columns four at a time, so that 4 * kRows loads are in flight
  int k = 0;
  for (; k + 4 <= j; k += 4) {
    double a[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int r = 0; r < kRows; ++r) {
      const int i = base + r * kBlock;
      if (i < p.n) {
#pragma unroll
        for (int q = 0; q < 4; ++q) a[q] += wv[r] * V[(k + q) * n + i];
      }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const double t = op2hip::wave_sum(a[q]);
      if (lane == 0) dsh[(k + q) * kWaves + wave] = t;
    }
  }
  for (; k < j; ++k) {
    const double* __restrict__ vk = V + k * n;
    double a = 0.0;
#pragma unroll
    for (int r = 0; r < kRows; ++r) {
      const int i = base + r * kBlock;
      if (i < p.n) a += wv[r] * vk[i];
    }
    a = op2hip::wave_sum(a);
    if (lane == 0) dsh[k * kWaves + wave] = a;
  }
  {
    double* __restrict__ vj = V + j * n;
    double a = 0.0;
#pragma unroll
    for (int r = 0; r < kRows; ++r) {
      const int i = base + r * kBlock;
      if (i < p.n) {
        const double v = vj[i] * sc;
        vj[i] = v;
        a += wv[r] * v;
      }
    }
    a = op2hip::wave_sum(a);
    if (lane == 0) dsh[j * kWaves + wave] = a;
  }
  __syncthreads();
  for (int c = threadIdx.x; c <= j; c += kBlock) {
    double t = 0.0;
    for (int q = 0; q < kWaves; ++q) t += dsh[c * kWaves + q];
    p.part[(size_t)c * p.ps + blockIdx.x] = t;
  }
}

// Gram-Schmidt coefficients h_k = w . v_k into column j of H; lhh = -h.
// One wavefront per column, all columns at once.  A lane adds its partials
// b = lane, lane + 64, ... in that order, loading 16 of them before the
// first wait: a partial past the end loads the last one (b < nb, so nb >= 1)
// and adds +0.0 instead, which leaves a sum that starts at +0.0 unchanged
// bit for bit (such a sum is never -0.0).
__global__ void __launch_bounds__(kRBlock) k_g_coef(Prob p) {
  constexpr int kBatch = 16;
  const State& s = *p.st;
  if (s.phase != PH_STEP) return;
  const int j = s.it;
  const int lane = threadIdx.x & 63;
  const int nb = p.nb;
  for (int k = threadIdx.x >> 6; k <= j; k += kRBlock / 64) {
    double v = 0.0;
    const double* __restrict__ pk = p.part + (size_t)k * p.ps;
    for (int b0 = lane; b0 < nb; b0 += kBatch * 64) {
      double t[kBatch];
#pragma unroll
      for (int q = 0; q < kBatch; ++q) {
        const int b = b0 + q * 64;
        t[q] = pk[b < nb ? b : nb - 1];
      }
#pragma unroll
      for (int q = 0; q < kBatch; ++q) v += b0 + q * 64 < nb ? t[q] : 0.0;
    }
    v = op2hip::wave_sum(v);
    if (lane == 0) {
      p.lhh[k] = -v;
      p.H[(size_t)j * (p.m + 1) + k] = 0.0 - (-v);
    }
  }
}

// w -= sum_k h_k v_k over the workgroup's tile, partial |w|^2
__global__ void __launch_bounds__(kBlock) k_g_orth(Prob p) {
  __shared__ double sh[kBlock / 64];
  const State& s = *p.st;
  if (s.phase != PH_STEP) return;
  const int j = s.it;
  const size_t n = p.n;
  double* __restrict__ w = p.V + (size_t)(j + 1) * n;
  const double* __restrict__ V = p.V;
  const double* __restrict__ lhh = p.lhh;
  const int base = blockIdx.x * kTile + threadIdx.x;
  double wv[kRows];
#pragma unroll
  for (int r = 0; r < kRows; ++r) {
    const int i = base + r * kBlock;
    wv[r] = i < p.n ? w[i] : 0.0;
  }
  // columns four at a time (loads issued together, applied in column order)
  int k = 0;
  for (; k + 4 <= j + 1; k += 4) {
    double v[4][kRows];
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
      for (int r = 0; r < kRows; ++r) {
        const int i = base + r * kBlock;
        v[q][r] = i < p.n ? V[(k + q) * n + i] : 0.0;
      }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const double c = lhh[k + q];
#pragma unroll
      for (int r = 0; r < kRows; ++r) wv[r] = wv[r] + c * v[q][r];
    }
  }
  for (; k <= j; ++k) {
    const double c = lhh[k];
    const double* __restrict__ vk = V + k * n;
#pragma unroll
    for (int r = 0; r < kRows; ++r) {
      const int i = base + r * kBlock;
      if (i < p.n) wv[r] = wv[r] + c * vk[i];
    }
  }
  double acc = 0.0;
#pragma unroll
  for (int r = 0; r < kRows; ++r) {
    const int i = base + r * kBlock;
    if (i < p.n) {
      w[i] = wv[r];
      acc += wv[r] * wv[r];
    }
  }
  store_part(p, 0, acc, sh);
}

// |w|, happy breakdown, Givens rotations of the new column, the residual
// estimate and the test (KSPGMRESCycle / KSPGMRESUpdateHessenberg).  The
// column and the rotations are staged in LDS for the serial chain.
// Dynamic LDS: 3 * (m + 1) doubles.
__global__ void __launch_bounds__(kRBlock) k_g_rot(Prob p) {
  __shared__ double sh[kRBlock / 64];
  extern __shared__ double rsh[];
  State& s = *p.st;
  if (s.phase != PH_STEP) return;
  const int j = s.it;
  const int ld = p.m + 1;
  double* hh = rsh;            // column j, j + 2 entries
  double* cs = rsh + ld;       // rotations 0 .. j-1
  double* sn = rsh + 2 * ld;
  double* Hj = p.H + (size_t)j * ld;
  for (int k = threadIdx.x; k <= j; k += kRBlock) {
    hh[k] = Hj[k];
    if (k < j) {
      cs[k] = p.cc[k];
      sn[k] = p.ss[k];
    }
  }
  double ww[1];
  reduce_parts<1>(p, p.nb, 0, ww, sh);  // synchronises the workgroup
  if (threadIdx.x != 0) return;
  const double tt = sqrt(ww[0]);
  hh[j + 1] = tt;
  const double g = p.grs[j];
  double hapbnd = fabs(tt / g);
  if (hapbnd > kHapTol) hapbnd = kHapTol;
  const bool hapend = tt < hapbnd;
  for (int k = 1; k <= j; ++k) {
    const double t0 = hh[k - 1];
    hh[k - 1] = cs[k - 1] * t0 + sn[k - 1] * hh[k];
    hh[k] = cs[k - 1] * hh[k] - sn[k - 1] * t0;
  }
  double res = s.res;
  if (!hapend) {
    const double t = sqrt(hh[j] * hh[j] + hh[j + 1] * hh[j + 1]);
    if (t == 0.0) {
      s.reason = KSP_DIVERGED_NULL;
    } else {
      const double c = hh[j] / t, sj = hh[j + 1] / t;
      p.cc[j] = c;
      p.ss[j] = sj;
      p.grs[j + 1] = -(sj * g);
      p.grs[j] = c * g;
      hh[j] = c * hh[j] + sj * hh[j + 1];
      res = fabs(p.grs[j + 1]);
    }
  } else {
    res = 0.0;
  }
  for (int k = 0; k <= j + 1; ++k) Hj[k] = hh[k];
  s.it = j + 1;
  s.its += 1;
  s.res = res;
  log_res(p, s.its, res);
  if (!s.reason) {
    s.reason = ksp_test(s, res);
    if (hapend && !s.reason) s.reason = KSP_DIVERGED_BREAKDOWN;
  }
  s.scale = tt != 0.0 ? 1.0 / tt : 1.0;
  if (s.reason || s.it >= p.m || s.its >= p.maxit) s.phase = PH_END;
}

// cycle end: y from the triangular system in grs (KSPGMRESBuildSoln), then
// DONE or RESTART.  The triangle and the right-hand side are staged in LDS
// for the serial back substitution when (m + 1)^2 doubles fit in
// kSolveLds (the launch then gives that much dynamic LDS).
__device__ inline bool solve_staged(int m) {
  return sizeof(double) * (size_t)(m + 1) * (m + 1) <= kSolveLds;
}

__global__ void __launch_bounds__(kBlock) k_g_solve(Prob p) {
  extern __shared__ double lds[];
  State& s = *p.st;
  if (s.phase != PH_END) return;
  const int top = s.it - 1;
  const int ld = p.m + 1;
  const bool staged = solve_staged(p.m);
  const double* tri = staged ? lds : p.H;
  double* g = staged ? lds + (size_t)ld * p.m : p.grs;  // after the m columns of H
  if (staged) {
    for (int k = threadIdx.x; k < (top + 1) * ld; k += kBlock) lds[k] = p.H[k];
    for (int k = threadIdx.x; k <= top; k += kBlock) g[k] = p.grs[k];
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  s.upd = 0;
  if (top >= 0) {
    bool ok = true;
    if (tri[top * ld + top] != 0.0) {
      g[top] = g[top] / tri[top * ld + top];
    } else {
      s.reason = KSP_DIVERGED_BREAKDOWN;
      ok = false;
    }
    for (int k = top - 1; ok && k >= 0; --k) {
      double t = g[k];
      for (int j = k + 1; j <= top; ++j) t = t - tri[j * ld + k] * g[j];
      if (tri[k * ld + k] == 0.0) {
        s.reason = KSP_DIVERGED_BREAKDOWN;
        ok = false;
        break;
      }
      g[k] = t / tri[k * ld + k];
    }
    if (ok) {
      if (staged)
        for (int k = 0; k <= top; ++k) p.grs[k] = g[k];
      s.upd = 1;
    }
  }
  if (!s.reason && s.its >= p.maxit) s.reason = KSP_DIVERGED_ITS;
  s.phase = s.reason ? PH_DONE : PH_RESTART;
}

// x += sum_k y_k v_k
__global__ void __launch_bounds__(kBlock) k_g_update(Prob p) {
  const State& s = *p.st;
  if (!s.upd) return;
  const int top = s.it - 1;
  const size_t n = p.n;
  const double* __restrict__ V = p.V;
  const double* __restrict__ y = p.grs;
  ROWS(i) {
    double t = 0.0;
    for (int k = 0; k <= top; ++k) t = t + y[k] * V[k * n + i];
    p.x[i] = p.x[i] + t;
  }
}

// --- BiCGStab ------------------------------------------------------------------------

// r0 = rp = D^-1 b (k_resid), |r0|, test at 0, rho = r0 . rp
__global__ void __launch_bounds__(kRBlock) k_b_start(Prob p) {
  __shared__ double sh[kRBlock / 64];
  State& s = *p.st;
  if (s.phase != PH_RESTART) return;
  const double rr = reduce_part(p, p.nbs, 0, sh);
  if (threadIdx.x != 0) return;
  const double dp = sqrt(rr);
  s.res = dp;
  s.its = 0;
  log_res(p, 0, dp);
  s.reason = ksp_test(s, dp);
  if (s.reason) {
    s.phase = PH_DONE;
    return;
  }
  s.rho = rr;
  s.rhoold = 1.0;
  s.alpha = 1.0;
  s.omegaold = 1.0;
  s.beta = (s.rho / s.rhoold) * (s.alpha / s.omegaold);
  s.phase = PH_STEP;
}

// p = r - omega_old beta v + beta p (p = r on the first iteration, where
// PETSc's p and v are zero)
__global__ void __launch_bounds__(kBlock) k_b_dir(Prob p) {
  const State& s = *p.st;
  if (s.phase != PH_STEP) return;
  const bool first = s.its == 0;
  const double beta = s.beta, wb = -s.omegaold * s.beta;
  double* __restrict__ P = p.P;
  const double* __restrict__ R = p.R;
  const double* __restrict__ V = p.Vb;
  ROWS(i) P[i] = first ? R[i] : R[i] + wb * V[i] + beta * P[i];
}

// v = D^-1 A p, partial v . rp
__global__ void __launch_bounds__(kBlock) k_b_spmv_p(Prob p) {
  __shared__ double sh[kBlock / 64];
  const State& s = *p.st;
  if (s.phase != PH_STEP) return;
  const double* __restrict__ P = p.P;
  double acc = 0.0;
  spmv_rows(p, [&](int c) { return P[c]; }, [&](int row, double ap) {
    const double v = p.inv[row] * ap;
    p.Vb[row] = v;
    acc += v * p.RP[row];
  });
  store_part(p, 0, acc, sh);
}

// (k_b_alpha and k_b_finish load the State fields they read at entry, in
// one round trip with the phase; thread 0 is the only writer of the State
// while they run, and it writes after the loads.)
__global__ void __launch_bounds__(kRBlock) k_b_alpha(Prob p) {
  __shared__ double sh[kRBlock / 64];
  State& s = *p.st;
  const int phase = s.phase;
  const double rho = s.rho;
  OP2HIP_LOADED("s"(rho));
  if (phase != PH_STEP) return;
  const double d1 = reduce_part(p, p.nbs, 0, sh);
  if (threadIdx.x != 0) return;
  if (d1 == 0.0) {
    s.reason = KSP_DIVERGED_BREAKDOWN;
    s.phase = PH_DONE;
    return;
  }
  s.alpha = rho / d1;
}

// s = r - alpha v (formed where gathered, stored per row), t = D^-1 A s,
// partials s.t, t.t, s.s
__global__ void __launch_bounds__(kBlock) k_b_spmv_s(Prob p) {
  __shared__ double sh[kBlock / 64];
  const State& s = *p.st;
  if (s.phase != PH_STEP) return;
  const double na = -s.alpha;
  const double* __restrict__ R = p.R;
  const double* __restrict__ V = p.Vb;
  double st = 0.0, tt = 0.0, ss = 0.0;
  spmv_rows(p, [&](int c) { return na * V[c] + R[c]; }, [&](int row, double as) {
    const double si = na * V[row] + R[row];
    const double ti = p.inv[row] * as;
    p.S[row] = si;
    p.T[row] = ti;
    st += si * ti;
    tt += ti * ti;
    ss += si * si;
  });
  store_part(p, 0, st, sh);
  store_part(p, 1, tt, sh);
  store_part(p, 2, ss, sh);
}

__global__ void __launch_bounds__(kRBlock) k_b_omega(Prob p) {
  __shared__ double sh[3 * kRBlock / 64];
  State& s = *p.st;
  if (s.phase != PH_STEP) return;
  double d[3];
  reduce_parts<3>(p, p.nbs, 0, d, sh);
  if (threadIdx.x != 0) return;
  const double d1 = d[0], d2 = d[1], ss = d[2];
  if (d2 == 0.0) {
    // t = 0: with s = 0 too, x + alpha p is the solution
    if (ss != 0.0) {
      s.reason = KSP_DIVERGED_BREAKDOWN;
      s.phase = PH_DONE;
      return;
    }
    s.its += 1;
    s.res = 0.0;
    s.reason = KSP_CONVERGED_RTOL;
    log_res(p, s.its, 0.0);
    s.phase = PH_FINAL_X;
    return;
  }
  s.omega = d1 / d2;
}

// x = alpha p + omega s + x; r = s - omega t; partials r.r, r.rp
__global__ void __launch_bounds__(kBlock) k_b_update(Prob p) {
  __shared__ double sh[kBlock / 64];
  const State& s = *p.st;
  double* __restrict__ x = p.x;
  const double* __restrict__ P = p.P;
  if (s.phase == PH_FINAL_X) {
    const double a = s.alpha;
    ROWS(i) x[i] = x[i] + a * P[i];
    return;
  }
  if (s.phase != PH_STEP) return;
  const double a = s.alpha, om = s.omega, nom = -s.omega;
  const double* __restrict__ S = p.S;
  const double* __restrict__ T = p.T;
  const double* __restrict__ RP = p.RP;
  double* __restrict__ R = p.R;
  double rr = 0.0, rrp = 0.0;
  ROWS(i) {
    const double si = S[i];
    x[i] = a * P[i] + om * si + x[i];
    const double ri = nom * T[i] + si;
    R[i] = ri;
    rr += ri * ri;
    rrp += ri * RP[i];
  }
  store_part(p, 0, rr, sh);
  store_part(p, 1, rrp, sh);
}

__global__ void __launch_bounds__(kRBlock) k_b_finish(Prob p) {
  __shared__ double sh[2 * kRBlock / 64];
  State& s = *p.st;
  const int phase = s.phase, its = s.its + 1;
  const double rho = s.rho, omega = s.omega, alpha = s.alpha;
  const double ttol = s.ttol, atol = s.atol, dtol = s.dtol, rnorm0 = s.rnorm0;
  OP2HIP_LOADED("s"(its), "s"(rho), "s"(omega), "s"(alpha), "s"(ttol), "s"(atol), "s"(dtol),
                "s"(rnorm0));
  if (phase == PH_FINAL_X) {
    if (threadIdx.x == 0) s.phase = PH_DONE;
    return;
  }
  if (phase != PH_STEP) return;
  double d[2];
  reduce_parts<2>(p, p.nb, 0, d, sh);
  if (threadIdx.x != 0) return;
  const double rr = d[0], rrp = d[1];
  const double dp = sqrt(rr);
  s.rhoold = rho;
  s.omegaold = omega;
  s.its = its;
  s.res = dp;
  log_res(p, its, dp);
  // its >= 1: rnorm0 and ttol were set by the first test
  int reason = ksp_test_rn(dp, ttol, atol, dtol, rnorm0);
  if (!reason && rho == 0.0) reason = KSP_DIVERGED_BREAKDOWN;
  if (!reason && its >= p.maxit) reason = KSP_DIVERGED_ITS;
  s.reason = reason;
  if (reason) {
    s.phase = PH_DONE;
    return;
  }
  s.rho = rrp;
  s.beta = (rrp / rho) * (alpha / omega);
}

// --- host -----------------------------------------------------------------------------

inline size_t round_up(size_t v) { return (v + 31) & ~size_t(31); }

// one workgroup (and one partial) per tile of kTile rows
int tiles(int n) {
  const int nb = (n + kTile - 1) / kTile;
  return nb < 1 ? 1 : nb;
}

// SpMV workgroups: one per kBlock rows, at most kMaxSpmvBlocks
int spmv_blocks(int n) {
  const int nb = (n + kBlock - 1) / kBlock;
  return nb < 1 ? 1 : (nb > kMaxSpmvBlocks ? kMaxSpmvBlocks : nb);
}

// workspace layout in doubles: state, partials, D^-1, then the method's
// vectors and (GMRES) the small dense arrays
size_t layout(int method, int n, int m, Prob* p, double* w) {
  const int nb = tiles(n), nbs = spmv_blocks(n);
  const int ps = nb > nbs ? nb : nbs;
  const int nq = m + 1 > 3 ? m + 1 : 3;  // quantities reduced at once
  size_t off = 0;
  auto take = [&](size_t len) {
    double* q = w ? w + off : nullptr;
    off += round_up(len);
    return q;
  };
  State* st = reinterpret_cast<State*>(take(32));
  double* part = take((size_t)nq * ps);
  double* inv = take(n);
  if (p) {
    p->st = st;
    p->part = part;
    p->inv = inv;
    p->nb = nb;
    p->nbs = nbs;
    p->ps = ps;
  }
  if (method == METHOD_GMRES) {
    double* H = take((size_t)(m + 1) * (m > 0 ? m : 1));
    double* cc = take(m + 1);
    double* ss = take(m + 1);
    double* grs = take(m + 2);
    double* lhh = take(m + 1);
    double* V = take((size_t)(m + 1) * n);
    if (p) {
      p->H = H;
      p->cc = cc;
      p->ss = ss;
      p->grs = grs;
      p->lhh = lhh;
      p->V = V;
    }
  } else {
    double* v[6];
    for (int k = 0; k < 6; ++k) v[k] = take(n);
    if (p) {
      p->R = v[0];
      p->RP = v[1];
      p->Vb = v[2];
      p->T = v[3];
      p->S = v[4];
      p->P = v[5];
    }
  }
  return off;
}

int read_state(const Prob& p, State* h, hipStream_t s) {
  OP2HIP_CHECK(hipMemcpyAsync(h, p.st, sizeof(State), hipMemcpyDeviceToHost, s));
  OP2HIP_CHECK(hipStreamSynchronize(s));
  return OP2HIP_OK;
}

// One round: for GMRES the cycle-end work (solution update, restart
// residual and norm) then `steps` Arnoldi steps; for BiCGStab `steps`
// iterations.
void enqueue_round(int method, const Prob& p, int steps, hipStream_t s) {
  const dim3 grid(p.nb), sgrid(p.nbs), blk(kBlock), rblk(kRBlock), one(1);
  if (method == METHOD_GMRES) {
    const size_t dots_lds = sizeof(double) * (p.m + 1) * kWaves;
    const size_t rot_lds = sizeof(double) * 3 * (p.m + 1);
    const size_t solve_lds = sizeof(double) * (size_t)(p.m + 1) * (p.m + 1) <= kSolveLds
                                 ? sizeof(double) * (size_t)(p.m + 1) * (p.m + 1) : 0;
    hipLaunchKernelGGL(k_g_solve, one, blk, solve_lds, s, p);
    hipLaunchKernelGGL(k_g_update, grid, blk, 0, s, p);
    hipLaunchKernelGGL(k_resid, sgrid, blk, 0, s, p, p.V, static_cast<double*>(nullptr));
    hipLaunchKernelGGL(k_g_start, one, rblk, 0, s, p);
    for (int k = 0; k < steps; ++k) {
      hipLaunchKernelGGL(k_g_spmv, sgrid, blk, 0, s, p);
      hipLaunchKernelGGL(k_g_dots, grid, blk, dots_lds, s, p);
      hipLaunchKernelGGL(k_g_coef, one, rblk, 0, s, p);
      hipLaunchKernelGGL(k_g_orth, grid, blk, 0, s, p);
      hipLaunchKernelGGL(k_g_rot, one, rblk, rot_lds, s, p);
    }
  } else {
    for (int k = 0; k < steps; ++k) {
      hipLaunchKernelGGL(k_b_dir, grid, blk, 0, s, p);
      hipLaunchKernelGGL(k_b_spmv_p, sgrid, blk, 0, s, p);
      hipLaunchKernelGGL(k_b_alpha, one, rblk, 0, s, p);
      hipLaunchKernelGGL(k_b_spmv_s, sgrid, blk, 0, s, p);
      hipLaunchKernelGGL(k_b_omega, one, rblk, 0, s, p);
      hipLaunchKernelGGL(k_b_update, grid, blk, 0, s, p);
      hipLaunchKernelGGL(k_b_finish, one, rblk, 0, s, p);
    }
  }
}

// the capture stream and the instantiated round graphs of one problem
struct Rounds {
  hipStream_t cap = nullptr;
  hipGraphExec_t exec[kChunk + 1] = {};
  ~Rounds() {
    for (hipGraphExec_t g : exec)
      if (g) hipGraphExecDestroy(g);
    if (cap) hipStreamDestroy(cap);
  }
};

// Round graphs kept across solves.  A graph's kernels take the Prob by
// value, so a later solve whose Prob (matrix, vectors, work space, sizes)
// is byte-for-byte the same on the same device replays the same graphs;
// repeated solves of one system (time stepping) skip the capture.  The
// least recently used of kMaxCached entries is dropped.  Not thread-safe,
// like the rest of the library.
constexpr size_t kMaxCached = 8;

struct CacheEntry {
  int method, device;
  Prob p;
  std::unique_ptr<Rounds> r;
  unsigned long long used;
};

// never destroyed: graphs and streams must not be released during process
// teardown, after the HIP runtime may have gone
std::vector<CacheEntry>& round_cache() {
  static std::vector<CacheEntry>* c = new std::vector<CacheEntry>();
  return *c;
}

// Field-by-field equality of two argument blocks.  Prob has padding (nine
// ints before the GMRES pointers), whose bytes the language
// leaves unspecified, so a byte comparison of whole structs is not a sound
// cache key.
bool same_prob(const Prob& a, const Prob& b) {
  return a.rowptr == b.rowptr && a.colidx == b.colidx && a.vals == b.vals && a.b == b.b &&
         a.x == b.x && a.inv == b.inv && a.st == b.st && a.part == b.part && a.hist == b.hist &&
         a.n == b.n && a.lanes == b.lanes && a.ps == b.ps && a.m == b.m && a.maxit == b.maxit &&
         a.hist_len == b.hist_len && a.jacobi == b.jacobi && a.nb == b.nb && a.nbs == b.nbs &&
         a.V == b.V && a.H == b.H && a.cc == b.cc && a.ss == b.ss && a.grs == b.grs &&
         a.lhh == b.lhh && a.R == b.R && a.RP == b.RP && a.Vb == b.Vb && a.T == b.T &&
         a.S == b.S && a.P == b.P;
}
// Every member of Prob is compared by same_prob; a new member must be added
// there (the size check below fails when one is added without updating it).
static_assert(sizeof(Prob) == 9 * sizeof(void*) + 10 * sizeof(int) + 12 * sizeof(void*),
              "Prob changed: update same_prob");

Rounds* cached_rounds(int method, int device, const Prob& p) {
  static unsigned long long tick = 0;
  auto& c = round_cache();
  for (CacheEntry& e : c)
    if (e.method == method && e.device == device && same_prob(e.p, p)) {
      e.used = ++tick;
      return e.r.get();
    }
  if (c.size() >= kMaxCached) {
    size_t old = 0;
    for (size_t k = 1; k < c.size(); ++k)
      if (c[k].used < c[old].used) old = k;
    c.erase(c.begin() + old);
  }
  c.push_back(CacheEntry{method, device, p, std::unique_ptr<Rounds>(new Rounds()), ++tick});
  return c.back().r.get();
}

// a cache entry whose capture failed part-way is not kept
void drop_rounds(const Rounds* r) {
  auto& c = round_cache();
  for (size_t k = 0; k < c.size(); ++k)
    if (c[k].r.get() == r) {
      c.erase(c.begin() + k);
      return;
    }
}

}  // namespace

extern "C" {

long long op2hip_krylov_work(int method, int n, int restart) {
  if ((method != METHOD_GMRES && method != METHOD_BCGS) || n < 0 || restart < 0) return -1;
  return static_cast<long long>(layout(method, n, restart, nullptr, nullptr));
}

int op2hip_krylov_solve(int method, int n, const int* rowptr, const int* colidx,
                        const double* vals, const double* b, double* x, double* work,
                        long long work_len, int restart, int jacobi, double rtol, double atol,
                        double dtol, int maxit, double* reshist, int hist_len, int* its,
                        double* rnorm, int* reason, void* stream) {
  OP2HIP_REQUIRE(method == METHOD_GMRES || method == METHOD_BCGS, "unknown Krylov method %d",
                 method);
  OP2HIP_REQUIRE(n >= 0 && restart >= 0 && its && rnorm && reason, "bad Krylov arguments");
  OP2HIP_REQUIRE(method != METHOD_GMRES || (restart >= 1 && restart <= kMaxRestart),
                 "GMRES restart must be in [1, %d]", kMaxRestart);
  OP2HIP_REQUIRE(n == 0 || (rowptr && colidx && vals && b && x && work),
                 "null array passed to op2hip_krylov_solve");
  OP2HIP_REQUIRE(hist_len == 0 || reshist, "history length without a history array");
  if (n == 0) {
    // |r0| = 0: converged on entry, as PETSc reports it
    *its = 0;
    *rnorm = 0.0;
    *reason = KSP_CONVERGED_ATOL;
    return OP2HIP_OK;
  }
  const size_t need = layout(method, n, restart, nullptr, nullptr);
  OP2HIP_REQUIRE(work_len >= 0 && static_cast<size_t>(work_len) >= need,
                 "Krylov workspace of %lld doubles, %zu needed", work_len, need);
  hipStream_t s = static_cast<hipStream_t>(stream);
  Prob p{};
  layout(method, n, restart, &p, work);
  p.rowptr = rowptr;
  p.colidx = colidx;
  p.vals = vals;
  p.b = b;
  p.x = x;
  p.hist = hist_len > 0 ? reshist : nullptr;
  p.hist_len = hist_len;
  p.n = n;
  p.m = restart;
  p.maxit = maxit;
  p.jacobi = jacobi ? 1 : 0;
  // lanes per SpMV row from the mean row length: about two entries a lane
  int nnz = 0;
  OP2HIP_CHECK(hipMemcpyAsync(&nnz, rowptr + n, sizeof(int), hipMemcpyDeviceToHost, s));
  OP2HIP_CHECK(hipStreamSynchronize(s));
  const int mean = (nnz + n - 1) / n;
  int lanes = 2;
  while (lanes < kLanes && 2 * lanes < mean) lanes <<= 1;
  p.lanes = lanes;

  State h{};
  h.phase = PH_RESTART;
  h.rtol = rtol;
  h.atol = atol;
  h.dtol = dtol;
  h.scale = 1.0;
  OP2HIP_CHECK(hipMemcpyAsync(p.st, &h, sizeof(State), hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_init, dim3(p.nb), dim3(kBlock), 0, s, p);
  if (method == METHOD_BCGS) {
    hipLaunchKernelGGL(k_resid, dim3(p.nbs), dim3(kBlock), 0, s, p, p.R, p.RP);
    hipLaunchKernelGGL(k_b_start, dim3(1), dim3(kRBlock), 0, s, p);
  }
  OP2HIP_CHECK(hipGetLastError());
  // Rounds of iterations replayed as hipGraphs (one per round length); the
  // state is read after each round.  Kernels outside their phase return at
  // once, so a round may run past a cycle's end or the solve's.
  int device = 0;
  OP2HIP_CHECK(hipGetDevice(&device));
  Rounds& rounds = *cached_rounds(method, device, p);
  if (!rounds.cap) {
    const hipError_t e = hipStreamCreateWithFlags(&rounds.cap, hipStreamNonBlocking);
    if (e != hipSuccess) {
      rounds.cap = nullptr;
      drop_rounds(&rounds);
      OP2HIP_CHECK(e);
    }
  }
  const long long max_rounds = 4LL * (maxit > 0 ? maxit : 1) + 16;
  int st = OP2HIP_OK;
  for (long long round = 0;; ++round) {
    int steps = kBcgsChunk;
    if (method == METHOD_GMRES) {
      steps = h.phase == PH_STEP ? restart - h.it : restart;
      if (steps > kChunk) steps = kChunk;
    }
    hipGraphExec_t& g = rounds.exec[steps];
    if (!g) {
      hipGraph_t graph = nullptr;
      hipError_t e = hipStreamBeginCapture(rounds.cap, hipStreamCaptureModeThreadLocal);
      if (e == hipSuccess) {
        enqueue_round(method, p, steps, rounds.cap);
        e = hipStreamEndCapture(rounds.cap, &graph);
      }
      if (e == hipSuccess) {
        e = hipGraphInstantiate(&g, graph, nullptr, nullptr, 0);
        hipGraphDestroy(graph);
      }
      if (e != hipSuccess) {
        g = nullptr;
        drop_rounds(&rounds);
        OP2HIP_CHECK(e);
      }
    }
    OP2HIP_CHECK(hipGraphLaunch(g, s));
    st = read_state(p, &h, s);
    if (st != OP2HIP_OK || h.phase == PH_DONE) break;
    OP2HIP_REQUIRE(round < max_rounds, "Krylov solve did not finish in %lld rounds", round);
  }
  if (st != OP2HIP_OK) return st;
  *its = h.its;
  *rnorm = h.res;
  *reason = h.reason;
  return OP2HIP_OK;
}

}  // extern "C"
