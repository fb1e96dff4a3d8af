// Batched Jacobi-preconditioned CG: many independent FP64 CSR systems advanced
// together, each with its own scalars, iteration count and convergence test.
//
// Semantics per system are PETSc's KSPSolve_CG with PCJACOBI and the
// preconditioned norm (pyop2/petsc_base.py:172-220; defaults
// pyop2/base.py:1591-1601), as in the single solve:
//   r = b, z = D^-1 r, dp = |z|, test at it 0, beta = z.r
//   loop: p = z (+ beta/betaold p); w = A p; dpi = p.w; a = beta/dpi;
//         x += a p; r -= a w; z = D^-1 r; dp = |z|; test; beta = z.r
// Four launches per iteration: w = A p with p formed from z and the old p
// as it is gathered, partial p.w (k_spmv); a (k_alpha); p stored, x, r, z,
// partials z.z and z.r (k_axpy); dp, the test and beta (k_finish).  z and p
// are stored interleaved (zp[2 i] = z_i, zp[2 i + 1] = p_i): one 16-byte
// gather per matrix entry and one load and one store per row in k_axpy.
//   ttol = max(rtol*|z0|, atol); converged dp <= ttol; diverged
//   dp >= dtol*|z0|; DIVERGED_ITS after max_it.
//
// Work is cut into chunks of kChunk rows counted from each system's first
// row; every reduction is per-chunk partials summed in chunk order by one
// workgroup per system.  A system's arithmetic therefore does not depend on
// which other systems share the launch: a solve gives bit-identical results
// alone and in a batch.  Systems join (state NEW) and leave (DONE) between
// iteration launches, which is what the serving loop's continuous batching
// uses; kernels skip chunks of systems in other states.
#include <cmath>
#include <vector>

#include "common.h"
#include "reduce.h"

namespace {

constexpr int kBlock = 256;
constexpr int kChunk = 256;   // rows per chunk (kChunk / kBlock rows per thread)
constexpr int kLanes = 16;    // most lanes per row in SpMV
constexpr int kRBlock = 1024; // workgroup of the per-system reductions of chunk partials

enum { SYS_EMPTY = 0, SYS_NEW = 1, SYS_RUNNING = 2, SYS_DONE = 3 };

enum {
  KSP_CONVERGED_RTOL = 2,
  KSP_CONVERGED_ATOL = 3,
  KSP_DIVERGED_ITS = -3,
  KSP_DIVERGED_DTOL = -4,
  KSP_DIVERGED_INDEFINITE_MAT = -8,
  KSP_DIVERGED_NANORINF = -9,
};

}  // namespace

extern "C" {

// One system.  Vectors are indexed by global row i in [row0, row0 + n);
// colidx holds global column numbers into the same vectors.
struct op2hip_cg_system {
  const int* rowptr;
  const int* colidx;
  const double* vals;
  const double* b;
  double* x;
  double* inv_diag;
  double* r;
  double* zp;  // z and p interleaved: zp[2 i] = z_i, zp[2 i + 1] = p_i (16-byte aligned)
  double* w;
  int row0;
  int n;
  int chunk0;    // first entry of this system in the chunk list
  int nchunks;
  double rtol, atol, dtol;
  int maxit;
  int state;
  int no_precond;  // 0: Jacobi, 1: none (D^-1 taken as I)
  // solver scalars
  double beta, betaold, dpi, dpiold, dp, rnorm0, ttol, alpha;
  int its, reason;
};

}  // extern "C"

namespace {

using Sys = op2hip_cg_system;

using op2hip::block_sum;

// The systems' arrays are global memory, but their pointers come out of a
// descriptor in memory, so the compiler cannot tell their address space and
// emits flat loads; cast to global, they load with global instructions.
#define OP2HIP_GLOBAL __attribute__((address_space(1)))
// z and p of a row (a clang vector type: it loads and stores as one 16-byte
// access in any address space)
typedef double zp_pair __attribute__((ext_vector_type(2)));
template <typename T>
__device__ inline OP2HIP_GLOBAL T* gmem(T* p) {
  return (OP2HIP_GLOBAL T*)p;
}

// chunk list entry: (system, chunk index within the system)
struct ChunkRef {
  int sys, chunk;
};

__device__ inline void chunk_rows(int row0, int n, int c, int& lo, int& hi) {
  lo = row0 + c * kChunk;
  hi = lo + kChunk;
  if (hi > row0 + n) hi = row0 + n;
}

__device__ inline void chunk_rows(const Sys& s, int c, int& lo, int& hi) {
  chunk_rows(s.row0, s.n, c, lo, hi);
}

// The sums over a workgroup of NB threads of two values, each in
// block_sum's order (valid in thread 0), with one barrier: the last use of
// sh (2 * NB / 64 doubles) in the kernel.
template <int NB>
__device__ inline void block_sum_pair(double a, double b, double* sh, double& sa, double& sb) {
  a = op2hip::wave_sum(a);
  b = op2hip::wave_sum(b);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) {
    sh[w] = a;
    sh[NB / 64 + w] = b;
  }
  __syncthreads();
  double ta = 0.0, tb = 0.0;
  if (threadIdx.x == 0)
    for (int k = 0; k < NB / 64; ++k) {
      ta += sh[k];
      tb += sh[NB / 64 + k];
    }
  sa = ta;
  sb = tb;
}

// NEW systems: Jacobi diagonal, x = 0, r = b, z = D^-1 r, partials z.z, z.r
__global__ void __launch_bounds__(kBlock) k_init(Sys* sys, const ChunkRef* chunks,
                                                 double* part) {
  __shared__ double sh[2 * kBlock / 64];
  const ChunkRef cr = chunks[blockIdx.x];
  Sys& s = sys[cr.sys];
  if (s.state != SYS_NEW) return;
  int lo, hi;
  chunk_rows(s, cr.chunk, lo, hi);
  double zz = 0.0, zr = 0.0;
  // the system's arrays as local pointers: stores through them cannot
  // change the descriptor, so it is read once
  const OP2HIP_GLOBAL int* __restrict__ rowptr = gmem(s.rowptr);
  const OP2HIP_GLOBAL int* __restrict__ colidx = gmem(s.colidx);
  const OP2HIP_GLOBAL double* __restrict__ vals = gmem(s.vals);
  const OP2HIP_GLOBAL double* __restrict__ b = gmem(s.b);
  OP2HIP_GLOBAL double* __restrict__ x = gmem(s.x);
  OP2HIP_GLOBAL double* __restrict__ inv_diag = gmem(s.inv_diag);
  OP2HIP_GLOBAL double* __restrict__ r = gmem(s.r);
  OP2HIP_GLOBAL zp_pair* __restrict__ zp = gmem(reinterpret_cast<zp_pair*>(s.zp));
  for (int i = lo + (int)threadIdx.x; i < hi; i += kBlock) {
    double d = 0.0;
    for (int k = rowptr[i]; k < rowptr[i + 1]; ++k)
      if (colidx[k] == i) d = vals[k];
    double inv = (!s.no_precond && d != 0.0) ? 1.0 / d : 1.0;
    inv_diag[i] = inv;
    double ri = b[i];
    double zi = inv * ri;
    x[i] = 0.0;
    r[i] = ri;
    // p is first formed from z alone; its slot starts at zero
    zp[i] = zp_pair{zi, 0.0};
    zz += zi * zi;
    zr += zi * ri;
  }
  double t, u;
  block_sum_pair<kBlock>(zz, zr, sh, t, u);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = t;
    part[2 * blockIdx.x + 1] = u;
  }
}

// One thread's sums of a system's chunk partials (part[2 (c0 + c)] and
// part[2 (c0 + c) + 1], chunks c = tid, tid + kRBlock, ... in that order),
// with four chunks' loads in flight at once: a chunk past the end loads the
// last chunk (c < nch, so nch >= 1) and adds +0.0 instead, which leaves a
// sum that starts at +0.0 unchanged bit for bit (such a sum is never -0.0).
__device__ inline void thread_partials(const double* part, int c0, int nch, double& a,
                                       double& b) {
  const OP2HIP_GLOBAL zp_pair* __restrict__ pp =
      gmem(reinterpret_cast<const zp_pair*>(part)) + c0;
  double va = 0.0, vb = 0.0;
  for (int c = threadIdx.x; c < nch; c += 4 * kRBlock) {
    zp_pair q[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int ck = c + k * kRBlock;
      q[k] = pp[ck < nch ? ck : nch - 1];
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const bool in = c + k * kRBlock < nch;
      va += in ? q[k].x : 0.0;
      vb += in ? q[k].y : 0.0;
    }
  }
  a = va;
  b = vb;
}

// KSPConvergedDefault's test of dp: 0 to go on, else the reason
__device__ inline int test_dp(double dp, double ttol, double atol, double dtol, double rnorm0) {
  if (!(dp == dp) || isinf(dp)) return KSP_DIVERGED_NANORINF;
  if (dp <= ttol) return dp < atol ? KSP_CONVERGED_ATOL : KSP_CONVERGED_RTOL;
  if (dp >= dtol * rnorm0) return KSP_DIVERGED_DTOL;
  return 0;
}

// The per-system kernels (k_start, k_alpha, k_finish) load every field of
// their system they read at entry, in one round trip (OP2HIP_LOADED keeps
// the loads ahead of the state test): thread 0 of the system's workgroup is
// the only writer of the descriptor while they run, and it writes after
// the loads.

// one workgroup per system
__global__ void __launch_bounds__(kRBlock) k_start(Sys* sys, const double* part,
                                                  double* reshist, int hist_stride) {
  __shared__ double sh[2 * kRBlock / 64];
  Sys& s = sys[blockIdx.x];
  const int state = s.state, c0 = s.chunk0, nch = s.nchunks, maxit = s.maxit;
  const double rtol = s.rtol, atol = s.atol, dtol = s.dtol;
  OP2HIP_LOADED("s"(c0), "s"(nch), "s"(maxit), "s"(rtol), "s"(atol), "s"(dtol));
  if (state != SYS_NEW) return;
  double va, vb;
  thread_partials(part, c0, nch, va, vb);
  double zz, zr;
  block_sum_pair<kRBlock>(va, vb, sh, zz, zr);
  if (threadIdx.x == 0) {
    const double dp = sqrt(zz);
    const double ttol = fmax(rtol * dp, atol);
    s.dp = dp;
    s.rnorm0 = dp;
    s.ttol = ttol;
    s.its = 0;
    s.dpi = 0.0;
    if (reshist) reshist[(long long)blockIdx.x * hist_stride] = dp;
    int reason = test_dp(dp, ttol, atol, dtol, dp);
    s.beta = zr;
    if (reason == 0 && maxit <= 0) reason = KSP_DIVERGED_ITS;
    s.reason = reason;
    s.state = reason != 0 ? SYS_DONE : SYS_RUNNING;
  }
}


// w = A p for the chunk's rows, partial p.w
__global__ void __launch_bounds__(kBlock) k_spmv(Sys* sys, const ChunkRef* chunks,
                                                 double* part) {
  __shared__ double sh[kBlock / 64];
  const ChunkRef cr = chunks[blockIdx.x];
  const Sys& s = sys[cr.sys];
  // every descriptor field read here, loaded with the state in one round
  // trip (OP2HIP_LOADED keeps the loads ahead of the state test)
  const int state = s.state, row0 = s.row0, nrows = s.n, its = s.its;
  const double beta = s.beta, betaold = s.betaold;
  const int* rp = s.rowptr;
  const int* ci = s.colidx;
  const double* va = s.vals;
  const double* zpp = s.zp;
  double* wp = s.w;
  OP2HIP_LOADED("s"(cr.chunk), "s"(row0), "s"(nrows), "s"(its), "s"(beta), "s"(betaold), "s"(rp),
                "s"(ci), "s"(va), "s"(zpp), "s"(wp));
  if (state != SYS_RUNNING) return;
  int lo, hi;
  chunk_rows(row0, nrows, cr.chunk, lo, hi);
  const OP2HIP_GLOBAL int* __restrict__ rowptr = gmem(rp);
  const OP2HIP_GLOBAL int* __restrict__ colidx = gmem(ci);
  const OP2HIP_GLOBAL double* __restrict__ vals = gmem(va);
  const OP2HIP_GLOBAL zp_pair* __restrict__ zp = gmem(reinterpret_cast<const zp_pair*>(zpp));
  OP2HIP_GLOBAL double* __restrict__ w = gmem(wp);
  // lanes per row from the chunk's mean row length (about two entries per
  // lane: 4 lanes for 2-D P1 rows, 8 for 3-D, 16 for 3-D vector rows); a
  // property of the chunk alone, so a system sums alike in any batch
  const int nzc = rowptr[hi] - rowptr[lo];
  const int mean = (nzc + (hi - lo) - 1) / (hi - lo);
  int lanes = 2;
  while (lanes < kLanes && 2 * lanes < mean) lanes <<= 1;
  const int sub = threadIdx.x & (lanes - 1);
  const int stride = kBlock / lanes;
  // the new direction p' = z (first iteration) or z + (beta/betaold) p,
  // formed where it is read: the gathered columns here (one 16-byte load
  // of z and p per entry), each row again in k_axpy, which stores it
  const bool first = its == 0;
  const double bb = first ? 0.0 : beta / betaold;
  double acc = 0.0;
  // Two rows per lane group (i0 and i0 + stride) and two entries of each
  // per lane are loaded together, so four gathers are in flight per lane.
  // A lane still adds its entries k, k + lanes, k + 2 lanes, ... of a row
  // in that order and the rows enter acc in ascending order, as one row at
  // a time would: the sums do not change.  Entries past a row's end load
  // row 0's z and p (present in every batch) and are not added.
  for (int i0 = lo + (int)threadIdx.x / lanes; i0 < hi; i0 += 2 * stride) {
    const int i1 = i0 + stride;
    const bool has1 = i1 < hi;
    const int kb0 = rowptr[i0], ke0 = rowptr[i0 + 1];
    const int kb1 = has1 ? rowptr[i1] : 0, ke1 = has1 ? rowptr[i1 + 1] : 0;
    double v0 = 0.0, v1 = 0.0;
    for (int t = sub;; t += 2 * lanes) {
      const int ka = kb0 + t, kc = kb1 + t;
      const bool a0 = ka < ke0, a1 = ka + lanes < ke0, c0 = kc < ke1, c1 = kc + lanes < ke1;
      if (!a0 && !c0) break;
      const int ca0 = a0 ? colidx[ka] : 0, ca1 = a1 ? colidx[ka + lanes] : 0;
      const int cc0 = c0 ? colidx[kc] : 0, cc1 = c1 ? colidx[kc + lanes] : 0;
      const double va0 = a0 ? vals[ka] : 0.0, va1 = a1 ? vals[ka + lanes] : 0.0;
      const double vc0 = c0 ? vals[kc] : 0.0, vc1 = c1 ? vals[kc + lanes] : 0.0;
      const zp_pair qa0 = zp[ca0], qa1 = zp[ca1], qc0 = zp[cc0], qc1 = zp[cc1];
      const double pa0 = first ? qa0.x : qa0.x + bb * qa0.y;
      const double pa1 = first ? qa1.x : qa1.x + bb * qa1.y;
      const double pc0 = first ? qc0.x : qc0.x + bb * qc0.y;
      const double pc1 = first ? qc1.x : qc1.x + bb * qc1.y;
      if (a0) v0 += va0 * pa0;
      if (a1) v0 += va1 * pa1;
      if (c0) v1 += vc0 * pc0;
      if (c1) v1 += vc1 * pc1;
    }
    for (int off = lanes / 2; off > 0; off >>= 1) {
      v0 += __shfl_xor(v0, off, lanes);
      v1 += __shfl_xor(v1, off, lanes);
    }
    if (sub == 0) {
      w[i0] = v0;
      const zp_pair q0 = zp[i0];
      acc += (first ? q0.x : q0.x + bb * q0.y) * v0;
      if (has1) {
        w[i1] = v1;
        const zp_pair q1 = zp[i1];
        acc += (first ? q1.x : q1.x + bb * q1.y) * v1;
      }
    }
  }
  double t = block_sum<kBlock>(acc, sh);
  if (threadIdx.x == 0) part[2 * blockIdx.x] = t;
}

__global__ void __launch_bounds__(kRBlock) k_alpha(Sys* sys, const double* part) {
  __shared__ double sh[2 * kRBlock / 64];
  Sys& s = sys[blockIdx.x];
  const int state = s.state, c0 = s.chunk0, nch = s.nchunks, its = s.its + 1;
  const double beta = s.beta, dpiold = s.dpi;
  OP2HIP_LOADED("s"(c0), "s"(nch), "s"(its), "s"(beta), "s"(dpiold));
  if (state != SYS_RUNNING) return;
  double va, vb;
  thread_partials(part, c0, nch, va, vb);
  double pw, unused;
  block_sum_pair<kRBlock>(va, 0.0, sh, pw, unused);
  if (threadIdx.x == 0) {
    s.its = its;
    if (beta == 0.0) {
      s.reason = KSP_CONVERGED_ATOL;
      s.state = SYS_DONE;
      return;
    }
    s.dpiold = dpiold;
    s.dpi = pw;
    if (pw == 0.0 || (its > 1 && pw * dpiold <= 0.0)) {
      s.reason = KSP_DIVERGED_INDEFINITE_MAT;
      s.state = SYS_DONE;
      return;
    }
    s.alpha = beta / pw;
  }
}

__global__ void __launch_bounds__(kBlock) k_axpy(Sys* sys, const ChunkRef* chunks, double* part) {
  __shared__ double sh[2 * kBlock / 64];
  const ChunkRef cr = chunks[blockIdx.x];
  const Sys& s = sys[cr.sys];
  // every descriptor field read here, loaded with the state in one round trip
  const int state = s.state, row0 = s.row0, nrows = s.n, its = s.its;
  const double a = s.alpha, beta = s.beta, betaold = s.betaold;
  double* xp = s.x;
  double* rp = s.r;
  const double* wp = s.w;
  const double* dp = s.inv_diag;
  double* zpp = s.zp;
  OP2HIP_LOADED("s"(cr.chunk), "s"(row0), "s"(nrows), "s"(its), "s"(a), "s"(beta), "s"(betaold),
                "s"(xp), "s"(rp), "s"(wp), "s"(dp), "s"(zpp));
  if (state != SYS_RUNNING) return;
  int lo, hi;
  chunk_rows(row0, nrows, cr.chunk, lo, hi);
  double zz = 0.0, zr = 0.0;
  OP2HIP_GLOBAL double* __restrict__ x = gmem(xp);
  OP2HIP_GLOBAL double* __restrict__ r = gmem(rp);
  const OP2HIP_GLOBAL double* __restrict__ w = gmem(wp);
  const OP2HIP_GLOBAL double* __restrict__ inv_diag = gmem(dp);
  OP2HIP_GLOBAL zp_pair* __restrict__ zp = gmem(reinterpret_cast<zp_pair*>(zpp));
  // k_alpha has counted this iteration: its == 1 on the first
  const bool first = its == 1;
  const double bb = first ? 0.0 : beta / betaold;
  for (int i = lo + (int)threadIdx.x; i < hi; i += kBlock) {
    const zp_pair q = zp[i];
    const double pi = first ? q.x : q.x + bb * q.y;
    x[i] = x[i] + a * pi;
    double ri = r[i] - a * w[i];
    r[i] = ri;
    double zi = inv_diag[i] * ri;
    zp[i] = zp_pair{zi, pi};
    zz += zi * zi;
    zr += zi * ri;
  }
  double t, u;
  block_sum_pair<kBlock>(zz, zr, sh, t, u);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = t;
    part[2 * blockIdx.x + 1] = u;
  }
}

__global__ void __launch_bounds__(kRBlock) k_finish(Sys* sys, const double* part, double* reshist,
                                                   int hist_stride) {
  __shared__ double sh[2 * kRBlock / 64];
  Sys& s = sys[blockIdx.x];
  const int state = s.state, c0 = s.chunk0, nch = s.nchunks, its = s.its, maxit = s.maxit;
  const double ttol = s.ttol, atol = s.atol, dtol = s.dtol, rnorm0 = s.rnorm0, beta = s.beta;
  OP2HIP_LOADED("s"(c0), "s"(nch), "s"(its), "s"(maxit), "s"(ttol), "s"(atol), "s"(dtol),
                "s"(rnorm0), "s"(beta));
  if (state != SYS_RUNNING) return;
  double va, vb;
  thread_partials(part, c0, nch, va, vb);
  double zz, zr;
  block_sum_pair<kRBlock>(va, vb, sh, zz, zr);
  if (threadIdx.x == 0) {
    const double dp = sqrt(zz);
    s.dp = dp;
    if (reshist && its < hist_stride) reshist[(long long)blockIdx.x * hist_stride + its] = dp;
    int reason = test_dp(dp, ttol, atol, dtol, rnorm0);
    if (reason == 0 && its >= maxit) reason = KSP_DIVERGED_ITS;
    if (reason != 0) {
      s.reason = reason;
      s.state = SYS_DONE;
    } else {
      s.betaold = beta;
      s.beta = zr;
    }
  }
}

struct Batch {
  int max_sys = 0;
  long long max_chunks = 0;
  int nsys = 0;
  long long nchunks = 0;
  Sys* d_sys = nullptr;
  ChunkRef* d_chunks = nullptr;
  double* d_part = nullptr;
  double* d_hist = nullptr;
  int hist_stride = 0;
  std::vector<Sys> h_sys;
  // a captured step (graph of `g_iters` iterations for the current chunk
  // list), rebuilt when the chunk list or the iteration count changes
  hipStream_t cap = nullptr;
  hipGraphExec_t gexec = nullptr;
  int g_iters = -1;
  long long g_nchunks = -1;
  int g_nsys = -1;
};

void enqueue_step(Batch* b, int iters, hipStream_t s) {
  dim3 gc(static_cast<unsigned>(b->nchunks)), gs(b->nsys), blk(kBlock), rblk(kRBlock);
  hipLaunchKernelGGL(k_init, gc, blk, 0, s, b->d_sys, b->d_chunks, b->d_part);
  hipLaunchKernelGGL(k_start, gs, rblk, 0, s, b->d_sys, b->d_part, b->d_hist, b->hist_stride);
  for (int it = 0; it < iters; ++it) {
    hipLaunchKernelGGL(k_spmv, gc, blk, 0, s, b->d_sys, b->d_chunks, b->d_part);
    hipLaunchKernelGGL(k_alpha, gs, rblk, 0, s, b->d_sys, b->d_part);
    hipLaunchKernelGGL(k_axpy, gc, blk, 0, s, b->d_sys, b->d_chunks, b->d_part);
    hipLaunchKernelGGL(k_finish, gs, rblk, 0, s, b->d_sys, b->d_part, b->d_hist, b->hist_stride);
  }
}

void drop_graph(Batch* b) {
  if (b->gexec) hipGraphExecDestroy(b->gexec);
  b->gexec = nullptr;
  b->g_iters = -1;
}

}  // namespace

extern "C" {

// A batch of up to max_sys systems and max_chunks chunks of 256 rows;
// hist_len > 0 keeps a residual history of hist_len entries per system.
int op2hip_cgb_create(int max_sys, long long max_chunks, int hist_len, void** out) {
  OP2HIP_REQUIRE(max_sys > 0 && max_chunks > 0 && out, "bad batch sizes");
  Batch* b = new Batch();
  b->max_sys = max_sys;
  b->max_chunks = max_chunks;
  b->hist_stride = hist_len;
  hipError_t e = hipMalloc(&b->d_sys, sizeof(Sys) * max_sys);
  if (e == hipSuccess) e = hipMalloc(&b->d_chunks, sizeof(ChunkRef) * max_chunks);
  if (e == hipSuccess) e = hipMalloc(&b->d_part, sizeof(double) * 2 * max_chunks);
  if (e == hipSuccess && hist_len > 0)
    e = hipMalloc(&b->d_hist, sizeof(double) * (size_t)hist_len * max_sys);
  if (e != hipSuccess) {
    if (b->d_sys) hipFree(b->d_sys);
    if (b->d_chunks) hipFree(b->d_chunks);
    if (b->d_part) hipFree(b->d_part);
    delete b;
    op2hip::set_error("batch allocation failed: %s", hipGetErrorString(e));
    return OP2HIP_ERR_HIP;
  }
  *out = b;
  return OP2HIP_OK;
}

int op2hip_cgb_destroy(void* h) {
  Batch* b = static_cast<Batch*>(h);
  if (!b) return OP2HIP_OK;
  drop_graph(b);
  if (b->cap) hipStreamDestroy(b->cap);
  hipFree(b->d_sys);
  hipFree(b->d_chunks);
  hipFree(b->d_part);
  if (b->d_hist) hipFree(b->d_hist);
  delete b;
  return OP2HIP_OK;
}

// Load the batch's systems (host array of nsys descriptors; state NEW for
// systems to start, RUNNING/DONE to keep, EMPTY for unused slots) and
// rebuild the chunk list (NEW and RUNNING systems only).  Solver scalars of RUNNING/DONE systems are taken
// from the descriptors, so a caller that keeps systems across calls passes
// back what op2hip_cgb_read returned.
int op2hip_cgb_load(void* h, int nsys, const op2hip_cg_system* systems, void* stream) {
  Batch* b = static_cast<Batch*>(h);
  OP2HIP_REQUIRE(b && nsys >= 0 && nsys <= b->max_sys, "bad batch load (%d systems)", nsys);
  hipStream_t s = static_cast<hipStream_t>(stream);
  b->h_sys.assign(systems, systems + nsys);
  std::vector<ChunkRef> chunks;
  for (int k = 0; k < nsys; ++k) {
    Sys& sy = b->h_sys[k];
    OP2HIP_REQUIRE(sy.state == SYS_EMPTY || (reinterpret_cast<uintptr_t>(sy.zp) & 15) == 0,
                   "system %d: zp must be 16-byte aligned", k);
    sy.chunk0 = static_cast<int>(chunks.size());
    // finished and empty slots get no chunks: launches cover live rows only
    sy.nchunks = (sy.state == SYS_EMPTY || sy.state == SYS_DONE) ? 0 : (sy.n + kChunk - 1) / kChunk;
    for (int c = 0; c < sy.nchunks; ++c) chunks.push_back({k, c});
  }
  OP2HIP_REQUIRE(static_cast<long long>(chunks.size()) <= b->max_chunks,
                 "batch needs %zu chunks, capacity %lld", chunks.size(), b->max_chunks);
  b->nsys = nsys;
  b->nchunks = static_cast<long long>(chunks.size());
  if (nsys) OP2HIP_CHECK(hipMemcpyAsync(b->d_sys, b->h_sys.data(), sizeof(Sys) * nsys,
                                        hipMemcpyHostToDevice, s));
  if (!chunks.empty())
    OP2HIP_CHECK(hipMemcpyAsync(b->d_chunks, chunks.data(), sizeof(ChunkRef) * chunks.size(),
                                hipMemcpyHostToDevice, s));
  OP2HIP_CHECK(hipStreamSynchronize(s));
  return OP2HIP_OK;
}

// Start NEW systems and run `iters` iterations of every RUNNING one.
// Enqueues only (no host synchronisation).
int op2hip_cgb_step(void* h, int iters, void* stream) {
  Batch* b = static_cast<Batch*>(h);
  OP2HIP_REQUIRE(b, "null batch");
  if (b->nsys == 0 || b->nchunks == 0) return OP2HIP_OK;
  enqueue_step(b, iters, static_cast<hipStream_t>(stream));
  OP2HIP_CHECK(hipGetLastError());
  return OP2HIP_OK;
}

// The same step replayed from a hipGraph: the 2 + 4 * iters launches are
// captured once per chunk list and iteration count (on a stream of the
// batch's own) and then launched on `stream` as one graph.  Results are
// those of op2hip_cgb_step.
int op2hip_cgb_step_graph(void* h, int iters, void* stream) {
  Batch* b = static_cast<Batch*>(h);
  OP2HIP_REQUIRE(b && iters >= 0, "bad batch step");
  if (b->nsys == 0 || b->nchunks == 0) return OP2HIP_OK;
  if (!b->gexec || b->g_iters != iters || b->g_nchunks != b->nchunks || b->g_nsys != b->nsys) {
    drop_graph(b);
    if (!b->cap) OP2HIP_CHECK(hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking));
    hipGraph_t g = nullptr;
    OP2HIP_CHECK(hipStreamBeginCapture(b->cap, hipStreamCaptureModeThreadLocal));
    enqueue_step(b, iters, b->cap);
    hipError_t e = hipStreamEndCapture(b->cap, &g);
    OP2HIP_CHECK(e);
    e = hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    OP2HIP_CHECK(e);
    b->g_iters = iters;
    b->g_nchunks = b->nchunks;
    b->g_nsys = b->nsys;
  }
  OP2HIP_CHECK(hipGraphLaunch(b->gexec, static_cast<hipStream_t>(stream)));
  return OP2HIP_OK;
}

// Copy the systems' descriptors (with state, its, reason, dp) back to the
// host; waits for the stream.
int op2hip_cgb_read(void* h, op2hip_cg_system* out, void* stream) {
  Batch* b = static_cast<Batch*>(h);
  OP2HIP_REQUIRE(b && (out || b->nsys == 0), "bad batch read");
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (b->nsys)
    OP2HIP_CHECK(hipMemcpyAsync(out, b->d_sys, sizeof(Sys) * b->nsys, hipMemcpyDeviceToHost, s));
  OP2HIP_CHECK(hipStreamSynchronize(s));
  return OP2HIP_OK;
}

// Residual history of system k (hist_len entries) into host `out`.
int op2hip_cgb_history(void* h, int k, double* out, void* stream) {
  Batch* b = static_cast<Batch*>(h);
  OP2HIP_REQUIRE(b && b->d_hist && k >= 0 && k < b->nsys, "no history for system %d", k);
  hipStream_t s = static_cast<hipStream_t>(stream);
  OP2HIP_CHECK(hipMemcpyAsync(out, b->d_hist + (size_t)k * b->hist_stride,
                              sizeof(double) * b->hist_stride, hipMemcpyDeviceToHost, s));
  OP2HIP_CHECK(hipStreamSynchronize(s));
  return OP2HIP_OK;
}

int op2hip_cgb_sizeof_system(void) { return static_cast<int>(sizeof(Sys)); }

}  // extern "C"

namespace {

// ---- the single solve without device allocation ----------------------------
//
// op2hip_pcg_noalloc keeps what op2hip_pcg allocates per call (the system
// descriptor, the chunk list and the chunk partials) in this file-scope
// array instead: one copy per GPU, resident until the process ends.  Host
// code must not pass the array's name to a kernel (in host code it names a
// host-side stand-in); it uses the device address that
// hipGetSymbolAddress returns for the current device, kept per device.
constexpr int kNoAllocMaxRows = 1 << 24;
constexpr long long kNoAllocChunks = kNoAllocMaxRows / kChunk + 1;
constexpr size_t kScratchSys = 256;  // one Sys, rounded up to 256 bytes
constexpr size_t kScratchChunks = ((sizeof(ChunkRef) * kNoAllocChunks + 255) / 256) * 256;
constexpr size_t kScratchBytes = kScratchSys + kScratchChunks + sizeof(double) * 2 * kNoAllocChunks;
static_assert(sizeof(Sys) <= kScratchSys, "system descriptor outgrew its scratch slot");

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
unsigned char* g_scratch_ptr[kMaxDevices] = {};

// Device address of this GPU's scratch (looked up once per device).
int scratch_for_current_device(unsigned char** out) {
  int dev = 0;
  OP2HIP_CHECK(hipGetDevice(&dev));
  OP2HIP_REQUIRE(dev >= 0 && dev < kMaxDevices, "device %d beyond the %d supported", dev,
                 kMaxDevices);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    OP2HIP_CHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    OP2HIP_REQUIRE(ptr != nullptr, "no device address for the solver scratch");
    g_scratch_ptr[dev] = static_cast<unsigned char*>(ptr);
  }
  *out = g_scratch_ptr[dev];
  return OP2HIP_OK;
}

// The descriptor and chunk list of one system, written by a kernel (no host
// copy): chunk c of system 0 is (0, c).
__global__ void k_noalloc_setup(Sys sy, Sys* __restrict__ sys, ChunkRef* __restrict__ chunks,
                                int nchunks) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c == 0) sys[0] = sy;
  if (c < nchunks) chunks[c] = ChunkRef{0, c};
}

}  // namespace

extern "C" {

int op2hip_pcg_noalloc_max_rows(void) { return kNoAllocMaxRows; }

// op2hip_pcg without device allocation (no counterpart in the reference:
// for callers that cannot allocate).  Same arguments, results bit for bit
// those of op2hip_pcg; n at most op2hip_pcg_noalloc_max_rows().  The
// solver's own state lives in a per-GPU static array: calls that overlap on
// one GPU (other threads, other streams) race.  Waits for the stream.
int op2hip_pcg_noalloc(int n, const int* rowptr, const int* colidx, const double* vals,
                       const double* b, double* x, double* work, double rtol, double atol,
                       double dtol, int maxit, int* its, double* rnorm, int* reason,
                       void* stream) {
  OP2HIP_REQUIRE(n >= 0 && rowptr && colidx && vals && b && x && work && its && rnorm && reason,
                 "null argument to op2hip_pcg_noalloc");
  OP2HIP_REQUIRE((reinterpret_cast<uintptr_t>(work) & 15) == 0,
                 "op2hip_pcg_noalloc: work must be 16-byte aligned");
  OP2HIP_REQUIRE(n <= kNoAllocMaxRows, "op2hip_pcg_noalloc takes at most %d rows, not %d",
                 kNoAllocMaxRows, n);
  unsigned char* base = nullptr;
  int st = scratch_for_current_device(&base);
  if (st != OP2HIP_OK) return st;
  hipStream_t s = static_cast<hipStream_t>(stream);
  Batch bt;
  bt.max_sys = 1;
  bt.max_chunks = kNoAllocChunks;
  bt.d_sys = reinterpret_cast<Sys*>(base);
  bt.d_chunks = reinterpret_cast<ChunkRef*>(base + kScratchSys);
  bt.d_part = reinterpret_cast<double*>(base + kScratchSys + kScratchChunks);
  bt.d_hist = nullptr;
  bt.hist_stride = 0;
  Sys sy{};
  sy.rowptr = rowptr;
  sy.colidx = colidx;
  sy.vals = vals;
  sy.b = b;
  sy.x = x;
  sy.inv_diag = work;
  sy.r = work + n;
  sy.zp = work + 2 * (size_t)n;
  sy.w = work + 4 * (size_t)n;
  sy.row0 = 0;
  sy.n = n;
  sy.chunk0 = 0;
  sy.nchunks = (n + kChunk - 1) / kChunk;
  sy.rtol = rtol;
  sy.atol = atol;
  sy.dtol = dtol;
  sy.maxit = maxit;
  sy.state = SYS_NEW;
  bt.nsys = 1;
  bt.nchunks = sy.nchunks;
  hipLaunchKernelGGL(k_noalloc_setup, dim3((sy.nchunks + 255) / 256 + 1), dim3(256), 0, s, sy,
                     bt.d_sys, bt.d_chunks, sy.nchunks);
  OP2HIP_CHECK(hipGetLastError());
  for (int done = 0;;) {
    const int k = done ? 16 : 4;
    if (bt.nchunks > 0) {
      enqueue_step(&bt, k, s);
      OP2HIP_CHECK(hipGetLastError());
    }
    OP2HIP_CHECK(hipMemcpyAsync(&sy, bt.d_sys, sizeof(Sys), hipMemcpyDeviceToHost, s));
    OP2HIP_CHECK(hipStreamSynchronize(s));
    done += k;
    if (sy.state == SYS_DONE || bt.nchunks == 0 || done > maxit + 16) break;
  }
  *its = sy.its;
  *rnorm = sy.dp;
  *reason = sy.reason;
  return OP2HIP_OK;
}

int op2hip_cgb_chunk_rows(void) { return kChunk; }


// One Jacobi-PCG solve (the single-system form of the batch): A (n x n CSR
// on the device), b, x (device, x overwritten), work (device, 5*n doubles,
// 16-byte aligned: D^-1, r, then z and p interleaved over 2*n, then w).  PETSc KSPCG + PCJACOBI semantics as above; its,
// rnorm (the last preconditioned residual norm) and reason (KSP converged
// reason) are written to host memory.  Waits for the stream.
int op2hip_pcg(int n, const int* rowptr, const int* colidx, const double* vals, const double* b,
               double* x, double* work, double rtol, double atol, double dtol, int maxit,
               int* its, double* rnorm, int* reason, void* stream) {
  OP2HIP_REQUIRE(n >= 0 && rowptr && colidx && vals && b && x && work && its && rnorm && reason,
                 "null argument to op2hip_pcg");
  OP2HIP_REQUIRE((reinterpret_cast<uintptr_t>(work) & 15) == 0,
                 "op2hip_pcg: work must be 16-byte aligned");
  void* h = nullptr;
  int st = op2hip_cgb_create(1, (n + kChunk - 1) / kChunk + 1, 0, &h);
  if (st != OP2HIP_OK) return st;
  Sys sy{};
  sy.rowptr = rowptr;
  sy.colidx = colidx;
  sy.vals = vals;
  sy.b = b;
  sy.x = x;
  sy.inv_diag = work;
  sy.r = work + n;
  sy.zp = work + 2 * (size_t)n;
  sy.w = work + 4 * (size_t)n;
  sy.row0 = 0;
  sy.n = n;
  sy.rtol = rtol;
  sy.atol = atol;
  sy.dtol = dtol;
  sy.maxit = maxit;
  sy.state = SYS_NEW;
  st = op2hip_cgb_load(h, 1, &sy, stream);
  for (int done = 0; st == OP2HIP_OK;) {
    const int k = done ? 16 : 4;
    st = op2hip_cgb_step(h, k, stream);
    if (st == OP2HIP_OK) st = op2hip_cgb_read(h, &sy, stream);
    done += k;
    if (st != OP2HIP_OK || sy.state == SYS_DONE || done > maxit + 16) break;
  }
  op2hip_cgb_destroy(h);
  if (st != OP2HIP_OK) return st;
  *its = sy.its;
  *rnorm = sy.dp;
  *reason = sy.reason;
  return OP2HIP_OK;
}

}  // extern "C"
