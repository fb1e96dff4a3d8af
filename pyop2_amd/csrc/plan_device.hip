// Device construction of owner-computes plans.
//
// Replaces, for loops whose conflicting writes are increments into one
// target set, the host plan of pyop2/plan.pyx:68-456 (partitioning, staging
// and colouring) by the owner-computes plan, built here on the device by
// sort and scan.  Targets are owned by blocks in contiguous ranges of k
// (b owns [b*k, min(ntargets, (b+1)*k)); contiguous positions are
// mesh-local under the device layout).  A block executes every element
// naming one of its targets and sums each owned target's increments in the
// order of the elements' ranks (their user numbering, the order of the
// reference's sequential loop), so no two blocks write one target and
// neither colours nor atomics are needed at run time.  The block owning an
// element's first race column is its primary block.  The arrays are the
// same, entry for entry, as the host build's (op2hip_owner_plan_build):
//
//   tp[ntargets+1], contrib   per target, its increments (slot * nac + ac)
//                             in ascending (rank, column) order
//   blk_t0/blk_t1, rng_lo/rng_n, ex_off/exec, tg_off/tgt, tg_cb
//                             per active block (one naming any element)
//
// Steps: count incidences per target, scan (tp), scatter 64-bit keys
// (rank, element * nac + column), sort each target's keys (a wavefront per
// target of up to 256 keys, a workgroup in global memory beyond); then one
// workgroup per block sorts its incidences' elements (LDS, or global memory
// for blocks of more than kCap incidences), keeps the distinct ones,
// classifies them primary or not, numbers their slots, writes the codes
// and lists; scans over the blocks compact the active ones.  The one host
// synchronisation reads the scalars the caller sizes its arrays by.
#include <vector>

#include "common.h"
#include "scan.h"

namespace {

using op2hip_scan::exclusive_scan;
using op2hip_scan::scan_tiles_for;

constexpr int kMaxCols = 16;  // race columns of a loop
constexpr int kThreads = 256;
constexpr int kCap = 4096;      // incidences a block sorts in LDS
constexpr int kSegWave = 256;   // target keys a wavefront sorts in registers
constexpr int kLongGrid = 512;  // workgroups of the long-segment sort
typedef unsigned long long u64;
constexpr u64 kKeyPad = ~0ull;

struct Cols {
  int n;  // nac
  const int* v[kMaxCols];
  int dim[kMaxCols];
  int col[kMaxCols];
};

__device__ inline int target_of(const Cols& C, int e, int ac) {
  return C.v[ac][static_cast<long long>(e) * C.dim[ac] + C.col[ac]];
}

__global__ void zero_ints(int* __restrict__ p, long long n) {
  long long k = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k < n) p[k] = 0;
}

__global__ void count_targets(Cols C, int n, int ntargets, int* __restrict__ cnt) {
  long long k = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= static_cast<long long>(n) * C.n) return;
  const int e = static_cast<int>(k / C.n), ac = static_cast<int>(k - static_cast<long long>(e) * C.n);
  const int t = target_of(C, e, ac);
  if (static_cast<unsigned>(t) < static_cast<unsigned>(ntargets)) atomicAdd(cnt + t, 1);
}

__global__ void scatter_keys(Cols C, int n, int ntargets, const int* __restrict__ rank,
                             int* __restrict__ cursor, u64* __restrict__ key) {
  long long k = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= static_cast<long long>(n) * C.n) return;
  const int e = static_cast<int>(k / C.n), ac = static_cast<int>(k - static_cast<long long>(e) * C.n);
  const int t = target_of(C, e, ac);
  if (static_cast<unsigned>(t) >= static_cast<unsigned>(ntargets)) return;
  const unsigned r = rank ? static_cast<unsigned>(rank[e]) : static_cast<unsigned>(e);
  const int pos = atomicAdd(cursor + t, 1);
  key[pos] = (static_cast<u64>(r) << 32) | static_cast<unsigned>(k);
}

// Ascending bitonic sort of 64*R keys v[r] (element r*64 + lane).
template <int R, typename T>
__device__ inline void wave_sort(T (&v)[R], int lane) {
  constexpr int N = 64 * R;
#pragma unroll
  for (int kk = 2; kk <= N; kk <<= 1) {
#pragma unroll
    for (int j = kk >> 1; j > 0; j >>= 1) {
      if (j >= 64) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
          int pr = r ^ (j >> 6);
          if (pr > r) {
            bool asc = (((r << 6) | lane) & kk) == 0;
            T a = v[r], b = v[pr];
            T lo = a < b ? a : b, hi = a < b ? b : a;
            v[r] = asc ? lo : hi;
            v[pr] = asc ? hi : lo;
          }
        }
      } else {
#pragma unroll
        for (int r = 0; r < R; ++r) {
          T w = __shfl_xor(v[r], j, 64);
          bool lower = (lane & j) == 0;
          bool asc = (((r << 6) | lane) & kk) == 0;
          T lo = v[r] < w ? v[r] : w, hi = v[r] < w ? w : v[r];
          v[r] = (lower == asc) ? lo : hi;
        }
      }
    }
  }
}

template <int R>
__device__ inline void sort_segment_wave(u64* __restrict__ key, int q0, int len, int lane) {
  u64 v[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    int i = (r << 6) | lane;
    v[r] = i < len ? key[q0 + i] : kKeyPad;
  }
  wave_sort<R>(v, lane);
#pragma unroll
  for (int r = 0; r < R; ++r) {
    int i = (r << 6) | lane;
    if (i < len) key[q0 + i] = v[r];
  }
}

// One wavefront per target: sort its keys; longer segments are listed.
__global__ void __launch_bounds__(256) sort_targets_wave(const int* __restrict__ tp, int ntargets,
                                                         u64* __restrict__ key,
                                                         int* __restrict__ longlist,
                                                         int* __restrict__ nlong) {
  const int lane = threadIdx.x & 63;
  long long t = (static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
  if (t >= ntargets) return;
  const int q0 = tp[t], len = tp[t + 1] - q0;
  if (len <= 1) return;
  if (len > kSegWave) {
    if (lane == 0) longlist[atomicAdd(nlong, 1)] = static_cast<int>(t);
    return;
  }
  if (len <= 64)
    sort_segment_wave<1>(key, q0, len, lane);
  else if (len <= 128)
    sort_segment_wave<2>(key, q0, len, lane);
  else
    sort_segment_wave<4>(key, q0, len, lane);
}

// Ascending sort of a[0, n) by one workgroup: the bitonic network in its
// all-ascending ("flip") form over the next power of two, comparisons with
// an index at or beyond n skipped (those entries stand for +infinity, which
// that form never moves).  a may be LDS or global memory.
template <typename T>
__device__ void block_sort(T* a, int n) {
  int P = 1;
  while (P < n) P <<= 1;
  const int tid = threadIdx.x;
  for (int size = 2; size <= P; size <<= 1) {
    const int half = size >> 1;
    for (int i = tid; i < (P >> 1); i += blockDim.x) {
      const int blk = i / half, off = i - blk * half;
      const int lo = blk * size + off, hi = blk * size + size - 1 - off;
      if (hi < n) {
        T x = a[lo], y = a[hi];
        if (y < x) {
          a[lo] = y;
          a[hi] = x;
        }
      }
    }
    __syncthreads();
    for (int j = half >> 1; j > 0; j >>= 1) {
      for (int i = tid; i < (P >> 1); i += blockDim.x) {
        const int blk = i / j, off = i - blk * j;
        const int lo = blk * 2 * j + off, hi = lo + j;
        if (hi < n) {
          T x = a[lo], y = a[hi];
          if (y < x) {
            a[lo] = y;
            a[hi] = x;
          }
        }
      }
      __syncthreads();
    }
  }
}

__global__ void __launch_bounds__(kThreads) sort_targets_long(const int* __restrict__ tp,
                                                              u64* key,
                                                              const int* __restrict__ longlist,
                                                              const int* __restrict__ nlong) {
  const int total = *nlong;
  for (int li = blockIdx.x; li < total; li += gridDim.x) {
    const int t = longlist[li];
    const int q0 = tp[t];
    block_sort<u64>(key + q0, tp[t + 1] - q0);
    __syncthreads();
  }
}

// Workgroup-wide exclusive prefix of one int per thread (kThreads threads);
// *all receives the total.
__device__ inline int block_prefix(int v, int* s_part, int* all) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int x = v;
  for (int off = 1; off < 64; off <<= 1) {
    int y = __shfl_up(x, off, 64);
    if (lane >= off) x += y;
  }
  if (lane == 63) s_part[w] = x;
  __syncthreads();
  int before = 0, tot = 0;
  for (int k = 0; k < kThreads / 64; ++k) {
    if (k < w) before += s_part[k];
    tot += s_part[k];
  }
  __syncthreads();
  *all = tot;
  return before + x - v;
}

// Per-block results of the block pass (raw block numbering).
struct BlockStats {
  int* nels;   // distinct elements
  int* rn;     // primary run length (0: no run)
  int* rlo;    // primary run start
  int* nexec;  // listed elements
  int* ntgt;   // targets receiving increments
  int* ncode;  // incidences of the block's targets
  int* active;
};

// One workgroup per block of k targets.
__global__ void __launch_bounds__(kThreads) block_pass(
    Cols C, int ntargets, int k, int nblk, const int* __restrict__ tp,
    const u64* __restrict__ key, int* gbuf, int* gslot, int* __restrict__ exec_tmp,
    int* __restrict__ tgt_tmp, int* __restrict__ contrib32, BlockStats st) {
  __shared__ int s_buf[kCap];
  __shared__ int s_slot[kCap];
  __shared__ int s_part[kThreads / 64];
  __shared__ int s_nprim, s_pmin, s_pmax;
  const int tid = threadIdx.x;
  const int nac = C.n;
  for (int b = blockIdx.x; b < nblk; b += gridDim.x) {
    const int t0 = b * k, t1 = t0 + k < ntargets ? t0 + k : ntargets;
    const int q0 = tp[t0], M = tp[t1] - q0;
    if (M == 0) {
      if (tid == 0) {
        st.nels[b] = st.rn[b] = st.rlo[b] = st.nexec[b] = st.ntgt[b] = st.ncode[b] = 0;
        st.active[b] = 0;
      }
      continue;
    }
    int* buf = M <= kCap ? s_buf : gbuf + q0;
    int* sl = M <= kCap ? s_slot : gslot + q0;
    for (int i = tid; i < M; i += kThreads)
      buf[i] = static_cast<int>(static_cast<unsigned>(key[q0 + i] & 0xffffffffull) / nac);
    if (tid == 0) {
      s_nprim = 0;
      s_pmin = 0x7fffffff;
      s_pmax = -1;
    }
    __syncthreads();
    block_sort<int>(buf, M);
    // first occurrences, flagged before the in-place compaction below
    for (int i = tid; i < M; i += kThreads) sl[i] = (i == 0 || buf[i - 1] != buf[i]) ? 1 : 0;
    __syncthreads();
    int nels = 0;
    for (int c = 0; c < M; c += kThreads) {
      const int i = c + tid;
      const int v = i < M ? buf[i] : 0;
      const int f = i < M ? sl[i] : 0;
      int tot;
      const int pos = block_prefix(f, s_part, &tot);
      if (f) buf[nels + pos] = v;  // nels + pos <= i: every later chunk is still unread
      nels += tot;
      __syncthreads();
    }
    // primaries: the element's first race column is owned here
    for (int u = tid; u < nels; u += kThreads) {
      const int e = buf[u];
      const int prim = target_of(C, e, 0) / k == b;
      sl[u] = prim;
      if (prim) {
        atomicAdd(&s_nprim, 1);
        atomicMin(&s_pmin, e);
        atomicMax(&s_pmax, e);
      }
    }
    __syncthreads();
    const int nprim = s_nprim;
    const bool run = nprim > 0 && s_pmax - s_pmin + 1 == nprim;
    const int rn = run ? nprim : 0, rlo = run ? s_pmin : 0;
    // slots: the primary run first (element - rlo), then the others in
    // ascending order; without a run, the ascending position
    int nonp_before = 0;
    for (int c = 0; c < nels; c += kThreads) {
      const int u = c + tid;
      const int prim = u < nels ? sl[u] : 1;
      const int e = u < nels ? buf[u] : 0;
      int tot;
      const int pos = block_prefix(prim ? 0 : 1, s_part, &tot);
      if (u < nels) {
        int slot;
        if (run)
          slot = prim ? e - rlo : rn + nonp_before + pos;
        else
          slot = u;
        sl[u] = slot;
        if (run) {
          if (!prim)
            exec_tmp[q0 + nonp_before + pos] =
                static_cast<int>(static_cast<unsigned>(e) | 0x80000000u);
        } else {
          exec_tmp[q0 + u] = prim ? e : static_cast<int>(static_cast<unsigned>(e) | 0x80000000u);
        }
      }
      nonp_before += tot;
      __syncthreads();
    }
    // contribution codes, in the targets' key order
    for (int i = tid; i < M; i += kThreads) {
      const unsigned low = static_cast<unsigned>(key[q0 + i] & 0xffffffffull);
      const int e = static_cast<int>(low / nac), ac = static_cast<int>(low - e * nac);
      int lo = 0, hi = nels - 1;
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (buf[mid] < e)
          lo = mid + 1;
        else
          hi = mid;
      }
      contrib32[q0 + i] = sl[lo] * nac + ac;
    }
    // targets receiving increments
    int ntgt = 0;
    for (int c = t0; c < t1; c += kThreads) {
      const int t = c + tid;
      const int has = t < t1 && tp[t + 1] > tp[t];
      int tot;
      const int pos = block_prefix(has, s_part, &tot);
      if (has) tgt_tmp[t0 + ntgt + pos] = t;
      ntgt += tot;
    }
    if (tid == 0) {
      st.nels[b] = nels;
      st.rn[b] = rn;
      st.rlo[b] = rlo;
      st.nexec[b] = run ? nels - nprim : nels;
      st.ntgt[b] = ntgt;
      st.ncode[b] = M;
      st.active[b] = 1;
    }
    __syncthreads();
  }
}

// Scalars the caller sizes its arrays by: 0 nactive, 1 exec length, 2 tgt
// length, 3 max_exec (largest nels), 4 max codes per block, 5 sum of nels.
__global__ void __launch_bounds__(1024) plan_scalars(int nblk, BlockStats st,
                                                     const int* __restrict__ aidx,
                                                     const int* __restrict__ exoff,
                                                     const int* __restrict__ tgoff,
                                                     long long* __restrict__ out) {
  __shared__ int s_me, s_mc;
  __shared__ unsigned long long s_sum;
  if (threadIdx.x == 0) {
    s_me = 0;
    s_mc = 0;
    s_sum = 0;
  }
  __syncthreads();
  int me = 0, mc = 0;
  unsigned long long sum = 0;
  for (int b = threadIdx.x; b < nblk; b += blockDim.x) {
    me = st.nels[b] > me ? st.nels[b] : me;
    if (st.active[b]) mc = st.ncode[b] > mc ? st.ncode[b] : mc;
    sum += static_cast<unsigned long long>(st.nels[b]);
  }
  atomicMax(&s_me, me);
  atomicMax(&s_mc, mc);
  atomicAdd(&s_sum, sum);
  __syncthreads();
  if (threadIdx.x == 0) {
    out[0] = aidx[nblk];
    out[1] = exoff[nblk];
    out[2] = tgoff[nblk];
    out[3] = s_me;
    out[4] = s_mc;
    out[5] = static_cast<long long>(s_sum);
  }
}

struct Outputs {
  int *blk_t0, *blk_t1, *rng_lo, *rng_n, *ex_off, *exec, *tg_off, *tgt, *tg_cb;
  void* contrib;
  int c16;
};

// Per-block arrays of the active blocks (one thread per raw block).
__global__ void compact_blocks(int ntargets, int k, int nblk, BlockStats st,
                               const int* __restrict__ aidx, const int* __restrict__ exoff,
                               const int* __restrict__ tgoff, Outputs o) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b > nblk) return;
  const int nact = aidx[nblk];
  if (b == nblk) {
    o.ex_off[nact] = exoff[nblk];
    o.tg_off[nact] = tgoff[nblk];
    return;
  }
  if (!st.active[b]) return;
  const int a = aidx[b];
  const int t0 = b * k;
  o.blk_t0[a] = t0;
  o.blk_t1[a] = t0 + k < ntargets ? t0 + k : ntargets;
  o.rng_lo[a] = st.rlo[b];
  o.rng_n[a] = st.rn[b];
  o.ex_off[a] = exoff[b];
  o.tg_off[a] = tgoff[b];
}

// Lists of the active blocks moved to their compact offsets (one workgroup
// per raw block), with each listed target's first contribution.
__global__ void __launch_bounds__(kThreads) compact_lists(int k, int nblk, BlockStats st,
                                                          const int* __restrict__ tp,
                                                          const int* __restrict__ exoff,
                                                          const int* __restrict__ tgoff,
                                                          const int* __restrict__ exec_tmp,
                                                          const int* __restrict__ tgt_tmp,
                                                          Outputs o) {
  for (int b = blockIdx.x; b < nblk; b += gridDim.x) {
    if (!st.active[b]) continue;
    const int t0 = b * k, q0 = tp[t0];
    const int ne = st.nexec[b], nt = st.ntgt[b];
    for (int i = threadIdx.x; i < ne; i += kThreads) o.exec[exoff[b] + i] = exec_tmp[q0 + i];
    for (int i = threadIdx.x; i < nt; i += kThreads) {
      const int t = tgt_tmp[t0 + i];
      o.tgt[tgoff[b] + i] = t;
      o.tg_cb[tgoff[b] + i] = tp[t];
    }
  }
}

// The padding entries of tgt and tg_cb (a kernel may load entry ng of a
// block unconditionally) and the contribution codes in their final width.
__global__ void finish_lists(int ntargets, const int* __restrict__ tp,
                             const int* __restrict__ tgoff, int nblk, Outputs o) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    const int nt = tgoff[nblk];
    o.tgt[nt] = 0;
    o.tg_cb[nt] = tp[ntargets];
    o.tg_cb[nt + 1] = tp[ntargets];
  }
}

__global__ void pack_contrib(const int* __restrict__ c32, long long n, Outputs o) {
  long long q = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (q >= n) return;
  if (o.c16)
    static_cast<unsigned short*>(o.contrib)[q] = static_cast<unsigned short>(c32[q]);
  else
    static_cast<int*>(o.contrib)[q] = c32[q];
}

size_t align256(size_t x) { return (x + 255) & ~static_cast<size_t>(255); }

struct Work {
  size_t cnt, key, longlist, nlong, gbuf, gslot, exec_tmp, tgt_tmp, contrib32, st, aidx, exoff,
      tgoff, sums, total, scal, bytes;
};

Work work_layout(long long ntargets, long long ninc, long long nblk) {
  Work w;
  size_t o = 0;
  auto take = [&](size_t b) {
    size_t at = o;
    o += align256(b > 0 ? b : 1);
    return at;
  };
  const size_t I = sizeof(int);
  w.cnt = take(I * (ntargets + 1));
  w.key = take(sizeof(u64) * ninc);
  w.longlist = take(I * (ntargets + 1));
  w.nlong = take(I);
  w.gbuf = take(I * ninc);
  w.gslot = take(I * ninc);
  w.exec_tmp = take(I * ninc);
  w.tgt_tmp = take(I * (ntargets + 1));
  w.contrib32 = take(I * ninc);
  w.st = take(I * 7 * (nblk + 1));
  w.aidx = take(I * (nblk + 1));
  w.exoff = take(I * (nblk + 1));
  w.tgoff = take(I * (nblk + 1));
  long long ns = scan_tiles_for(ntargets + 1);
  long long nb2 = scan_tiles_for(nblk + 1);
  w.sums = take(sizeof(long long) * (ns > nb2 ? ns : nb2));
  w.total = take(sizeof(long long) * 2);
  w.scal = take(sizeof(long long) * 8);
  w.bytes = o;
  return w;
}

struct DevicePlan {
  int nelems = 0, ntargets = 0, k = 0, nac = 0, nblk = 0;
  long long ninc = 0;
  char* work = nullptr;
  Work w{};
  long long scal[6] = {0, 0, 0, 0, 0, 0};
};

template <typename T>
T* at(char* base, size_t off) {
  return reinterpret_cast<T*>(base + off);
}

BlockStats stats_of(char* base, const Work& w, int nblk) {
  int* s = at<int>(base, w.st);
  const int n = nblk + 1;
  return BlockStats{s, s + n, s + 2 * n, s + 3 * n, s + 4 * n, s + 5 * n, s + 6 * n};
}

}  // namespace

extern "C" {

long long op2hip_owner_plan_workspace(int nelems, int ntargets, int nac, int k) {
  if (nelems < 0 || ntargets < 0 || nac <= 0 || k <= 0) return -1;
  const long long nblk = (static_cast<long long>(ntargets) + k - 1) / k;
  return static_cast<long long>(
      work_layout(ntargets, static_cast<long long>(nelems) * nac, nblk).bytes);
}

// Phase 1 of the device build: everything but the compaction.  maps: host
// array of device int32 map pointers (rows in the loop's element order, at
// least nelems of them); idxs[m] the column used (-1: all dims[m]); rank
// (device, nelems ints, may be NULL for the identity) orders each target's
// increments; tp (device, ntargets + 1) receives the contribution offsets.
// scalars (host, 6): nactive, exec length, tgt length, max_exec, largest
// codes per block, sum of elements executed per block.  Waits for the
// stream once.
int op2hip_owner_plan_dbuild(int nelems, int ntargets, int nmaps, const int* const* maps,
                             const int* dims, const int* idxs, int k, const int* rank,
                             void* work, long long work_bytes, int* tp, void** handle,
                             long long* scalars, void* stream) {
  OP2HIP_REQUIRE(handle && scalars && tp && work && nelems >= 0 && ntargets >= 0 && k > 0 &&
                     nmaps > 0,
                 "bad owner plan arguments");
  Cols C{};
  for (int m = 0; m < nmaps; ++m) {
    OP2HIP_REQUIRE(maps[m] && dims[m] > 0 && idxs[m] < dims[m], "bad race map %d", m);
    const int c0 = idxs[m] >= 0 ? idxs[m] : 0, c1 = idxs[m] >= 0 ? idxs[m] + 1 : dims[m];
    for (int c = c0; c < c1; ++c) {
      OP2HIP_REQUIRE(C.n < kMaxCols, "more than %d race columns", kMaxCols);
      C.v[C.n] = maps[m];
      C.dim[C.n] = dims[m];
      C.col[C.n] = c;
      ++C.n;
    }
  }
  const long long ninc = static_cast<long long>(nelems) * C.n;
  OP2HIP_REQUIRE(ninc <= 0x7fffffffLL, "owner plan of %lld incidences, beyond int32", ninc);
  const int nblk = static_cast<int>((static_cast<long long>(ntargets) + k - 1) / k);
  Work w = work_layout(ntargets, ninc, nblk);
  OP2HIP_REQUIRE(work_bytes >= static_cast<long long>(w.bytes),
                 "owner plan workspace of %lld bytes, %zu needed", work_bytes, w.bytes);
  hipStream_t s = static_cast<hipStream_t>(stream);
  char* base = static_cast<char*>(work);
  int* cnt = at<int>(base, w.cnt);
  u64* key = at<u64>(base, w.key);
  int* longlist = at<int>(base, w.longlist);
  int* nlong = at<int>(base, w.nlong);
  long long* sums = at<long long>(base, w.sums);
  long long* total = at<long long>(base, w.total);
  BlockStats st = stats_of(base, w, nblk);
  int* aidx = at<int>(base, w.aidx);
  int* exoff = at<int>(base, w.exoff);
  int* tgoff = at<int>(base, w.tgoff);
  long long* scal = at<long long>(base, w.scal);

  hipLaunchKernelGGL(zero_ints, dim3((ntargets + 1 + 255) / 256), dim3(256), 0, s, cnt,
                     static_cast<long long>(ntargets) + 1);
  hipLaunchKernelGGL(zero_ints, dim3(1), dim3(64), 0, s, nlong, 1LL);
  if (ninc > 0)
    hipLaunchKernelGGL(count_targets, dim3((ninc + 255) / 256), dim3(256), 0, s, C, nelems,
                       ntargets, cnt);
  exclusive_scan(cnt, ntargets, tp, cnt, sums, total, s);
  if (ninc > 0)
    hipLaunchKernelGGL(scatter_keys, dim3((ninc + 255) / 256), dim3(256), 0, s, C, nelems,
                       ntargets, rank, cnt, key);
  if (ntargets > 0) {
    long long threads = static_cast<long long>(ntargets) * 64;
    hipLaunchKernelGGL(sort_targets_wave, dim3((threads + 255) / 256), dim3(256), 0, s, tp,
                       ntargets, key, longlist, nlong);
    hipLaunchKernelGGL(sort_targets_long, dim3(kLongGrid), dim3(kThreads), 0, s, tp, key,
                       longlist, nlong);
  }
  if (nblk > 0) {
    const int grid = nblk < 65536 ? nblk : 65536;
    hipLaunchKernelGGL(block_pass, dim3(grid), dim3(kThreads), 0, s, C, ntargets, k, nblk, tp,
                       key, at<int>(base, w.gbuf), at<int>(base, w.gslot),
                       at<int>(base, w.exec_tmp), at<int>(base, w.tgt_tmp),
                       at<int>(base, w.contrib32), st);
  }
  exclusive_scan(st.active, nblk, aidx, nullptr, sums, total, s);
  exclusive_scan(st.nexec, nblk, exoff, nullptr, sums, total, s);
  exclusive_scan(st.ntgt, nblk, tgoff, nullptr, sums, total, s);
  hipLaunchKernelGGL(plan_scalars, dim3(1), dim3(1024), 0, s, nblk, st, aidx, exoff, tgoff, scal);
  long long host[6];
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(host, scal, sizeof(host), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) {
    op2hip::set_error("owner plan build failed: %s", hipGetErrorString(e));
    return OP2HIP_ERR_HIP;
  }
  DevicePlan* p = new DevicePlan();
  p->nelems = nelems;
  p->ntargets = ntargets;
  p->k = k;
  p->nac = C.n;
  p->nblk = nblk;
  p->ninc = ninc;
  p->work = base;
  p->w = w;
  for (int i = 0; i < 6; ++i) p->scal[i] = scalars[i] = host[i];
  *handle = p;
  return OP2HIP_OK;
}

// Phase 2: write the compacted arrays (device, caller-allocated): blk_t0,
// blk_t1, rng_lo, rng_n (nactive), ex_off, tg_off (nactive + 1), exec
// (exec length), tgt (tgt length + 1), tg_cb (tgt length + 2) and contrib
// (ninc entries, uint16 when c16 else int32); tp was written by phase 1.
// Enqueued on the stream; releases the handle.
int op2hip_owner_plan_dfill(void* handle, const int* tp, int* blk_t0, int* blk_t1, int* rng_lo,
                            int* rng_n, int* ex_off, int* exec, int* tg_off, int* tgt,
                            int* tg_cb, void* contrib, int c16, void* stream) {
  DevicePlan* p = static_cast<DevicePlan*>(handle);
  OP2HIP_REQUIRE(p && tp && ex_off && tg_off && tgt && tg_cb, "null owner plan output");
  OP2HIP_REQUIRE(p->ninc == 0 || contrib, "null contribution array");
  hipStream_t s = static_cast<hipStream_t>(stream);
  char* base = p->work;
  const Work& w = p->w;
  BlockStats st = stats_of(base, w, p->nblk);
  Outputs o{blk_t0, blk_t1, rng_lo, rng_n, ex_off, exec, tg_off, tgt, tg_cb, contrib, c16};
  hipLaunchKernelGGL(compact_blocks, dim3((p->nblk + 1 + 255) / 256), dim3(256), 0, s,
                     p->ntargets, p->k, p->nblk, st, at<int>(base, w.aidx),
                     at<int>(base, w.exoff), at<int>(base, w.tgoff), o);
  if (p->nblk > 0) {
    const int grid = p->nblk < 65536 ? p->nblk : 65536;
    hipLaunchKernelGGL(compact_lists, dim3(grid), dim3(kThreads), 0, s, p->k, p->nblk, st, tp,
                       at<int>(base, w.exoff), at<int>(base, w.tgoff),
                       at<int>(base, w.exec_tmp), at<int>(base, w.tgt_tmp), o);
  }
  hipLaunchKernelGGL(finish_lists, dim3(1), dim3(64), 0, s, p->ntargets, tp,
                     at<int>(base, w.tgoff), p->nblk, o);
  if (p->ninc > 0)
    hipLaunchKernelGGL(pack_contrib, dim3((p->ninc + 255) / 256), dim3(256), 0, s,
                       at<int>(base, w.contrib32), p->ninc, o);
  hipError_t e = hipGetLastError();
  delete p;
  if (e != hipSuccess) {
    op2hip::set_error("owner plan fill failed: %s", hipGetErrorString(e));
    return OP2HIP_ERR_HIP;
  }
  return OP2HIP_OK;
}

int op2hip_owner_plan_ddiscard(void* handle) {
  delete static_cast<DevicePlan*>(handle);
  return OP2HIP_OK;
}

}  // extern "C"

// ---- owner Mat assembly program on the device -----------------------------

namespace {

struct MatPlan {
  const int *blk_t0, *blk_t1, *ex_off, *exec, *rng_lo, *rng_n, *tp;
  const void* contrib;
  int c16, nac, nactive;
};

__device__ inline int contrib_at(const MatPlan& P, int c) {
  return P.c16 ? static_cast<int>(static_cast<const unsigned short*>(P.contrib)[c])
               : static_cast<const int*>(P.contrib)[c];
}

// Pair of (row node n, column node m): the position of m among the node
// columns of scalar row n*rd (every cd-th entry), as a global pair index;
// -1 when the pattern lacks it.
__device__ inline long long pair_of(const int* __restrict__ rowptr,
                                    const int* __restrict__ colidx, int rd, int cd, int n, int m) {
  const int s0 = rowptr[static_cast<long long>(n) * rd];
  const int s1 = rowptr[static_cast<long long>(n) * rd + 1];
  int lo = 0, hi = (s1 - s0) / cd - 1;
  while (lo <= hi) {
    const int mid = (lo + hi) >> 1;
    const int v = colidx[s0 + mid * cd] / cd;
    if (v == m) return static_cast<long long>(s0) / (static_cast<long long>(rd) * cd) + mid;
    if (v < m)
      lo = mid + 1;
    else
      hi = mid - 1;
  }
  return -1;
}

// One thread per row node: its contributions in plan order (ascending user
// element, then row position) and each one's column positions j, so the
// pairs of node n, touched by this thread alone, receive their entries in
// sequential order without atomics.  COUNT: entries per pair; else the
// codes, through per-pair cursors (the scanned offsets).
template <bool COUNT>
__global__ void mat_program_rows(MatPlan P, const int* __restrict__ cmap, int carity, int e1,
                                 int nnodes, const int* __restrict__ rowptr,
                                 const int* __restrict__ colidx, int rd, int cd,
                                 int* __restrict__ cnt, int wide, void* __restrict__ prog,
                                 int* __restrict__ err) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nnodes) return;
  const int c0 = P.tp[n], c1 = P.tp[n + 1];
  if (c0 == c1) return;
  // the active block owning n: the last whose first target is <= n
  int lo = 0, hi = P.nactive - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (P.blk_t0[mid] <= n)
      lo = mid;
    else
      hi = mid - 1;
  }
  const int b = lo;
  const int rn = P.rng_n[b], rlo = P.rng_lo[b], ex0 = P.ex_off[b];
  const int cbase = P.tp[P.blk_t0[b]];
  for (int c = c0; c < c1; ++c) {
    const int slot = contrib_at(P, c) / P.nac;
    const int e = slot < rn ? rlo + slot
                            : static_cast<int>(static_cast<unsigned>(P.exec[ex0 + slot - rn]) &
                                               0x7fffffffu);
    const long long code0 = static_cast<long long>(c - cbase) * e1;
    for (int j = 0; j < e1; ++j) {
      const long long p =
          pair_of(rowptr, colidx, rd, cd, n, cmap[static_cast<long long>(e) * carity + j]);
      if (p < 0) {
        atomicOr(err, 1);
        continue;
      }
      if (COUNT) {
        ++cnt[p];
      } else {
        const int at = cnt[p]++;
        const long long code = code0 + j;
        if (wide)
          static_cast<int*>(prog)[at] = static_cast<int>(code);
        else if (code > 0xffff)
          atomicOr(err, 2);
        else
          static_cast<unsigned short*>(prog)[at] = static_cast<unsigned short>(code);
      }
    }
  }
}

}  // namespace

extern "C" {

long long op2hip_mat_program_workspace(long long npairs) {
  if (npairs < 0) return -1;
  return static_cast<long long>(align256(sizeof(int) * (npairs + 1)) +
                                align256(sizeof(long long) * scan_tiles_for(npairs + 1)) +
                                align256(sizeof(long long) * 2) + align256(sizeof(int)));
}

// The scatter program of op2hip_mat_program built on the device from the
// device owner plan (contrib uint16 when c16, else int32) and the device
// CSR: poff (device, npairs + 1) and prog (device, tp[nnodes] * e1 codes,
// uint16 unless wide).  work: op2hip_mat_program_workspace(npairs) bytes.
// Waits for the stream once, to report an element naming an entry outside
// the pattern or a code that does not fit 16 bits.
int op2hip_mat_program_device(int nactive, const int* blk_t0, const int* blk_t1,
                              const int* ex_off, const int* exec, const int* rng_lo,
                              const int* rng_n, const int* tp, const void* contrib, int c16,
                              int nac, const int* cmap, int carity, int e1, int nnodes,
                              const int* rowptr, const int* colidx, int rd, int cd,
                              long long npairs, int wide, void* work, long long work_bytes,
                              int* poff, void* prog, void* stream) {
  OP2HIP_REQUIRE(nactive >= 0 && nac > 0 && carity > 0 && e1 > 0 && e1 <= carity && rd > 0 &&
                     cd > 0 && npairs >= 0 && poff && prog && work,
                 "bad Mat program arguments");
  OP2HIP_REQUIRE(npairs <= 0x7fffffffLL, "Mat pattern of %lld node pairs", npairs);
  OP2HIP_REQUIRE(work_bytes >= op2hip_mat_program_workspace(npairs), "Mat program workspace");
  hipStream_t s = static_cast<hipStream_t>(stream);
  char* base = static_cast<char*>(work);
  int* cnt = reinterpret_cast<int*>(base);
  size_t off = align256(sizeof(int) * (npairs + 1));
  long long* sums = reinterpret_cast<long long*>(base + off);
  off += align256(sizeof(long long) * scan_tiles_for(npairs + 1));
  long long* total = reinterpret_cast<long long*>(base + off);
  off += align256(sizeof(long long) * 2);
  int* err = reinterpret_cast<int*>(base + off);
  MatPlan P{blk_t0, blk_t1, ex_off, exec, rng_lo, rng_n, tp, contrib, c16, nac, nactive};
  hipLaunchKernelGGL(zero_ints, dim3((npairs + 1 + 255) / 256), dim3(256), 0, s, cnt,
                     npairs + 1);
  hipLaunchKernelGGL(zero_ints, dim3(1), dim3(64), 0, s, err, 1LL);
  const dim3 grid((nnodes + 255) / 256 > 0 ? (nnodes + 255) / 256 : 1);
  if (nactive > 0)
    hipLaunchKernelGGL(mat_program_rows<true>, grid, dim3(256), 0, s, P, cmap, carity, e1,
                       nnodes, rowptr, colidx, rd, cd, cnt, wide, prog, err);
  exclusive_scan(cnt, static_cast<int>(npairs), poff, cnt, sums, total, s);
  if (nactive > 0)
    hipLaunchKernelGGL(mat_program_rows<false>, grid, dim3(256), 0, s, P, cmap, carity, e1,
                       nnodes, rowptr, colidx, rd, cd, cnt, wide, prog, err);
  int herr = 0;
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(&herr, err, sizeof(int), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) {
    op2hip::set_error("Mat program build failed: %s", hipGetErrorString(e));
    return OP2HIP_ERR_HIP;
  }
  OP2HIP_REQUIRE(!(herr & 1), "an element names a matrix entry outside the sparsity pattern");
  OP2HIP_REQUIRE(!(herr & 2), "Mat program codes do not fit 16 bits");
  return OP2HIP_OK;
}

}  // extern "C"

// ---- row-block Mat plans ----------------------------------------------------
//
// The plan of the owner Mat "rows" template.  Block b owns the k row nodes
// [b*k, min(nnodes, (b+1)*k)) of the device layout and runs every element
// naming one of them as a row, in ascending user order (rank), in chunks of
// a workgroup; after each chunk one thread per owned node adds the chunk's
// entries of its row into the block's CSR slice, in its records' order.
// Arrays:
//
//   tp[nnodes+1], code   per row node, its records in ascending (rank, row
//                        position i0): code = slot * e0 + i0, slot the
//                        element's position in its block's exec list
//   pos[ninc * W]        per record, W = ceil(e1 / 4) words of e1 bytes:
//                        byte j is the position of column node cmap[e][j]
//                        among the node columns of row node n
//   xoff[nblk+1], exec   per block, its elements (device numbering) in
//                        ascending rank
//
// Steps: count records per node, scan (tp), scatter 64-bit keys (rank,
// element * e0 + i0) and sort each node's keys (the owner-plan kernels
// above); count the distinct blocks of each element, scan (xoff), scatter
// (rank, element) keys per block and sort each block's keys (LDS); then one
// thread per node writes its records' codes and column positions.  The one
// host synchronisation reads four scalars and the error flags.

namespace {

constexpr int kCapX = 4096;  // exec keys a workgroup sorts in LDS

// Distinct blocks owning a row node of element e: count them (COUNT) or
// scatter the key (rank << 32 | e) to each block's cursor.
template <bool COUNT>
__global__ void rows_exec_entries(const int* __restrict__ rmap, int rarity, int e0, int n,
                                  int nnodes, int k, const int* __restrict__ rank,
                                  int* __restrict__ cnt, u64* __restrict__ xkey) {
  const long long e = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const int* row = rmap + e * rarity;
  for (int i = 0; i < e0; ++i) {
    const int t = row[i];
    if (static_cast<unsigned>(t) >= static_cast<unsigned>(nnodes)) continue;
    const int b = t / k;
    bool dup = false;
    for (int j = 0; j < i; ++j) {
      const int u = row[j];
      dup |= static_cast<unsigned>(u) < static_cast<unsigned>(nnodes) && u / k == b;
    }
    if (dup) continue;
    if (COUNT) {
      atomicAdd(cnt + b, 1);
    } else {
      const unsigned r = rank ? static_cast<unsigned>(rank[e]) : static_cast<unsigned>(e);
      const int at = atomicAdd(cnt + b, 1);
      xkey[at] = (static_cast<u64>(r) << 32) | static_cast<unsigned>(e);
    }
  }
}

// One workgroup per block: sort its exec keys (LDS up to kCapX keys, global
// memory beyond) and write the element numbers.
__global__ void __launch_bounds__(kThreads) rows_sort_exec(const int* __restrict__ xoff, int nblk,
                                                           u64* xkey, int* __restrict__ exec) {
  __shared__ u64 s_k[kCapX];
  for (int b = blockIdx.x; b < nblk; b += gridDim.x) {
    const int q0 = xoff[b], len = xoff[b + 1] - q0;
    u64* a = len <= kCapX ? s_k : xkey + q0;
    if (len <= kCapX)
      for (int i = threadIdx.x; i < len; i += kThreads) s_k[i] = xkey[q0 + i];
    __syncthreads();
    block_sort<u64>(a, len);
    __syncthreads();
    for (int i = threadIdx.x; i < len; i += kThreads) {
      const u64 v = a[i];
      xkey[q0 + i] = v;
      exec[q0 + i] = static_cast<int>(static_cast<unsigned>(v & 0xffffffffull));
    }
    __syncthreads();
  }
}

// Per block: the largest CSR slice and exec list (scal[2], scal[1]).
__global__ void rows_stats(const int* __restrict__ rowptr, int nnodes, int rd, int k, int nblk,
                           const int* __restrict__ xoff, int* __restrict__ scal) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nblk) return;
  const long long t0 = static_cast<long long>(b) * k;
  const long long t1 = t0 + k < nnodes ? t0 + k : nnodes;
  atomicMax(scal + 2, rowptr[t1 * rd] - rowptr[t0 * rd]);
  atomicMax(scal + 1, xoff[b + 1] - xoff[b]);
}

// One thread per row node: its records' codes and column positions.  Flags
// (scal[4]): 1 a column outside the pattern, 2 a position beyond 255, 4 an
// element missing from its block's exec list.  scal[0]: the largest code.
__global__ void rows_records(const u64* __restrict__ key, const int* __restrict__ tp, int nnodes,
                             int e0, int k, const u64* __restrict__ xkey,
                             const int* __restrict__ xoff, const int* __restrict__ cmap,
                             int carity, int e1, const int* __restrict__ rowptr,
                             const int* __restrict__ colidx, int rd, int cd, int W,
                             int* __restrict__ code32, unsigned* __restrict__ pos,
                             int* __restrict__ scal) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nnodes) return;
  const int c0 = tp[n], c1 = tp[n + 1];
  if (c0 == c1) return;
  const int b = n / k, x0 = xoff[b], x1 = xoff[b + 1];
  const int s0 = rowptr[static_cast<long long>(n) * rd];
  const int ncn = (rowptr[static_cast<long long>(n) * rd + 1] - s0) / cd;
  int flags = 0, cmax = 0;
  for (int c = c0; c < c1; ++c) {
    const u64 kk = key[c];
    const unsigned low = static_cast<unsigned>(kk & 0xffffffffull);
    const int e = static_cast<int>(low / static_cast<unsigned>(e0));
    const int i0 = static_cast<int>(low - static_cast<unsigned>(e) * e0);
    const u64 want = (kk & 0xffffffff00000000ull) | static_cast<unsigned>(e);
    int lo = x0, hi = x1 - 1;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (xkey[mid] < want)
        lo = mid + 1;
      else
        hi = mid;
    }
    if (lo > hi || xkey[lo] != want) flags |= 4;
    const int code = (lo - x0) * e0 + i0;
    code32[c] = code;
    cmax = code > cmax ? code : cmax;
    for (int w = 0; w < W; ++w) {
      unsigned word = 0;
      for (int j = 4 * w; j < 4 * w + 4 && j < e1; ++j) {
        const int m = cmap[static_cast<long long>(e) * carity + j];
        int a = 0, z = ncn - 1, at = -1;
        while (a <= z) {
          const int mid = (a + z) >> 1;
          const int v = colidx[s0 + mid * cd] / cd;
          if (v == m) {
            at = mid;
            break;
          }
          if (v < m)
            a = mid + 1;
          else
            z = mid - 1;
        }
        if (at < 0) {
          flags |= 1;
          at = 0;
        }
        if (at > 255) {
          flags |= 2;
          at = 0;
        }
        word |= static_cast<unsigned>(at) << (8 * (j - 4 * w));
      }
      pos[static_cast<long long>(c) * W + w] = word;
    }
  }
  atomicMax(scal, cmax);
  if (flags) atomicOr(scal + 4, flags);
}

__global__ void rows_pack16(const int* __restrict__ c32, long long n,
                            unsigned short* __restrict__ out) {
  const long long q = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (q < n) out[q] = static_cast<unsigned short>(c32[q]);
}

struct RowsWork {
  size_t cnt, key, longlist, nlong, sums, total, cntb, xkey, scal, bytes;
};

RowsWork rows_layout(long long nnodes, long long ninc, long long nblk) {
  RowsWork w;
  size_t o = 0;
  auto take = [&](size_t b) {
    size_t at = o;
    o += align256(b > 0 ? b : 1);
    return at;
  };
  const size_t I = sizeof(int);
  w.cnt = take(I * (nnodes + 1));
  w.key = take(sizeof(u64) * ninc);
  w.longlist = take(I * (nnodes + 1));
  w.nlong = take(I);
  long long ns = scan_tiles_for(nnodes + 1);
  long long nb2 = scan_tiles_for(nblk + 1);
  w.sums = take(sizeof(long long) * (ns > nb2 ? ns : nb2));
  w.total = take(sizeof(long long) * 2);
  w.cntb = take(I * (nblk + 1));
  w.xkey = take(sizeof(u64) * ninc);
  w.scal = take(I * 8);
  w.bytes = o;
  return w;
}

}  // namespace

extern "C" {

long long op2hip_mat_rows_workspace(int nelems, int nnodes, int e0, int k) {
  if (nelems < 0 || nnodes < 0 || e0 <= 0 || k <= 0) return -1;
  const long long nblk = (static_cast<long long>(nnodes) + k - 1) / k;
  return static_cast<long long>(
      rows_layout(nnodes, static_cast<long long>(nelems) * e0, nblk).bytes);
}

// Build a row-block Mat plan on the device.  rmap / cmap: device row and
// column maps (arities rarity / carity, device numbering), rows i0 < e0 and
// columns j < e1 of each element; rank (device, may be NULL for the
// identity) the user order of each element; rowptr / colidx: the device
// CSR pattern with rd x cd blocks (scalar rows).  Outputs (device,
// caller-allocated): tp (nnodes + 1), xoff (nblk + 1), exec and code32
// (nelems * e0 each, an upper bound for exec), pos (nelems * e0 * W words,
// W = ceil(e1 / 4)).  scalars (host, 5): the largest code, the largest exec
// list, the largest CSR slice, the exec entries summed over blocks, and the
// flags of rows_records.  Waits for the stream once.
int op2hip_mat_rows_build(int nelems, int nnodes, const int* rmap, int rarity, int e0,
                          const int* cmap, int carity, int e1, const int* rank,
                          const int* rowptr, const int* colidx, int rd, int cd, int k, void* work,
                          long long work_bytes, int* tp, int* xoff, int* exec, int* code32,
                          unsigned* pos, long long* scalars, void* stream) {
  OP2HIP_REQUIRE(nelems >= 0 && nnodes >= 0 && rmap && cmap && rowptr && colidx && tp && xoff &&
                     exec && code32 && pos && scalars && work && k > 0,
                 "bad row-block plan arguments");
  OP2HIP_REQUIRE(e0 > 0 && e0 <= rarity && e0 <= kMaxCols && e1 > 0 && e1 <= carity && rd > 0 &&
                     cd > 0,
                 "bad row-block plan extents");
  const long long ninc = static_cast<long long>(nelems) * e0;
  OP2HIP_REQUIRE(ninc <= 0x7fffffffLL, "row-block plan of %lld records, beyond int32", ninc);
  const int nblk = static_cast<int>((static_cast<long long>(nnodes) + k - 1) / k);
  RowsWork w = rows_layout(nnodes, ninc, nblk);
  OP2HIP_REQUIRE(work_bytes >= static_cast<long long>(w.bytes),
                 "row-block plan workspace of %lld bytes, %zu needed", work_bytes, w.bytes);
  const int W = (e1 + 3) / 4;
  hipStream_t s = static_cast<hipStream_t>(stream);
  char* base = static_cast<char*>(work);
  int* cnt = at<int>(base, w.cnt);
  u64* key = at<u64>(base, w.key);
  int* longlist = at<int>(base, w.longlist);
  int* nlong = at<int>(base, w.nlong);
  long long* sums = at<long long>(base, w.sums);
  long long* total = at<long long>(base, w.total);
  int* cntb = at<int>(base, w.cntb);
  u64* xkey = at<u64>(base, w.xkey);
  int* scal = at<int>(base, w.scal);
  Cols C{};
  C.n = e0;
  for (int c = 0; c < e0; ++c) {
    C.v[c] = rmap;
    C.dim[c] = rarity;
    C.col[c] = c;
  }
  hipLaunchKernelGGL(zero_ints, dim3((nnodes + 1 + 255) / 256), dim3(256), 0, s, cnt,
                     static_cast<long long>(nnodes) + 1);
  hipLaunchKernelGGL(zero_ints, dim3((nblk + 1 + 255) / 256), dim3(256), 0, s, cntb,
                     static_cast<long long>(nblk) + 1);
  hipLaunchKernelGGL(zero_ints, dim3(1), dim3(64), 0, s, nlong, 1LL);
  hipLaunchKernelGGL(zero_ints, dim3(1), dim3(64), 0, s, scal, 8LL);
  // records per node, in ascending (rank, i0)
  if (ninc > 0)
    hipLaunchKernelGGL(count_targets, dim3((ninc + 255) / 256), dim3(256), 0, s, C, nelems,
                       nnodes, cnt);
  exclusive_scan(cnt, nnodes, tp, cnt, sums, total, s);
  if (ninc > 0)
    hipLaunchKernelGGL(scatter_keys, dim3((ninc + 255) / 256), dim3(256), 0, s, C, nelems,
                       nnodes, rank, cnt, key);
  if (nnodes > 0) {
    long long threads = static_cast<long long>(nnodes) * 64;
    hipLaunchKernelGGL(sort_targets_wave, dim3((threads + 255) / 256), dim3(256), 0, s, tp,
                       nnodes, key, longlist, nlong);
    hipLaunchKernelGGL(sort_targets_long, dim3(kLongGrid), dim3(kThreads), 0, s, tp, key,
                       longlist, nlong);
  }
  // exec lists per block, in ascending rank
  const dim3 egrid((nelems + 255) / 256 > 0 ? (nelems + 255) / 256 : 1);
  if (nelems > 0)
    hipLaunchKernelGGL(rows_exec_entries<true>, egrid, dim3(256), 0, s, rmap, rarity, e0, nelems,
                       nnodes, k, rank, cntb, xkey);
  exclusive_scan(cntb, nblk, xoff, cntb, sums, total, s);
  if (nelems > 0)
    hipLaunchKernelGGL(rows_exec_entries<false>, egrid, dim3(256), 0, s, rmap, rarity, e0, nelems,
                       nnodes, k, rank, cntb, xkey);
  if (nblk > 0) {
    const int grid = nblk < 65536 ? nblk : 65536;
    hipLaunchKernelGGL(rows_sort_exec, dim3(grid), dim3(kThreads), 0, s, xoff, nblk, xkey, exec);
    hipLaunchKernelGGL(rows_stats, dim3((nblk + 255) / 256), dim3(256), 0, s, rowptr, nnodes, rd,
                       k, nblk, xoff, scal);
  }
  if (nnodes > 0)
    hipLaunchKernelGGL(rows_records, dim3((nnodes + 255) / 256), dim3(256), 0, s, key, tp, nnodes,
                       e0, k, xkey, xoff, cmap, carity, e1, rowptr, colidx, rd, cd, W, code32, pos,
                       scal);
  int host[8];
  int hsum = 0;
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(host, scal, sizeof(host), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipMemcpyAsync(&hsum, xoff + nblk, sizeof(int), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) {
    op2hip::set_error("row-block plan build failed: %s", hipGetErrorString(e));
    return OP2HIP_ERR_HIP;
  }
  scalars[0] = host[0];
  scalars[1] = host[1];
  scalars[2] = host[2];
  scalars[3] = hsum;
  scalars[4] = host[4];
  return OP2HIP_OK;
}

// The codes of a row-block plan narrowed to 16 bits (every code < 65536).
int op2hip_mat_rows_pack16(const int* code32, long long n, unsigned short* out, void* stream) {
  OP2HIP_REQUIRE(n >= 0 && (n == 0 || (code32 && out)), "bad pack arguments");
  if (n > 0)
    hipLaunchKernelGGL(rows_pack16, dim3((n + 255) / 256), dim3(256), 0,
                       static_cast<hipStream_t>(stream), code32, n, out);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    op2hip::set_error("pack failed: %s", hipGetErrorString(e));
    return OP2HIP_ERR_HIP;
  }
  return OP2HIP_OK;
}

}  // extern "C"
