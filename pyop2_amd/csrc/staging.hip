// Per-block staging lists of a plan: for block b, whose elements are
// elems[xoff[b] .. xoff[b + 1]), and a map of arity dim, the
// sorted distinct targets its elements name (the reference's ind_map, with
// ind_sizes and ind_offs) and, per element slot and map column, the position
// of that target in the block's list (its loc_map): pyop2/plan.pyx:112-212,
// there built with numpy.unique per block.  Here one workgroup per block
// sorts the block's (target, slot) keys in LDS (bitonic), flags the first
// key of every target and scans the flags.
//
// Layout: block b's slots are xoff[b] * dim .. xoff[b + 1] * dim (ind_offs =
// xoff * dim); ind_map holds the block's list at the start
// of that range (ind_sizes[b] entries), loc_map the position of each slot's
// target.  op2hip_stage_compact packs the lists end to end (so, sn).
#include "common.h"

namespace {

constexpr int kThreads = 256;
// most keys a block may hold in LDS (8 bytes each), further bounded by the
// device's LDS per workgroup (op2hip_stage_max_keys)
constexpr int kMaxKeys = 8192;

typedef unsigned long long u64;

__global__ void __launch_bounds__(kThreads) stage_lists(
    const int* __restrict__ xoff, const int* __restrict__ elems,
    const int* __restrict__ map, int dim, int cap,
    long long nslots, int* __restrict__ ind_map, unsigned short* __restrict__ loc_map,
    int* __restrict__ ind_sizes, int* __restrict__ err) {
  extern __shared__ u64 keys[];
  __shared__ int s_cnt[kThreads];
  const int b = blockIdx.x, tid = threadIdx.x;
  const long long off = xoff[b];
  const int ne = xoff[b + 1] - xoff[b];
  const long long nk = static_cast<long long>(ne) * dim;
  const long long base = off * dim;
  // the block's keys must fit the LDS and its slots the outputs (the same
  // answer for every thread of the block)
  if (ne < 0 || nk > cap || base < 0 || base + nk > nslots) {
    if (tid == 0) {
      atomicOr(err, 1);
      ind_sizes[b] = 0;
    }
    return;
  }
  int P = 1;
  while (P < nk) P <<= 1;
  // key: target in the high word, slot within the block in the low word
  for (int i = tid; i < P; i += kThreads) {
    u64 k = ~0ull;
    if (i < nk) {
      const int e = elems[off + i / dim];
      const int t = map[static_cast<long long>(e) * dim + i % dim];
      k = (static_cast<u64>(static_cast<unsigned>(t)) << 32) | static_cast<unsigned>(i);
    }
    keys[i] = k;
  }
  __syncthreads();
  // bitonic sort of keys[0, P), ascending
  for (int k = 2; k <= P; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = tid; i < P; i += kThreads) {
        const int ixj = i ^ j;
        if (ixj > i) {
          const u64 a = keys[i], c = keys[ixj];
          const bool up = (i & k) == 0;
          if ((a > c) == up) {
            keys[i] = c;
            keys[ixj] = a;
          }
        }
      }
      __syncthreads();
    }
  }
  // the first key of each target starts a list entry: thread tid takes the
  // consecutive range [lo, hi) of the sorted keys
  const int per = (P + kThreads - 1) / kThreads;
  const int lo = tid * per < P ? tid * per : P;
  const int hi = lo + per < P ? lo + per : P;
  auto first = [&](int i) -> bool {
    return i < nk && (i == 0 || (keys[i] >> 32) != (keys[i - 1] >> 32));
  };
  int cnt = 0;
  for (int i = lo; i < hi; ++i) cnt += first(i) ? 1 : 0;
  s_cnt[tid] = cnt;
  __syncthreads();
  if (tid == 0) {
    int run = 0;
    for (int t = 0; t < kThreads; ++t) {
      const int c = s_cnt[t];
      s_cnt[t] = run;
      run += c;
    }
    ind_sizes[b] = run;
  }
  __syncthreads();
  int u = s_cnt[tid] - 1;
  for (int i = lo; i < hi && i < nk; ++i) {
    const u64 k = keys[i];
    if (first(i)) {
      ++u;
      ind_map[base + u] = static_cast<int>(k >> 32);
    }
    const long long slot = static_cast<long long>(k & 0xffffffffull);
    loc_map[base + slot] = static_cast<unsigned short>(u);
  }
}

// so[0..n) = exclusive scan of sizes, so[n] = the total: one workgroup.
__global__ void __launch_bounds__(1024) stage_scan(const int* __restrict__ sizes, int n,
                                                   int* __restrict__ so) {
  __shared__ long long s[1024];
  __shared__ long long carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int b0 = 0; b0 < n; b0 += 1024) {
    const int i = b0 + threadIdx.x;
    const long long v = i < n ? sizes[i] : 0;
    s[threadIdx.x] = v;
    __syncthreads();
    // inclusive Hillis-Steele scan of the 1024 values
    for (int d = 1; d < 1024; d <<= 1) {
      const long long add = threadIdx.x >= d ? s[threadIdx.x - d] : 0;
      __syncthreads();
      s[threadIdx.x] += add;
      __syncthreads();
    }
    if (i < n) so[i] = static_cast<int>(carry + s[threadIdx.x] - v);
    __syncthreads();
    if (threadIdx.x == 1023) carry += s[1023];
    __syncthreads();
  }
  if (threadIdx.x == 0) so[n] = static_cast<int>(carry);
}

// sn[so[b] + i] = ind_map[xoff[b] * dim + i] for i < sizes[b]
__global__ void __launch_bounds__(kThreads) stage_pack(const int* __restrict__ xoff, int dim,
                                                       const int* __restrict__ sizes,
                                                       const int* __restrict__ so,
                                                       const int* __restrict__ ind_map,
                                                       int* __restrict__ sn) {
  const int b = blockIdx.x;
  const long long base = static_cast<long long>(xoff[b]) * dim;
  const int n = sizes[b], dst = so[b];
  for (int i = threadIdx.x; i < n; i += kThreads) sn[dst + i] = ind_map[base + i];
}

}  // namespace

extern "C" {

// The most keys ((xoff[b + 1] - xoff[b]) * dim) one block may have on the current
// device: a power of two whose 8-byte keys fit its LDS per workgroup beside
// the kernel's 1 KB of counters, at most 8192; -1 when the device cannot be
// queried.
int op2hip_stage_max_keys(void) {
  int dev = 0, lds = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
    return -1;
  int cap = kMaxKeys;
  while (cap > 1 && static_cast<long long>(cap) * 8 + 4 * kThreads > lds) cap >>= 1;
  return cap;
}

// Staging lists of nblocks blocks (device arrays: xoff, nblocks + 1
// ascending offsets into elems; elems; the map, arity dim; device
// numbering).  maxkeys: at least the largest (xoff[b + 1] - xoff[b]) * dim,
// at most op2hip_stage_max_keys().  Outputs (device, caller-allocated, nslots
// = entries of ind_map and loc_map, at least xoff[nblocks] * dim): ind_map,
// loc_map, ind_sizes (nblocks); *err (device int) gets bit 1 set for a block
// with more than maxkeys keys or slots beyond nslots (its size is then 0).
int op2hip_stage_lists(int nblocks, const int* xoff, const int* elems, const int* map, int dim,
                       int maxkeys, long long nslots, int* ind_map, unsigned short* loc_map,
                       int* ind_sizes, int* err, void* stream) {
  const int most = op2hip_stage_max_keys();
  OP2HIP_REQUIRE(nblocks >= 0 && dim >= 1 && dim <= 1024 && maxkeys >= 0 && maxkeys <= most &&
                     nslots >= 0,
                 "bad staging list arguments (%d blocks, dim %d, %d keys, at most %d)", nblocks,
                 dim, maxkeys, most);
  if (nblocks == 0) return OP2HIP_OK;
  OP2HIP_REQUIRE(xoff && elems && map && ind_map && loc_map && ind_sizes && err,
                 "null staging list argument");
  int cap = 1;
  while (cap < maxkeys) cap <<= 1;
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(stage_lists, dim3(nblocks), dim3(kThreads), sizeof(u64) * cap, s, xoff,
                     elems, map, dim, cap, nslots, ind_map, loc_map, ind_sizes, err);
  OP2HIP_CHECK(hipGetLastError());
  return OP2HIP_OK;
}

// The lists end to end: so (nblocks + 1) the exclusive scan of ind_sizes,
// sn[so[b] ..] block b's list (sn: at least the sum of ind_sizes entries,
// which the caller bounds by nslots).
int op2hip_stage_compact(int nblocks, const int* xoff, int dim, const int* ind_sizes,
                         const int* ind_map, int* so, int* sn, void* stream) {
  OP2HIP_REQUIRE(nblocks >= 0 && dim >= 1, "bad staging compaction arguments");
  OP2HIP_REQUIRE(xoff && ind_sizes && ind_map && so && sn, "null staging compaction argument");
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(stage_scan, dim3(1), dim3(1024), 0, s, ind_sizes, nblocks, so);
  if (nblocks > 0)
    hipLaunchKernelGGL(stage_pack, dim3(nblocks), dim3(kThreads), 0, s, xoff, dim, ind_sizes,
                       so, ind_map, sn);
  OP2HIP_CHECK(hipGetLastError());
  return OP2HIP_OK;
}

}  // extern "C"
