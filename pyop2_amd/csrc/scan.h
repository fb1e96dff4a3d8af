// Exclusive scan of int32 counts on the device (64-bit partial sums, int32
// offsets), shared by the sparsity and plan builds.  Three launches: tile
// sums, one workgroup scanning them in tile order, and the tiles' own scans.
// Results do not depend on timing (fixed summation order).
#pragma once

#include <hip/hip_runtime.h>

namespace op2hip_scan {

constexpr int kScanItems = 4;  // values per thread
constexpr int kScanBlock = 256;
constexpr int kScanTile = kScanItems * kScanBlock;

// Tile sums a scan of n values needs.
inline long long scan_tiles_for(long long n) { return (n + kScanTile - 1) / kScanTile + 1; }

static __device__ inline long long block_exclusive(long long v, long long* s_warp, long long* total) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  long long x = v;
  for (int off = 1; off < 64; off <<= 1) {
    long long y = __shfl_up(x, off, 64);
    if (lane >= off) x += y;
  }
  if (lane == 63) s_warp[w] = x;
  __syncthreads();
  long long before = 0, all = 0;
  for (int k = 0; k < kScanBlock / 64; ++k) {
    if (k < w) before += s_warp[k];
    all += s_warp[k];
  }
  __syncthreads();
  *total = all;
  return before + x - v;
}

static __global__ void __launch_bounds__(kScanBlock) scan_tiles(const int* __restrict__ in, int n,
                                                         long long* __restrict__ sums) {
  __shared__ long long s_warp[kScanBlock / 64];
  long long base = static_cast<long long>(blockIdx.x) * kScanTile + threadIdx.x * kScanItems;
  long long v = 0;
  for (int j = 0; j < kScanItems; ++j)
    if (base + j < n) v += in[base + j];
  long long total;
  block_exclusive(v, s_warp, &total);
  if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// One workgroup: exclusive scan of the tile sums in tile order; the grand
// total goes to *total and, as an int32, to out_last (the array's entry n).
static __global__ void __launch_bounds__(1024) scan_sums(long long* __restrict__ sums, int nb,
                                                  long long* __restrict__ total,
                                                  int* __restrict__ out_last) {
  __shared__ long long s_warp[16];
  __shared__ long long carry;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int b0 = 0; b0 < nb; b0 += 1024) {
    int i = b0 + threadIdx.x;
    long long v = i < nb ? sums[i] : 0;
    long long x = v;
    for (int off = 1; off < 64; off <<= 1) {
      long long y = __shfl_up(x, off, 64);
      if (lane >= off) x += y;
    }
    if (lane == 63) s_warp[w] = x;
    __syncthreads();
    long long before = carry, all = 0;
    for (int k = 0; k < 16; ++k) {
      if (k < w) before += s_warp[k];
      all += s_warp[k];
    }
    if (i < nb) sums[i] = before + x - v;
    __syncthreads();
    if (threadIdx.x == 0) carry += all;
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    *total = carry;
    if (out_last) *out_last = carry <= 0x7fffffffLL ? static_cast<int>(carry) : -1;
  }
}

// in and out2 may be the same array (a counter array turned into the
// offsets it is scanned to): each thread loads its values before any store.
static __global__ void __launch_bounds__(kScanBlock) scan_apply(const int* in, int n,
                                                         const long long* __restrict__ sums,
                                                         int* __restrict__ out, int* out2) {
  __shared__ long long s_warp[kScanBlock / 64];
  long long base = static_cast<long long>(blockIdx.x) * kScanTile + threadIdx.x * kScanItems;
  int x[kScanItems];
  long long v = 0;
  for (int j = 0; j < kScanItems; ++j) {
    x[j] = base + j < n ? in[base + j] : 0;
    v += x[j];
  }
  long long total;
  long long run = sums[blockIdx.x] + block_exclusive(v, s_warp, &total);
  for (int j = 0; j < kScanItems; ++j) {
    if (base + j < n) {
      int o = run <= 0x7fffffffLL ? static_cast<int>(run) : -1;
      out[base + j] = o;
      if (out2) out2[base + j] = o;
    }
    run += x[j];
  }
}

// out[0..n) = exclusive scan of in[0..n), out[n] = total (int32, -1 when
// beyond int32), *total = the total as int64; out2 (may be null) receives a
// copy of out[0..n).  sums needs ceil(n / kScanTile) entries.
static inline void exclusive_scan(const int* in, int n, int* out, int* out2, long long* sums,
                    long long* total, hipStream_t s) {
  int nb = (n + kScanTile - 1) / kScanTile;
  if (nb > 0)
    hipLaunchKernelGGL(scan_tiles, dim3(nb), dim3(kScanBlock), 0, s, in, n, sums);
  hipLaunchKernelGGL(scan_sums, dim3(1), dim3(1024), 0, s, sums, nb, total, out + n);
  if (nb > 0)
    hipLaunchKernelGGL(scan_apply, dim3(nb), dim3(kScanBlock), 0, s, in, n, sums, out, out2);
}

}  // namespace op2hip_scan
