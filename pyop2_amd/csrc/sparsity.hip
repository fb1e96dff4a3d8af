// Device construction of a Sparsity's CSR structure.
//
// Replaces build_sparsity_pattern_seq (pyop2/sparsity_utils.cxx:5-51), which
// inserts, for every map pair, element e, row-map column i, r < rmult,
// col-map column d and c < cmult, the column cmult*cmap[e][d]+c into the
// std::set of row rmult*rmap[e][i]+r, then emits rowptr and the sorted
// unique colidx.  The result is a pure function of the maps, so the device
// build below is bit-exact to it:
//
//   1. count node-level candidates per row (cdim per (e, i) incidence),
//   2. exclusive scan -> candidate offsets, scatter candidates,
//   3. per row, one wavefront ranks every candidate: a candidate is kept
//      when no earlier candidate has its value, and lands at the number of
//      kept values smaller than it (sorted unique, no sort network needed),
//   4. scan of the unique counts -> node rowptr,
//   5. expand to rmult x cmult blocks: row rmult*n+r holds
//      {cmult*m+c : m in cols(n), c < cmult} in order.
#include <vector>

#include "common.h"

namespace {

constexpr int kScanBlock = 256;

// Inclusive-then-shifted three-phase exclusive scan of n int64 values.
__global__ void scan_block_sums(const long long* __restrict__ in, long long n,
                                long long* __restrict__ sums) {
  __shared__ long long s[kScanBlock];
  long long i = static_cast<long long>(blockIdx.x) * kScanBlock + threadIdx.x;
  s[threadIdx.x] = i < n ? in[i] : 0;
  __syncthreads();
  for (int off = kScanBlock / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) s[threadIdx.x] += s[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) sums[blockIdx.x] = s[0];
}

__global__ void scan_sums_serial(long long* __restrict__ sums, long long nb,
                                 long long* __restrict__ total) {
  // one block: chunked scan of the block sums in fixed order
  __shared__ long long s[kScanBlock];
  __shared__ long long carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (long long base = 0; base < nb; base += kScanBlock) {
    long long i = base + threadIdx.x;
    long long v = i < nb ? sums[i] : 0;
    s[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < kScanBlock; off <<= 1) {
      long long t = threadIdx.x >= off ? s[threadIdx.x - off] : 0;
      __syncthreads();
      s[threadIdx.x] += t;
      __syncthreads();
    }
    if (i < nb) sums[i] = carry + s[threadIdx.x] - v;  // exclusive
    __syncthreads();
    if (threadIdx.x == kScanBlock - 1) carry += s[kScanBlock - 1];
    __syncthreads();
  }
  if (threadIdx.x == 0) *total = carry;
}

template <typename Tout>
__global__ void scan_apply(const long long* __restrict__ in, long long n,
                           const long long* __restrict__ sums, Tout* __restrict__ out) {
  __shared__ long long s[kScanBlock];
  long long i = static_cast<long long>(blockIdx.x) * kScanBlock + threadIdx.x;
  long long v = i < n ? in[i] : 0;
  s[threadIdx.x] = v;
  __syncthreads();
  for (int off = 1; off < kScanBlock; off <<= 1) {
    long long t = threadIdx.x >= off ? s[threadIdx.x - off] : 0;
    __syncthreads();
    s[threadIdx.x] += t;
    __syncthreads();
  }
  if (i < n) out[i] = static_cast<Tout>(sums[blockIdx.x] + s[threadIdx.x] - v);
}

// out[0..n) = exclusive scan of in, out[n] = total; `sums` needs
// ceil(n/256) entries and `total` one.
template <typename Tout>
int exclusive_scan(const long long* in, long long n, Tout* out, long long* sums,
                   long long* total_dev, hipStream_t s) {
  long long nb = (n + kScanBlock - 1) / kScanBlock;
  if (nb == 0) nb = 1;
  hipLaunchKernelGGL(scan_block_sums, dim3(nb), dim3(kScanBlock), 0, s, in, n, sums);
  hipLaunchKernelGGL(scan_sums_serial, dim3(1), dim3(kScanBlock), 0, s, sums, nb, total_dev);
  hipLaunchKernelGGL(scan_apply<Tout>, dim3(nb), dim3(kScanBlock), 0, s, in, n, sums, out);
  OP2HIP_CHECK(hipGetLastError());
  return OP2HIP_OK;
}

__global__ void count_candidates(const int* __restrict__ rmap, int rdim, int cdim, long long nel,
                                 long long* __restrict__ count) {
  long long k = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= nel * rdim) return;
  atomicAdd(reinterpret_cast<unsigned long long*>(count + rmap[k]),
            static_cast<unsigned long long>(cdim));
}

__global__ void scatter_candidates(const int* __restrict__ rmap, int rdim,
                                   const int* __restrict__ cmap, int cdim, long long nel,
                                   unsigned long long* __restrict__ cursor,
                                   int* __restrict__ cand) {
  long long k = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= nel * rdim) return;
  long long e = k / rdim;
  int row = rmap[k];
  unsigned long long pos = atomicAdd(cursor + row, static_cast<unsigned long long>(cdim));
  const int* c = cmap + e * cdim;
  for (int d = 0; d < cdim; ++d) cand[pos + d] = c[d];
}

// One wavefront per row, two passes of O(L^2) comparisons over the row's L
// candidates.  Pass 1 flags the first occurrence of each value.
__global__ void __launch_bounds__(256) flag_first(const long long* __restrict__ cptr,
                                                  long long nrows, const int* __restrict__ cand,
                                                  unsigned char* __restrict__ first) {
  const int lane = threadIdx.x & 63;
  long long row = (static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
  if (row >= nrows) return;
  const long long b = cptr[row];
  const long long L = cptr[row + 1] - b;
  const int* c = cand + b;
  for (long long j = lane; j < L; j += 64) {
    int v = c[j];
    unsigned char f = 1;
    for (long long k = 0; k < j; ++k)
      if (c[k] == v) {
        f = 0;
        break;
      }
    first[b + j] = f;
  }
}

// Pass 2: a first occurrence lands at the number of distinct smaller values
// of its row, which yields the row's sorted unique columns.
__global__ void __launch_bounds__(256) rank_unique(const long long* __restrict__ cptr,
                                                   long long nrows, const int* __restrict__ cand,
                                                   const unsigned char* __restrict__ first,
                                                   int* __restrict__ sorted,
                                                   long long* __restrict__ ucount) {
  const int lane = threadIdx.x & 63;
  long long row = (static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
  if (row >= nrows) return;
  const long long b = cptr[row];
  const long long L = cptr[row + 1] - b;
  const int* c = cand + b;
  const unsigned char* f = first + b;
  int kept = 0;
  for (long long j = lane; j < L; j += 64) {
    if (!f[j]) continue;
    int v = c[j];
    long long smaller = 0;
    for (long long k = 0; k < L; ++k) smaller += (f[k] && c[k] < v) ? 1 : 0;
    sorted[b + smaller] = v;
    ++kept;
  }
  for (int off = 32; off > 0; off >>= 1) kept += __shfl_xor(kept, off, 64);
  if (lane == 0) ucount[row] = kept;
}

// Expand the node-level pattern into rmult x cmult blocks.
__global__ void expand_rowptr(const int* __restrict__ nptr, long long nrows, int rmult, int cmult,
                              int* __restrict__ rowptr) {
  long long R = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  long long lsize = nrows * rmult;
  if (R > lsize) return;
  if (R == lsize) {
    rowptr[R] = nptr[nrows] * rmult * cmult;
    return;
  }
  long long n = R / rmult;
  int r = static_cast<int>(R % rmult);
  long long u = nptr[n + 1] - nptr[n];
  rowptr[R] = static_cast<int>(static_cast<long long>(nptr[n]) * rmult * cmult + r * u * cmult);
}

__global__ void expand_colidx(const int* __restrict__ nptr, const long long* __restrict__ cptr,
                              const int* __restrict__ sorted, long long nrows, int rmult,
                              int cmult, const int* __restrict__ rowptr,
                              int* __restrict__ colidx, int* __restrict__ nnz_rows) {
  long long R = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (R >= nrows * rmult) return;
  long long n = R / rmult;
  long long u = nptr[n + 1] - nptr[n];
  const int* src = sorted + cptr[n];
  int* dst = colidx + rowptr[R];
  for (long long k = 0; k < u; ++k)
    for (int c = 0; c < cmult; ++c) dst[k * cmult + c] = cmult * src[k] + c;
  if (nnz_rows) nnz_rows[R] = static_cast<int>(u * cmult);
}

struct SparsityBuild {
  long long nrows = 0;
  int rmult = 1, cmult = 1;
  long long* cptr = nullptr;    // nrows + 1 candidate offsets
  int* sorted = nullptr;        // candidates, sorted unique prefix per row
  int* nptr = nullptr;          // nrows + 1 node-level rowptr
  long long nnz = 0;
};

void free_build(SparsityBuild* sb) {
  if (!sb) return;
  if (sb->cptr) hipFree(sb->cptr);
  if (sb->sorted) hipFree(sb->sorted);
  if (sb->nptr) hipFree(sb->nptr);
  delete sb;
}

}  // namespace

extern "C" {

// Phase 1: build the pattern and write rowptr (device, nrows*rmult + 1).
// rowmaps/colmaps are host arrays of device pointers to int32 map values;
// nelems[m] is the iteration-set size of pair m.  *nnz_out receives the
// number of nonzeros; the handle keeps what phase 2 needs.
int op2hip_sparsity_build(int rmult, int cmult, int nrows, int nmaps, const int* const* rowmaps,
                          const int* rdims, const int* const* colmaps, const int* cdims,
                          const long long* nelems, int* rowptr, void** handle,
                          long long* nnz_out, void* stream) {
  OP2HIP_REQUIRE(rmult > 0 && cmult > 0 && nrows >= 0 && nmaps > 0, "bad sparsity arguments");
  OP2HIP_REQUIRE(rowptr && handle && nnz_out, "null sparsity output");
  hipStream_t s = static_cast<hipStream_t>(stream);
  SparsityBuild* sb = new SparsityBuild();
  sb->nrows = nrows;
  sb->rmult = rmult;
  sb->cmult = cmult;
  long long* count = nullptr;
  long long* sums = nullptr;
  long long* total = nullptr;
  long long* ucount = nullptr;
  int* cand = nullptr;
  unsigned char* first = nullptr;
  int rc = OP2HIP_OK;
  long long ncand = 0;
  long long nb = (static_cast<long long>(nrows) + 1 + kScanBlock - 1) / kScanBlock + 1;
  auto fail = [&](int code) {
    if (count) hipFree(count);
    if (sums) hipFree(sums);
    if (total) hipFree(total);
    if (ucount) hipFree(ucount);
    if (cand) hipFree(cand);
    if (first) hipFree(first);
    free_build(sb);
    return code;
  };
#define SP_CHECK(x)                                                          \
  do {                                                                       \
    hipError_t _e = (x);                                                     \
    if (_e != hipSuccess) {                                                  \
      op2hip::set_error("%s failed: %s", #x, hipGetErrorString(_e));         \
      return fail(OP2HIP_ERR_HIP);                                           \
    }                                                                        \
  } while (0)
  SP_CHECK(hipMalloc(&count, sizeof(long long) * (nrows + 1)));
  SP_CHECK(hipMalloc(&sums, sizeof(long long) * nb));
  SP_CHECK(hipMalloc(&total, sizeof(long long)));
  SP_CHECK(hipMalloc(&sb->cptr, sizeof(long long) * (nrows + 1)));
  SP_CHECK(hipMemsetAsync(count, 0, sizeof(long long) * (nrows + 1), s));
  for (int m = 0; m < nmaps; ++m) {
    long long work = nelems[m] * rdims[m];
    if (work == 0) continue;
    hipLaunchKernelGGL(count_candidates, dim3((work + 255) / 256), dim3(256), 0, s, rowmaps[m],
                       rdims[m], cdims[m], nelems[m], count);
  }
  rc = exclusive_scan<long long>(count, nrows, sb->cptr, sums, total, s);
  if (rc) return fail(rc);
  SP_CHECK(hipMemcpyAsync(&ncand, total, sizeof(long long), hipMemcpyDeviceToHost, s));
  SP_CHECK(hipStreamSynchronize(s));
  SP_CHECK(hipMemcpyAsync(sb->cptr + nrows, total, sizeof(long long), hipMemcpyDeviceToDevice, s));
  SP_CHECK(hipMalloc(&cand, sizeof(int) * (ncand > 0 ? ncand : 1)));
  SP_CHECK(hipMalloc(&sb->sorted, sizeof(int) * (ncand > 0 ? ncand : 1)));
  // cursor = copy of cptr
  SP_CHECK(hipMemcpyAsync(count, sb->cptr, sizeof(long long) * (nrows + 1), hipMemcpyDeviceToDevice,
                          s));
  for (int m = 0; m < nmaps; ++m) {
    long long work = nelems[m] * rdims[m];
    if (work == 0) continue;
    hipLaunchKernelGGL(scatter_candidates, dim3((work + 255) / 256), dim3(256), 0, s, rowmaps[m],
                       rdims[m], colmaps[m], cdims[m], nelems[m],
                       reinterpret_cast<unsigned long long*>(count), cand);
  }
  SP_CHECK(hipMalloc(&ucount, sizeof(long long) * (nrows + 1)));
  SP_CHECK(hipMalloc(&first, ncand > 0 ? ncand : 1));
  if (nrows > 0) {
    long long threads = static_cast<long long>(nrows) * 64;
    hipLaunchKernelGGL(flag_first, dim3((threads + 255) / 256), dim3(256), 0, s, sb->cptr,
                       static_cast<long long>(nrows), cand, first);
    hipLaunchKernelGGL(rank_unique, dim3((threads + 255) / 256), dim3(256), 0, s, sb->cptr,
                       static_cast<long long>(nrows), cand, first, sb->sorted, ucount);
  }
  SP_CHECK(hipMalloc(&sb->nptr, sizeof(int) * (nrows + 1)));
  rc = exclusive_scan<int>(ucount, nrows, sb->nptr, sums, total, s);
  if (rc) return fail(rc);
  long long unode = 0;
  SP_CHECK(hipMemcpyAsync(&unode, total, sizeof(long long), hipMemcpyDeviceToHost, s));
  SP_CHECK(hipStreamSynchronize(s));
  int un = static_cast<int>(unode);
  SP_CHECK(hipMemcpyAsync(sb->nptr + nrows, &un, sizeof(int), hipMemcpyHostToDevice, s));
  sb->nnz = unode * rmult * cmult;
  if (sb->nnz > 0x7fffffffLL) {
    op2hip::set_error("sparsity has %lld nonzeros, beyond int32 CSR indices", sb->nnz);
    return fail(OP2HIP_ERR_ARG);
  }
  long long lsize = static_cast<long long>(nrows) * rmult;
  hipLaunchKernelGGL(expand_rowptr, dim3((lsize + 1 + 255) / 256), dim3(256), 0, s, sb->nptr,
                     static_cast<long long>(nrows), rmult, cmult, rowptr);
  SP_CHECK(hipGetLastError());
  SP_CHECK(hipStreamSynchronize(s));
  hipFree(count);
  hipFree(sums);
  hipFree(total);
  hipFree(ucount);
  hipFree(cand);
  hipFree(first);
#undef SP_CHECK
  *nnz_out = sb->nnz;
  *handle = sb;
  return OP2HIP_OK;
}

// Phase 2: write colidx (device, nnz) and, when non-null, the per-row counts
// nnz_rows (device, nrows*rmult), then release the handle.
int op2hip_sparsity_fill(void* handle, const int* rowptr, int* colidx, int* nnz_rows,
                         void* stream) {
  SparsityBuild* sb = static_cast<SparsityBuild*>(handle);
  OP2HIP_REQUIRE(sb && rowptr && colidx, "null sparsity fill argument");
  hipStream_t s = static_cast<hipStream_t>(stream);
  long long lsize = sb->nrows * sb->rmult;
  if (lsize > 0) {
    hipLaunchKernelGGL(expand_colidx, dim3((lsize + 255) / 256), dim3(256), 0, s, sb->nptr,
                       sb->cptr, sb->sorted, sb->nrows, sb->rmult, sb->cmult, rowptr, colidx,
                       nnz_rows);
  }
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  free_build(sb);
  if (e != hipSuccess) {
    op2hip::set_error("sparsity fill failed: %s", hipGetErrorString(e));
    return OP2HIP_ERR_HIP;
  }
  return OP2HIP_OK;
}

int op2hip_sparsity_discard(void* handle) {
  free_build(static_cast<SparsityBuild*>(handle));
  return OP2HIP_OK;
}

}  // extern "C"
