// Device construction of a Sparsity's CSR structure.
//
// Replaces build_sparsity_pattern_seq (pyop2/sparsity_utils.cxx:5-51), which
// inserts, for every map pair, element e, row-map column i, r < rmult,
// col-map column d and c < cmult, the column cmult*cmap[e][d]+c into the
// std::set of row rmult*rmap[e][i]+r, then emits rowptr and the sorted
// unique colidx.  The result is a pure function of the maps, so the device
// build below is bit-exact to it.
//
// The node-level pattern is found from the transpose of the row maps:
//
//   1. count the (element, row-map column) incidences of every row (atomic
//      adds), scan them into per-row offsets, scatter the element ids into
//      per-row incidence lists (their order inside a row does not matter);
//   2. per row, gather the column-map entries of the row's elements (the
//      row's candidates), sort them and count the distinct values:
//        - rows of at most 256 candidates: one wavefront, the candidates in
//          registers (1, 2 or 4 per lane) sorted by a bitonic network whose
//          exchanges between lanes are shuffles;
//        - longer rows: one workgroup, the distinct values collected in an
//          LDS hash table, compacted and sorted by a bitonic network in LDS;
//          rows with more distinct values than the table holds fall back to
//          repeated minimum searches (one pass per distinct value);
//   3. scan the distinct counts into the node-level rowptr (the one host
//      synchronisation: the caller needs nnz to allocate colidx), expand to
//      rmult x cmult blocks: row rmult*n+r holds {cmult*m+c : m in cols(n),
//      c < cmult} in order;
//   4. the fill pass expands each row's sorted distinct columns (kept by step
//      2 at the row's first candidate slot) into colidx; long rows repeat
//      step 2.
//
// Work is linear in the candidates (times log^2 of a row's length) and the
// scratch is the caller's workspace (incidence lists, a slot per candidate
// and per-row arrays: about 4 * (4 * nrows + incidences * (1 + widest
// column map)) bytes), so no build allocates device memory.
#include <vector>

#include "common.h"
#include "scan.h"

namespace {

using op2hip_scan::exclusive_scan;
using op2hip_scan::kScanTile;

constexpr int kMaxPairs = 8;
constexpr int kPad = 0x7fffffff;     // sorts after every column
constexpr int kWaveSlots = 256;      // candidates a wavefront row holds (4 per lane)
constexpr int kLongBlock = 256;      // threads of the long-row kernel
constexpr int kHash = 8192;          // LDS hash slots of the long-row kernel
constexpr int kHashMax = 6144;       // distinct values it takes before falling back
constexpr int kLongGrid = 1024;      // workgroups of the long-row kernel (grid-stride)

// The map pairs of a build: column maps, their arities, and each pair's
// first element in the concatenated element numbering of the incidence
// lists.
struct Pairs {
  int n;
  int cmax;  // widest column map; a row's candidate slots are incidences x cmax
  const int* cmap[kMaxPairs];
  int cdim[kMaxPairs];
  int eoff[kMaxPairs + 1];
};

__global__ void fill_zero(int* __restrict__ p, long long n) {
  long long k = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k < n) p[k] = 0;
}

__global__ void count_incidences(const int* __restrict__ rmap, long long nk, int nrows,
                                 int* __restrict__ cnt) {
  long long k = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= nk) return;
  int row = rmap[k];
  if (static_cast<unsigned>(row) < static_cast<unsigned>(nrows)) atomicAdd(cnt + row, 1);
}

__global__ void scatter_incidences(const int* __restrict__ rmap, long long nk, int rdim,
                                   int eoff, int nrows, int* __restrict__ cursor,
                                   int* __restrict__ inc) {
  long long k = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= nk) return;
  int row = rmap[k];
  if (static_cast<unsigned>(row) >= static_cast<unsigned>(nrows)) return;
  int pos = atomicAdd(cursor + row, 1);
  inc[pos] = eoff + static_cast<int>(k / rdim);
}

// ---- per-row distinct columns --------------------------------------------

// Candidate `slot` of a row whose incidence list is inc[ib ...]: incidence
// slot / cmax, column-map column slot % cmax (kPad beyond that map's arity).
__device__ inline int candidate(const Pairs& P, const int* __restrict__ inc, int ib, int slot) {
  int t = slot / P.cmax;
  int d = slot - t * P.cmax;
  int g = inc[ib + t];
  int m = 0;
  while (m + 1 < P.n && g >= P.eoff[m + 1]) ++m;
  if (d >= P.cdim[m]) return kPad;
  return P.cmap[m][static_cast<long long>(g - P.eoff[m]) * P.cdim[m] + d];
}

// Ascending bitonic sort of the 64*R values v[r] (element r*64 + lane).
template <int R>
__device__ inline void wave_sort(int (&v)[R], int lane) {
  constexpr int N = 64 * R;
#pragma unroll
  for (int k = 2; k <= N; k <<= 1) {
#pragma unroll
    for (int j = k >> 1; j > 0; j >>= 1) {
      if (j >= 64) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
          int pr = r ^ (j >> 6);
          if (pr > r) {
            bool asc = (((r << 6) | lane) & k) == 0;
            int a = v[r], b = v[pr];
            int lo = a < b ? a : b, hi = a < b ? b : a;
            v[r] = asc ? lo : hi;
            v[pr] = asc ? hi : lo;
          }
        }
      } else {
#pragma unroll
        for (int r = 0; r < R; ++r) {
          int w = __shfl_xor(v[r], j, 64);
          bool lower = (lane & j) == 0;
          bool asc = (((r << 6) | lane) & k) == 0;
          int lo = v[r] < w ? v[r] : w, hi = v[r] < w ? w : v[r];
          v[r] = (lower == asc) ? lo : hi;
        }
      }
    }
  }
}

struct RowOut {
  const int* nptr;   // node-level row offsets (fill pass)
  int rmult, cmult;
  int* colidx;       // fill pass: the expanded CSR columns
  int* nnz_rows;     // fill pass, may be null
  int* ucnt;         // count pass: distinct values per row
  int* err;          // set when a fill pass finds another count than the build
  int* uniq;         // each wavefront row's sorted distinct values, at its first
                     // candidate slot (written by the count pass, read by the fill)
};

// Write the sorted distinct values of node row n (rank-th value v) or count
// them; `u` is the row's distinct count.
__device__ inline void put_value(const RowOut& o, int n, int u, int rank, int v) {
  long long base = static_cast<long long>(o.nptr[n]) * o.rmult * o.cmult;
  for (int rr = 0; rr < o.rmult; ++rr) {
    int* dst = o.colidx + base + static_cast<long long>(rr) * u * o.cmult +
               static_cast<long long>(rank) * o.cmult;
    for (int c = 0; c < o.cmult; ++c) dst[c] = o.cmult * v + c;
  }
}

__device__ inline void put_count(const RowOut& o, int n, int u) {
  for (int rr = 0; rr < o.rmult; ++rr) o.nnz_rows[n * o.rmult + rr] = u * o.cmult;
}

template <int R>
__device__ inline void wave_row(const Pairs& P, const int* __restrict__ inc, int ib, int slots,
                                int n, int lane, const RowOut& o) {
  int v[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    int s = (r << 6) | lane;
    v[r] = s < slots ? candidate(P, inc, ib, s) : kPad;
  }
  wave_sort<R>(v, lane);
  // first occurrences: differ from the element before (lane - 1, or lane 63
  // of the previous register)
  unsigned long long bal[R];
  int u = 0;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    int up = __shfl_up(v[r], 1, 64);
    int prev_last = r > 0 ? __shfl(v[r > 0 ? r - 1 : 0], 63, 64) : -1;
    int prev = lane > 0 ? up : prev_last;
    bool first = v[r] != kPad && v[r] != prev;
    bal[r] = __ballot(first);
    u += __popcll(bal[r]);
  }
  if (lane == 0) o.ucnt[n] = u;
  int* dst = o.uniq + static_cast<long long>(ib) * P.cmax;
  int before = 0;
  const unsigned long long lt = __lanemask_lt();
#pragma unroll
  for (int r = 0; r < R; ++r) {
    if ((bal[r] >> lane) & 1ull) dst[before + __popcll(bal[r] & lt)] = v[r];
    before += __popcll(bal[r]);
  }
}

// One wavefront per row; rows of more than kWaveSlots candidates are listed
// for the long-row kernel.
template <bool FILL>
__global__ void __launch_bounds__(256) rows_wave(const int* __restrict__ iptr,
                                                 const int* __restrict__ inc, Pairs P, int nrows,
                                                 RowOut o, int* __restrict__ longlist,
                                                 int* __restrict__ nlong) {
  const int lane = threadIdx.x & 63;
  long long n = (static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
  if (n >= nrows) return;
  const int ib = iptr[n];
  const long long slots = static_cast<long long>(iptr[n + 1] - ib) * P.cmax;
  if (slots > kWaveSlots) {
    if (!FILL && lane == 0) longlist[atomicAdd(nlong, 1)] = static_cast<int>(n);
    return;
  }
  const int sl = static_cast<int>(slots);
  const int row = static_cast<int>(n);
  if (FILL) {
    // the build's sorted distinct values, expanded into colidx
    const int u = o.nptr[n + 1] - o.nptr[n];
    if (u > sl) {
      if (lane == 0) atomicOr(o.err, 1);
      return;
    }
    const int* src = o.uniq + static_cast<long long>(ib) * P.cmax;
    for (int q = lane; q < u; q += 64) put_value(o, row, u, q, src[q]);
    if (o.nnz_rows && lane == 0) put_count(o, row, u);
    return;
  }
  if (sl <= 64)
    wave_row<1>(P, inc, ib, sl, row, lane, o);
  else if (sl <= 128)
    wave_row<2>(P, inc, ib, sl, row, lane, o);
  else
    wave_row<4>(P, inc, ib, sl, row, lane, o);
}

__device__ inline unsigned hash_slot(int v) {
  return (static_cast<unsigned>(v) * 2654435761u) >> (32 - 13);  // kHash = 2^13
}

// One workgroup per listed row (grid-stride over the list).
template <bool FILL>
__global__ void __launch_bounds__(kLongBlock) rows_long(const int* __restrict__ iptr,
                                                        const int* __restrict__ inc, Pairs P,
                                                        RowOut o,
                                                        const int* __restrict__ longlist,
                                                        const int* __restrict__ nlong) {
  __shared__ int tab[kHash];
  __shared__ int s_cnt, s_over, s_min;
  __shared__ int s_part[kLongBlock / 64];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int total_long = *nlong;
  for (int li = blockIdx.x; li < total_long; li += gridDim.x) {
    const int n = longlist[li];
    const int ib = iptr[n];
    const long long slots = static_cast<long long>(iptr[n + 1] - ib) * P.cmax;
    for (int k = tid; k < kHash; k += kLongBlock) tab[k] = -1;
    if (tid == 0) {
      s_cnt = 0;
      s_over = 0;
    }
    __syncthreads();
    for (long long s = tid; s < slots; s += kLongBlock) {
      if (s_over) break;
      int v = candidate(P, inc, ib, static_cast<int>(s));
      if (v == kPad) continue;
      unsigned h = hash_slot(v);
      for (int probe = 0; probe < kHash; ++probe) {
        int prev = atomicCAS(&tab[h], -1, v);
        if (prev == -1) {
          if (atomicAdd(&s_cnt, 1) + 1 > kHashMax) s_over = 1;
          break;
        }
        if (prev == v) break;
        h = (h + 1) & (kHash - 1);
      }
    }
    __syncthreads();
    const bool over = s_over != 0;
    int u = s_cnt;
    if (!over) {
      if (!FILL) {
        if (tid == 0) o.ucnt[n] = u;
      } else if (o.nptr[n + 1] - o.nptr[n] != u) {
        if (tid == 0) atomicOr(o.err, 1);
      } else {
        // compact the table (each thread its run of kHash / kLongBlock
        // slots, read before any write), then sort the u values
        constexpr int kRun = kHash / kLongBlock;
        int mine[kRun];
        int c = 0;
        for (int k = 0; k < kRun; ++k) {
          mine[k] = tab[tid * kRun + k];
          c += mine[k] != -1;
        }
        int x = c;
        for (int off = 1; off < 64; off <<= 1) {
          int y = __shfl_up(x, off, 64);
          if (lane >= off) x += y;
        }
        if (lane == 63) s_part[w] = x;
        __syncthreads();
        int pos = x - c;
        for (int k = 0; k < w; ++k) pos += s_part[k];
        for (int k = 0; k < kRun; ++k)
          if (mine[k] != -1) tab[pos++] = mine[k];
        int p2 = 1;
        while (p2 < u) p2 <<= 1;
        for (int k = u + tid; k < p2; k += kLongBlock) tab[k] = kPad;
        __syncthreads();
        for (int kk = 2; kk <= p2; kk <<= 1) {
          for (int j = kk >> 1; j > 0; j >>= 1) {
            for (int i = tid; i < p2; i += kLongBlock) {
              int l = i ^ j;
              if (l > i) {
                int a = tab[i], b = tab[l];
                bool asc = (i & kk) == 0;
                if ((a > b) == asc) {
                  tab[i] = b;
                  tab[l] = a;
                }
              }
            }
            __syncthreads();
          }
        }
        for (int k = tid; k < u; k += kLongBlock) put_value(o, n, u, k, tab[k]);
        if (o.nnz_rows && tid == 0) put_count(o, n, u);
      }
    } else {
      // more distinct values than the table holds: emit them in ascending
      // order, one minimum search over the candidates per value
      int last = -1, rank = 0;
      const int expect = FILL ? o.nptr[n + 1] - o.nptr[n] : 0;
      bool bad = false;
      for (;;) {
        int m = kPad;
        for (long long s = tid; s < slots; s += kLongBlock) {
          int v = candidate(P, inc, ib, static_cast<int>(s));
          if (v > last && v < m) m = v;
        }
        for (int off = 32; off > 0; off >>= 1) {
          int y = __shfl_xor(m, off, 64);
          m = y < m ? y : m;
        }
        __syncthreads();
        if (tid == 0) s_min = kPad;
        __syncthreads();
        if (lane == 0) atomicMin(&s_min, m);
        __syncthreads();
        m = s_min;
        if (m == kPad) break;
        if (FILL) {
          if (rank >= expect) {
            bad = true;
            break;
          }
          if (tid == 0) put_value(o, n, expect, rank, m);
        }
        ++rank;
        last = m;
      }
      if (!FILL) {
        if (tid == 0) o.ucnt[n] = rank;
      } else if (bad || rank != expect) {
        if (tid == 0) atomicOr(o.err, 1);
      } else if (o.nnz_rows && tid == 0) {
        put_count(o, n, expect);
      }
    }
    __syncthreads();
  }
}

// Expand the node-level pattern's row offsets into rmult x cmult blocks.
__global__ void expand_rowptr(const int* __restrict__ nptr, int nrows, int rmult, int cmult,
                              int* __restrict__ rowptr) {
  long long R = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  long long lsize = static_cast<long long>(nrows) * rmult;
  if (R > lsize) return;
  if (R == lsize) {
    rowptr[R] = static_cast<int>(static_cast<long long>(nptr[nrows]) * rmult * cmult);
    return;
  }
  long long n = R / rmult;
  int r = static_cast<int>(R % rmult);
  long long u = nptr[n + 1] - nptr[n];
  rowptr[R] = static_cast<int>(static_cast<long long>(nptr[n]) * rmult * cmult + r * u * cmult);
}

// Workspace layout (offsets in bytes, each segment 256-byte aligned).
struct Work {
  size_t cnt, iptr, inc, uniq, ucnt, nptr, longlist, nlong, err, sums, total, bytes;
};

size_t align256(size_t x) { return (x + 255) & ~static_cast<size_t>(255); }

Work work_layout(long long nrows, long long ninc, int cmax) {
  Work w;
  size_t o = 0;
  auto take = [&](size_t b) {
    size_t at = o;
    o += align256(b > 0 ? b : 1);
    return at;
  };
  size_t rows1 = sizeof(int) * static_cast<size_t>(nrows + 1);
  long long nb = (nrows + 1 + kScanTile - 1) / kScanTile + 1;
  w.cnt = take(rows1);
  w.iptr = take(rows1);
  w.inc = take(sizeof(int) * static_cast<size_t>(ninc));
  w.uniq = take(sizeof(int) * static_cast<size_t>(ninc) * cmax);
  w.ucnt = take(rows1);
  w.nptr = take(rows1);
  w.longlist = take(sizeof(int) * static_cast<size_t>(nrows + 1));
  w.nlong = take(sizeof(int));
  w.err = take(sizeof(int));
  w.sums = take(sizeof(long long) * static_cast<size_t>(nb));
  w.total = take(sizeof(long long) * 2);
  w.bytes = o;
  return w;
}

struct SparsityBuild {
  int nrows = 0;
  int rmult = 1, cmult = 1;
  Pairs pairs{};
  char* work = nullptr;
  Work w{};
  long long nnz = 0;
};

template <typename T>
T* at(char* base, size_t off) {
  return reinterpret_cast<T*>(base + off);
}

long long incidences(int nmaps, const long long* nelems, const int* rdims) {
  long long ninc = 0;
  for (int m = 0; m < nmaps; ++m) ninc += nelems[m] * rdims[m];
  return ninc;
}

}  // namespace

extern "C" {

long long op2hip_sparsity_workspace(int nrows, int nmaps, const long long* nelems,
                                    const int* rdims, const int* cdims) {
  if (nrows < 0 || nmaps <= 0 || nmaps > kMaxPairs || !nelems || !rdims || !cdims) return -1;
  int cmax = 1;
  for (int m = 0; m < nmaps; ++m) {
    if (nelems[m] < 0 || rdims[m] <= 0 || cdims[m] <= 0) return -1;
    cmax = cdims[m] > cmax ? cdims[m] : cmax;
  }
  return static_cast<long long>(
      work_layout(nrows, incidences(nmaps, nelems, rdims), cmax).bytes);
}

// Phase 1: build the pattern and write rowptr (device, nrows*rmult + 1).
// rowmaps/colmaps are host arrays of device pointers to int32 map values;
// nelems[m] is the iteration-set size of pair m.  work is a device buffer of
// op2hip_sparsity_workspace() bytes that must stay untouched until the
// handle is filled or discarded.  *nnz_out receives the number of nonzeros.
int op2hip_sparsity_build(int rmult, int cmult, int nrows, int nmaps, const int* const* rowmaps,
                          const int* rdims, const int* const* colmaps, const int* cdims,
                          const long long* nelems, void* work, long long work_bytes,
                          int* rowptr, void** handle, long long* nnz_out, void* stream) {
  OP2HIP_REQUIRE(rmult > 0 && cmult > 0 && nrows >= 0 && nmaps > 0, "bad sparsity arguments");
  OP2HIP_REQUIRE(nmaps <= kMaxPairs, "a sparsity of at most %d map pairs is supported",
                 kMaxPairs);
  OP2HIP_REQUIRE(rowptr && handle && nnz_out && work, "null sparsity argument");
  long long ninc = 0;
  int cmax = 1;
  for (int m = 0; m < nmaps; ++m) {
    OP2HIP_REQUIRE(nelems[m] >= 0 && rdims[m] > 0 && cdims[m] > 0, "bad map pair %d", m);
    ninc += nelems[m] * rdims[m];
    cmax = cdims[m] > cmax ? cdims[m] : cmax;
  }
  long long etot = 0;
  for (int m = 0; m < nmaps; ++m) etot += nelems[m];
  OP2HIP_REQUIRE(ninc <= 0x7fffffffLL && etot <= 0x7fffffffLL,
                 "sparsity of %lld incidences, beyond int32 offsets", ninc);
  Work w = work_layout(nrows, ninc, cmax);
  OP2HIP_REQUIRE(work_bytes >= static_cast<long long>(w.bytes),
                 "sparsity workspace of %lld bytes, %zu needed", work_bytes, w.bytes);
  hipStream_t s = static_cast<hipStream_t>(stream);
  char* base = static_cast<char*>(work);
  SparsityBuild* sb = new SparsityBuild();
  sb->nrows = nrows;
  sb->rmult = rmult;
  sb->cmult = cmult;
  sb->work = base;
  sb->w = w;
  Pairs& P = sb->pairs;
  P.n = nmaps;
  P.cmax = cmax;
  P.eoff[0] = 0;
  for (int m = 0; m < nmaps; ++m) {
    P.cmap[m] = colmaps[m];
    P.cdim[m] = cdims[m];
    P.eoff[m + 1] = P.eoff[m] + static_cast<int>(nelems[m]);
  }
  int* cnt = at<int>(base, w.cnt);
  int* iptr = at<int>(base, w.iptr);
  int* inc = at<int>(base, w.inc);
  int* ucnt = at<int>(base, w.ucnt);
  int* nptr = at<int>(base, w.nptr);
  int* longlist = at<int>(base, w.longlist);
  int* nlong = at<int>(base, w.nlong);
  int* err = at<int>(base, w.err);
  long long* sums = at<long long>(base, w.sums);
  long long* total = at<long long>(base, w.total);

  // counters start at zero (fill kernels, so the sequence can be captured)
  hipLaunchKernelGGL(fill_zero, dim3((nrows + 1 + 255) / 256), dim3(256), 0, s, cnt,
                     static_cast<long long>(nrows) + 1);
  hipLaunchKernelGGL(fill_zero, dim3(1), dim3(64), 0, s, nlong, 1LL);
  hipLaunchKernelGGL(fill_zero, dim3(1), dim3(64), 0, s, err, 1LL);
  for (int m = 0; m < nmaps; ++m) {
    long long nk = nelems[m] * rdims[m];
    if (nk == 0) continue;
    hipLaunchKernelGGL(count_incidences, dim3((nk + 255) / 256), dim3(256), 0, s, rowmaps[m], nk,
                       nrows, cnt);
  }
  // row offsets into iptr; cnt becomes the scatter cursor
  exclusive_scan(cnt, nrows, iptr, cnt, sums, total, s);
  for (int m = 0; m < nmaps; ++m) {
    long long nk = nelems[m] * rdims[m];
    if (nk == 0) continue;
    hipLaunchKernelGGL(scatter_incidences, dim3((nk + 255) / 256), dim3(256), 0, s, rowmaps[m],
                       nk, rdims[m], P.eoff[m], nrows, cnt, inc);
  }
  RowOut o{nptr, rmult, cmult, nullptr, nullptr, ucnt, err, at<int>(base, w.uniq)};
  if (nrows > 0) {
    long long threads = static_cast<long long>(nrows) * 64;
    hipLaunchKernelGGL(rows_wave<false>, dim3((threads + 255) / 256), dim3(256), 0, s, iptr, inc,
                       P, nrows, o, longlist, nlong);
    hipLaunchKernelGGL(rows_long<false>, dim3(kLongGrid), dim3(kLongBlock), 0, s, iptr, inc, P,
                       o, longlist, nlong);
  }
  exclusive_scan(ucnt, nrows, nptr, nullptr, sums, total, s);
  hipError_t e = hipGetLastError();
  long long unode = 0;
  if (e == hipSuccess) e = hipMemcpyAsync(&unode, total, sizeof(long long),
                                          hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) {
    delete sb;
    op2hip::set_error("sparsity build failed: %s", hipGetErrorString(e));
    return OP2HIP_ERR_HIP;
  }
  sb->nnz = unode * rmult * cmult;
  if (sb->nnz > 0x7fffffffLL) {
    op2hip::set_error("sparsity has %lld nonzeros, beyond int32 CSR indices", sb->nnz);
    delete sb;
    return OP2HIP_ERR_ARG;
  }
  long long lsize = static_cast<long long>(nrows) * rmult;
  hipLaunchKernelGGL(expand_rowptr, dim3((lsize + 1 + 255) / 256), dim3(256), 0, s, nptr, nrows,
                     rmult, cmult, rowptr);
  e = hipGetLastError();
  if (e != hipSuccess) {
    delete sb;
    op2hip::set_error("sparsity rowptr failed: %s", hipGetErrorString(e));
    return OP2HIP_ERR_HIP;
  }
  *nnz_out = sb->nnz;
  *handle = sb;
  return OP2HIP_OK;
}

// Phase 2: write colidx (device, nnz) and, when non-null, the per-row counts
// nnz_rows (device, nrows*rmult), then release the handle.  A row whose
// recount differs from the build's (the maps changed in between) is not
// written and makes the call fail; the call waits for the stream to report
// that.
int op2hip_sparsity_fill(void* handle, const int* rowptr, int* colidx, int* nnz_rows,
                         void* stream) {
  SparsityBuild* sb = static_cast<SparsityBuild*>(handle);
  OP2HIP_REQUIRE(sb && rowptr && colidx, "null sparsity fill argument");
  hipStream_t s = static_cast<hipStream_t>(stream);
  char* base = sb->work;
  const Work& w = sb->w;
  int* err = at<int>(base, w.err);
  RowOut o{at<int>(base, w.nptr), sb->rmult, sb->cmult, colidx, nnz_rows, nullptr, err,
           at<int>(base, w.uniq)};
  if (sb->nrows > 0) {
    long long threads = static_cast<long long>(sb->nrows) * 64;
    hipLaunchKernelGGL(rows_wave<true>, dim3((threads + 255) / 256), dim3(256), 0, s,
                       at<int>(base, w.iptr), at<int>(base, w.inc), sb->pairs, sb->nrows, o,
                       at<int>(base, w.longlist), at<int>(base, w.nlong));
    hipLaunchKernelGGL(rows_long<true>, dim3(kLongGrid), dim3(kLongBlock), 0, s,
                       at<int>(base, w.iptr), at<int>(base, w.inc), sb->pairs, o,
                       at<int>(base, w.longlist), at<int>(base, w.nlong));
  }
  int herr = 0;
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(&herr, err, sizeof(int), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  delete sb;
  if (e != hipSuccess) {
    op2hip::set_error("sparsity fill failed: %s", hipGetErrorString(e));
    return OP2HIP_ERR_HIP;
  }
  if (herr) {
    op2hip::set_error("sparsity fill: the maps changed between build and fill");
    return OP2HIP_ERR_ARG;
  }
  return OP2HIP_OK;
}

int op2hip_sparsity_discard(void* handle) {
  delete static_cast<SparsityBuild*>(handle);
  return OP2HIP_OK;
}

}  // extern "C"
