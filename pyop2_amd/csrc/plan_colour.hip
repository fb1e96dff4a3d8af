// Device construction of coloured plans.
//
// Replaces pyop2/plan.pyx:68-456 (partition, block and thread colouring)
// for the coloured loop template, on the device.  Two forms, as the host
// build (op2hip_plan_build) offers them:
//
//   mode 0  two-level colouring: blocks are the contiguous runs of
//           part_size elements of the loop's element order (plan.pyx:107-
//           110; under the device layout a run is a compact patch of the
//           mesh); blocks sharing a race target get different colours, and
//           inside a block elements sharing a target get different thread
//           colours, given greedily in element order as plan.pyx does;
//   mode 2  element colouring: elements sharing a race target get different
//           colours; each colour's elements, in ascending order, are cut
//           into blocks of part_size (one thread colour per block).
//
// Block and element colours come from a Jones-Plassmann-Luby iteration:
// an uncoloured item takes the smallest colour none of its coloured
// neighbours has once no uncoloured neighbour has a higher priority (a
// fixed hash of its index), so the colouring depends on the inputs only.
// Items update their colour in place: a neighbour that sees the new colour
// accounts for it, one that does not still waits, so one kernel per round
// suffices.  Neighbours are found through the race targets' incidence
// lists (count, scan, scatter, as for the sparsity).  Arrays, as the host
// build's (checked by op2hip_plan_verify):
//
//   elems[n]        execution order (blocks are consecutive runs of it)
//   offset/nelem[b] the run of block b inside elems
//   thrcol[n]       thread colour of elems[i]; nthrcol[b] colours in b
//   blkcol[b]       block colour; blocks of one colour share no target
//   blkmap[nblocks] block ids sorted (stably) by colour; ncolblk[c]
#include <vector>

#include "common.h"
#include "scan.h"

namespace {

using op2hip_scan::exclusive_scan;
using op2hip_scan::scan_tiles_for;

constexpr int kMaxCols = 16;
constexpr int kMaxColours = 256;  // colours a JPL item may take
constexpr int kThreads = 256;
constexpr int kTile = 256;        // elements per tile of the colour sort
constexpr int kRoundsPerPoll = 8;
constexpr int kMaxRounds = 4096;

struct Cols {
  int n;
  const int* v[kMaxCols];
  int dim[kMaxCols];
  int col[kMaxCols];
  int off[kMaxCols];  // the column's space offset in the combined target numbering
};

__device__ inline int target_of(const Cols& C, int e, int ac) {
  return C.off[ac] + C.v[ac][static_cast<long long>(e) * C.dim[ac] + C.col[ac]];
}

__device__ inline unsigned prio(unsigned x) {
  // a bijective integer mix: distinct items get distinct priorities
  x ^= x >> 16;
  x *= 0x7feb352dU;
  x ^= x >> 15;
  x *= 0x846ca68bU;
  x ^= x >> 16;
  return x;
}

__global__ void zero_ints(int* __restrict__ p, long long n) {
  long long k = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k < n) p[k] = 0;
}

__global__ void fill_ints(int* __restrict__ p, long long n, int v) {
  long long k = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k < n) p[k] = v;
}

__global__ void count_inc(Cols C, int n, int* __restrict__ cnt) {
  long long k = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= static_cast<long long>(n) * C.n) return;
  const int e = static_cast<int>(k / C.n), ac = static_cast<int>(k - static_cast<long long>(e) * C.n);
  atomicAdd(cnt + target_of(C, e, ac), 1);
}

__global__ void scatter_inc(Cols C, int n, int* __restrict__ cursor, int* __restrict__ inc) {
  long long k = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= static_cast<long long>(n) * C.n) return;
  const int e = static_cast<int>(k / C.n), ac = static_cast<int>(k - static_cast<long long>(e) * C.n);
  inc[atomicAdd(cursor + target_of(C, e, ac), 1)] = e;
}

// A 256-bit set of colours held by one lane.
struct ColourSet {
  unsigned long long w[kMaxColours / 64];
  __device__ void clear() {
    for (int i = 0; i < kMaxColours / 64; ++i) w[i] = 0;
  }
  __device__ void add(int c) {
    if (c >= 0 && c < kMaxColours) w[c >> 6] |= 1ull << (c & 63);
  }
  // smallest colour not in the set (kMaxColours: none free)
  __device__ int first_free() const {
    for (int i = 0; i < kMaxColours / 64; ++i)
      if (~w[i]) return i * 64 + __builtin_ctzll(~w[i]);
    return kMaxColours;
  }
};

// One JPL round over elements (mode 2): a thread per uncoloured element.
__global__ void jpl_elements(Cols C, int n, const int* __restrict__ tptr,
                             const int* __restrict__ inc, int* colour, int* __restrict__ left,
                             int* __restrict__ err) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  if (__atomic_load_n(colour + e, __ATOMIC_RELAXED) >= 0) return;
  const unsigned pe = prio(static_cast<unsigned>(e));
  ColourSet used;
  used.clear();
  bool ready = true;
  for (int ac = 0; ac < C.n && ready; ++ac) {
    const int t = target_of(C, e, ac);
    for (int q = tptr[t]; q < tptr[t + 1]; ++q) {
      const int o = inc[q];
      if (o == e) continue;
      const int co = __atomic_load_n(colour + o, __ATOMIC_RELAXED);
      if (co >= 0) {
        used.add(co);
      } else if (prio(static_cast<unsigned>(o)) > pe) {
        ready = false;
        break;
      }
    }
  }
  if (!ready) {
    atomicAdd(left, 1);
    return;
  }
  const int c = used.first_free();
  if (c >= kMaxColours) {
    atomicOr(err, 1);
    return;
  }
  __atomic_store_n(colour + e, c, __ATOMIC_RELAXED);
}

// One JPL round over blocks (mode 0): a workgroup per uncoloured block of
// part_size consecutive elements; neighbours are the blocks of the
// elements sharing a target with one of its elements.
__global__ void __launch_bounds__(kThreads) jpl_blocks(Cols C, int n, int ps, int nblocks,
                                                       const int* __restrict__ tptr,
                                                       const int* __restrict__ inc,
                                                       int* blkcol, int* __restrict__ left,
                                                       int* __restrict__ err) {
  __shared__ unsigned long long s_used[kMaxColours / 64];
  __shared__ int s_wait, s_done;
  for (int b = blockIdx.x; b < nblocks; b += gridDim.x) {
    // one read of the block's state, shared, so the workgroup agrees
    if (threadIdx.x == 0) s_done = __atomic_load_n(blkcol + b, __ATOMIC_RELAXED) >= 0;
    if (threadIdx.x < kMaxColours / 64) s_used[threadIdx.x] = 0;
    if (threadIdx.x == 0) s_wait = 0;
    __syncthreads();
    if (s_done) {
      __syncthreads();
      continue;
    }
    const unsigned pb = prio(static_cast<unsigned>(b));
    const int e0 = b * ps, e1 = e0 + ps < n ? e0 + ps : n;
    for (int e = e0 + threadIdx.x; e < e1; e += kThreads) {
      for (int ac = 0; ac < C.n; ++ac) {
        const int t = target_of(C, e, ac);
        for (int q = tptr[t]; q < tptr[t + 1]; ++q) {
          const int o = inc[q] / ps;
          if (o == b) continue;
          const int co = __atomic_load_n(blkcol + o, __ATOMIC_RELAXED);
          if (co >= 0) {
            if (co < kMaxColours) atomicOr(&s_used[co >> 6], 1ull << (co & 63));
          } else if (prio(static_cast<unsigned>(o)) > pb) {
            s_wait = 1;
          }
        }
      }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      if (s_wait) {
        atomicAdd(left, 1);
      } else {
        int c = kMaxColours;
        for (int i = 0; i < kMaxColours / 64; ++i)
          if (~s_used[i]) {
            c = i * 64 + __builtin_ctzll(~s_used[i]);
            break;
          }
        if (c >= kMaxColours)
          atomicOr(err, 1);
        else
          __atomic_store_n(blkcol + b, c, __ATOMIC_RELAXED);
      }
    }
    __syncthreads();
  }
}

// Thread colours inside each block (mode 0): greedy in element order, as
// plan.pyx colours threads, by one thread per block over an LDS table of
// the block's targets (open addressing) with 64-bit colour masks in rounds
// of 64 colours.
constexpr int kTab = 4096;
__global__ void __launch_bounds__(64) thread_colours(Cols C, int n, int ps, int nblocks,
                                                     int* __restrict__ thrcol,
                                                     int* __restrict__ nthrcol,
                                                     int* __restrict__ err) {
  __shared__ int s_key[kTab];
  __shared__ unsigned long long s_mask[kTab];
  for (int b = blockIdx.x; b < nblocks; b += gridDim.x) {
    const int e0 = b * ps, e1 = e0 + ps < n ? e0 + ps : n;
    for (int i = e0 + threadIdx.x; i < e1; i += 64) thrcol[i] = -1;
    int remaining = e1 - e0, base = 0, ncol = 0;
    while (remaining > 0) {
      for (int k = threadIdx.x; k < kTab; k += 64) {
        s_key[k] = -1;
        s_mask[k] = 0;
      }
      __syncthreads();
      if (threadIdx.x == 0) {
        for (int e = e0; e < e1; ++e) {
          if (thrcol[e] >= 0) continue;
          unsigned long long m = 0;
          int slot[kMaxCols];
          bool full = false;
          for (int ac = 0; ac < C.n; ++ac) {
            const int t = target_of(C, e, ac);
            unsigned h = (static_cast<unsigned>(t) * 2654435761u) & (kTab - 1);
            int probe = 0;
            while (s_key[h] != -1 && s_key[h] != t && probe < kTab) {
              h = (h + 1) & (kTab - 1);
              ++probe;
            }
            if (probe >= kTab) {
              full = true;
              break;
            }
            slot[ac] = static_cast<int>(h);
            if (s_key[h] == t) m |= s_mask[h];
          }
          if (full) {
            atomicOr(err, 2);
            remaining = 0;
            break;
          }
          if (m == ~0ull) continue;
          const int c = __builtin_ctzll(~m);
          thrcol[e] = base + c;
          ncol = base + c + 1 > ncol ? base + c + 1 : ncol;
          --remaining;
          for (int ac = 0; ac < C.n; ++ac) {
            s_key[slot[ac]] = target_of(C, e, ac);
            s_mask[slot[ac]] |= 1ull << c;
          }
        }
      }
      __syncthreads();
      remaining = __shfl(remaining, 0, 64);
      base += 64;
      __syncthreads();
    }
    if (threadIdx.x == 0) nthrcol[b] = ncol;
    __syncthreads();
  }
}

__global__ void iota(int* __restrict__ p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = i;
}

// Largest colour + 1 of items[0, n) into *out (atomic max).
__global__ void max_colour(const int* __restrict__ c, int n, int* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  int v = i < n ? c[i] + 1 : 0;
  for (int off = 32; off > 0; off >>= 1) {
    int y = __shfl_xor(v, off, 64);
    v = y > v ? y : v;
  }
  if ((threadIdx.x & 63) == 0) atomicMax(out, v);
}

// Per tile of elements, the count of each colour: cnt[c * ntiles + tile].
__global__ void tile_counts(const int* __restrict__ colour, int n, int ncolours, int ntiles,
                            int* __restrict__ cnt) {
  __shared__ int s_c[kMaxColours];
  for (int k = threadIdx.x; k < ncolours; k += blockDim.x) s_c[k] = 0;
  __syncthreads();
  const int e = blockIdx.x * kTile + threadIdx.x;
  if (e < n) atomicAdd(&s_c[colour[e]], 1);
  __syncthreads();
  for (int k = threadIdx.x; k < ncolours; k += blockDim.x)
    cnt[static_cast<long long>(k) * ntiles + blockIdx.x] = s_c[k];
}

// Stable placement: each tile's elements, in ascending order, at their
// colour's scanned offset (one lane places them in order).
__global__ void __launch_bounds__(64) tile_place(const int* __restrict__ colour, int n,
                                                 int ncolours, int ntiles,
                                                 const int* __restrict__ off,
                                                 int* __restrict__ elems) {
  __shared__ int s_cur[kMaxColours];
  const int tile = blockIdx.x;
  for (int k = threadIdx.x; k < ncolours; k += 64)
    s_cur[k] = off[static_cast<long long>(k) * ntiles + tile];
  __syncthreads();
  if (threadIdx.x == 0) {
    const int e0 = tile * kTile, e1 = e0 + kTile < n ? e0 + kTile : n;
    for (int e = e0; e < e1; ++e) elems[s_cur[colour[e]]++] = e;
  }
}

size_t align256(size_t x) { return (x + 255) & ~static_cast<size_t>(255); }

struct Work {
  size_t tptr, inc, colour, left, err, ncol, sums, total, tcnt, toff, bytes;
};

Work work_layout(long long ntarget, long long ninc, long long n, long long nblocks) {
  Work w;
  size_t o = 0;
  auto take = [&](size_t b) {
    size_t at = o;
    o += align256(b > 0 ? b : 1);
    return at;
  };
  const size_t I = sizeof(int);
  const long long ntiles = (n + kTile - 1) / kTile;
  const long long items = n > nblocks ? n : nblocks;
  const long long tc = static_cast<long long>(kMaxColours) * (ntiles > 0 ? ntiles : 1);
  w.tptr = take(I * (ntarget + 1));
  w.inc = take(I * ninc);
  w.colour = take(I * items);
  w.left = take(I);
  w.err = take(I);
  w.ncol = take(I);
  long long ns = scan_tiles_for((ntarget + 1) > tc ? (ntarget + 1) : tc);
  w.sums = take(sizeof(long long) * ns);
  w.total = take(sizeof(long long) * 2);
  // target counts and cursors, then per-block thread colours or tile counts
  w.tcnt = take(I * ((tc > ntarget ? tc : ntarget) + 1));
  w.toff = take(I * (tc + 1));
  w.bytes = o;
  return w;
}

template <typename T>
T* at(char* base, size_t off) {
  return reinterpret_cast<T*>(base + off);
}

struct ColourPlan {
  int n = 0, ps = 1, mode = 0, nblocks = 0, ncolors = 0;
  char* work = nullptr;
  Work w{};
  Cols C{};
  std::vector<int> counts;  // mode 2: elements per colour
};

int build_cols(int nmaps, const int* const* maps, const int* dims, const int* idxs,
               const int* spaces, int nspaces, const int* space_size, Cols* C,
               long long* ntarget) {
  std::vector<long long> soff(nspaces + 1, 0);
  for (int s = 0; s < nspaces; ++s) soff[s + 1] = soff[s] + space_size[s];
  OP2HIP_REQUIRE(soff[nspaces] <= 0x7fffffffLL, "too many race targets");
  *C = Cols{};
  for (int m = 0; m < nmaps; ++m) {
    OP2HIP_REQUIRE(maps[m] && dims[m] > 0 && idxs[m] < dims[m] && spaces[m] >= 0 &&
                       spaces[m] < nspaces,
                   "bad race map %d", m);
    const int c0 = idxs[m] >= 0 ? idxs[m] : 0, c1 = idxs[m] >= 0 ? idxs[m] + 1 : dims[m];
    for (int c = c0; c < c1; ++c) {
      OP2HIP_REQUIRE(C->n < kMaxCols, "more than %d race columns", kMaxCols);
      C->v[C->n] = maps[m];
      C->dim[C->n] = dims[m];
      C->col[C->n] = c;
      C->off[C->n] = static_cast<int>(soff[spaces[m]]);
      ++C->n;
    }
  }
  *ntarget = soff[nspaces];
  return OP2HIP_OK;
}

}  // namespace

extern "C" {

long long op2hip_plan_workspace(int nelems, int part_size, int nac, const int* space_size,
                                int nspaces) {
  if (nelems < 0 || part_size <= 0 || nac <= 0 || nspaces <= 0 || !space_size) return -1;
  long long nt = 0;
  for (int s = 0; s < nspaces; ++s) nt += space_size[s];
  const long long nb = (static_cast<long long>(nelems) + part_size - 1) / part_size;
  return static_cast<long long>(
      work_layout(nt, static_cast<long long>(nelems) * nac, nelems, nb).bytes);
}

// Phase 1 of the device coloured plan: colours (blocks and threads for mode
// 0, elements for mode 2).  maps: host array of device int32 race map
// pointers (rows in the loop's element order) with the arguments of
// op2hip_plan_build; work: op2hip_plan_workspace() bytes kept until fill or
// discard.  Returns nblocks and ncolors (host) for the caller to size the
// arrays; waits for the stream once per kRoundsPerPoll colouring rounds.
int op2hip_plan_dbuild(int nelems, int part_size, int mode, int nmaps, const int* const* maps,
                       const int* dims, const int* idxs, const int* spaces, int nspaces,
                       const int* space_size, void* work, long long work_bytes,
                       int* thrcol, void** handle, int* nblocks_out, int* ncolors_out,
                       void* stream) {
  OP2HIP_REQUIRE(handle && nblocks_out && ncolors_out && work && nelems >= 0 && part_size > 0 &&
                     nmaps > 0 && (mode == 0 || mode == 2),
                 "bad device plan arguments");
  OP2HIP_REQUIRE(mode == 2 || nelems == 0 || thrcol, "two-level colouring needs thrcol");
  Cols C;
  long long nt = 0;
  int st = build_cols(nmaps, maps, dims, idxs, spaces, nspaces, space_size, &C, &nt);
  if (st != OP2HIP_OK) return st;
  const long long ninc = static_cast<long long>(nelems) * C.n;
  OP2HIP_REQUIRE(ninc <= 0x7fffffffLL, "plan of %lld incidences, beyond int32", ninc);
  const int nblocks = static_cast<int>((static_cast<long long>(nelems) + part_size - 1) / part_size);
  Work w = work_layout(nt, ninc, nelems, nblocks);
  OP2HIP_REQUIRE(work_bytes >= static_cast<long long>(w.bytes), "plan workspace too small");
  hipStream_t s = static_cast<hipStream_t>(stream);
  char* base = static_cast<char*>(work);
  int* tptr = at<int>(base, w.tptr);
  int* inc = at<int>(base, w.inc);
  int* colour = at<int>(base, w.colour);
  int* left = at<int>(base, w.left);
  int* err = at<int>(base, w.err);
  int* ncol = at<int>(base, w.ncol);
  long long* sums = at<long long>(base, w.sums);
  long long* total = at<long long>(base, w.total);
  int* cnt = at<int>(base, w.tcnt);  // target counts, then scan cursors
  auto blocks = [](long long k) { return dim3(static_cast<unsigned>((k + 255) / 256 > 0 ? (k + 255) / 256 : 1)); };

  hipLaunchKernelGGL(zero_ints, blocks(nt + 1), dim3(256), 0, s, cnt, nt + 1);
  hipLaunchKernelGGL(zero_ints, dim3(1), dim3(64), 0, s, err, 1LL);
  if (ninc > 0) hipLaunchKernelGGL(count_inc, blocks(ninc), dim3(256), 0, s, C, nelems, cnt);
  exclusive_scan(cnt, static_cast<int>(nt), tptr, cnt, sums, total, s);
  if (ninc > 0) hipLaunchKernelGGL(scatter_inc, blocks(ninc), dim3(256), 0, s, C, nelems, cnt, inc);
  const long long items = mode == 2 ? nelems : nblocks;
  hipLaunchKernelGGL(fill_ints, blocks(items), dim3(256), 0, s, colour, items, -1);
  // colouring rounds, polled every kRoundsPerPoll
  int rounds = 0;
  for (;;) {
    for (int r = 0; r < kRoundsPerPoll; ++r) {
      hipLaunchKernelGGL(zero_ints, dim3(1), dim3(64), 0, s, left, 1LL);
      if (mode == 2) {
        if (nelems > 0)
          hipLaunchKernelGGL(jpl_elements, blocks(nelems), dim3(256), 0, s, C, nelems, tptr, inc,
                             colour, left, err);
      } else if (nblocks > 0) {
        const int grid = nblocks < 16384 ? nblocks : 16384;
        hipLaunchKernelGGL(jpl_blocks, dim3(grid), dim3(kThreads), 0, s, C, nelems, part_size,
                           nblocks, tptr, inc, colour, left, err);
      }
    }
    rounds += kRoundsPerPoll;
    int hl[2] = {0, 0};
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(&hl[0], left, sizeof(int), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(&hl[1], err, sizeof(int), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
      op2hip::set_error("device plan colouring failed: %s", hipGetErrorString(e));
      return OP2HIP_ERR_HIP;
    }
    OP2HIP_REQUIRE(!hl[1], "device plan: more than %d colours needed", kMaxColours);
    if (hl[0] == 0) break;
    OP2HIP_REQUIRE(rounds < kMaxRounds, "device plan colouring did not finish");
  }
  if (mode == 0 && nblocks > 0) {
    const int grid = nblocks < 16384 ? nblocks : 16384;
    hipLaunchKernelGGL(thread_colours, dim3(grid), dim3(64), 0, s, C, nelems, part_size, nblocks,
                       thrcol, cnt /* nthrcol staged in the count array */, err);
  }
  hipLaunchKernelGGL(zero_ints, dim3(1), dim3(64), 0, s, ncol, 1LL);
  if (items > 0) hipLaunchKernelGGL(max_colour, blocks(items), dim3(256), 0, s, colour,
                                    static_cast<int>(items), ncol);
  int hc[2] = {0, 0};
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(&hc[0], ncol, sizeof(int), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipMemcpyAsync(&hc[1], err, sizeof(int), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) {
    op2hip::set_error("device plan failed: %s", hipGetErrorString(e));
    return OP2HIP_ERR_HIP;
  }
  OP2HIP_REQUIRE(!hc[1], "device plan: a block touches more targets than its table holds");
  ColourPlan* p = new ColourPlan();
  p->n = nelems;
  p->ps = part_size;
  p->mode = mode;
  p->work = base;
  p->w = w;
  p->C = C;
  p->ncolors = hc[0];
  if (mode == 2) {
    // elements per colour, on the host: the blocks are cut per colour
    const int ntiles = (nelems + kTile - 1) / kTile;
    int* toff = at<int>(base, w.toff);
    const long long tc = static_cast<long long>(p->ncolors) * ntiles;
    if (ntiles > 0)
      hipLaunchKernelGGL(tile_counts, dim3(ntiles), dim3(kTile), 0, s, colour, nelems, p->ncolors,
                         ntiles, cnt);
    exclusive_scan(cnt, static_cast<int>(tc), toff, nullptr, sums, total, s);
    std::vector<int> starts(p->ncolors + 1, 0);
    e = hipGetLastError();
    for (int c = 0; c < p->ncolors && e == hipSuccess; ++c)
      e = hipMemcpyAsync(&starts[c], toff + static_cast<long long>(c) * ntiles, sizeof(int),
                         hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
      delete p;
      op2hip::set_error("device plan ordering failed: %s", hipGetErrorString(e));
      return OP2HIP_ERR_HIP;
    }
    starts[p->ncolors] = nelems;
    p->counts.resize(p->ncolors);
    int nb = 0;
    for (int c = 0; c < p->ncolors; ++c) {
      p->counts[c] = starts[c + 1] - starts[c];
      nb += (p->counts[c] + part_size - 1) / part_size;
    }
    p->nblocks = nb;
  } else {
    p->nblocks = nblocks;
  }
  *nblocks_out = p->nblocks;
  *ncolors_out = p->ncolors;
  *handle = p;
  return OP2HIP_OK;
}

// Phase 2: the plan's arrays.  elems (device, n); thrcol (device, n:
// written by phase 1 in mode 0, zeroed here in mode 2); offset, nelem,
// nthrcol, blkcol, blkmap (host, nblocks) and ncolblk (host, ncolors).
// Waits for the stream; releases the handle.
int op2hip_plan_dfill(void* handle, int* elems, int* thrcol, int* offset, int* nelem,
                      int* nthrcol, int* blkcol, int* blkmap, int* ncolblk, void* stream) {
  ColourPlan* p = static_cast<ColourPlan*>(handle);
  OP2HIP_REQUIRE(p && (p->n == 0 || (elems && thrcol)), "null device plan output");
  hipStream_t s = static_cast<hipStream_t>(stream);
  char* base = p->work;
  const Work& w = p->w;
  int* colour = at<int>(base, w.colour);
  int* cnt = at<int>(base, w.tcnt);
  auto blocks = [](long long k) { return dim3(static_cast<unsigned>((k + 255) / 256 > 0 ? (k + 255) / 256 : 1)); };
  hipError_t e = hipSuccess;
  if (p->mode == 2) {
    const int ntiles = (p->n + kTile - 1) / kTile;
    if (ntiles > 0) {
      hipLaunchKernelGGL(tile_place, dim3(ntiles), dim3(64), 0, s, colour, p->n, p->ncolors,
                         ntiles, at<int>(base, w.toff), elems);
      hipLaunchKernelGGL(zero_ints, blocks(p->n), dim3(256), 0, s, thrcol,
                         static_cast<long long>(p->n));
    }
    int b = 0, start = 0;
    for (int c = 0; c < p->ncolors; ++c) {
      int nbc = 0;
      for (int o = 0; o < p->counts[c]; o += p->ps, ++b, ++nbc) {
        offset[b] = start + o;
        nelem[b] = p->counts[c] - o < p->ps ? p->counts[c] - o : p->ps;
        nthrcol[b] = 1;
        blkcol[b] = c;
        blkmap[b] = b;
      }
      ncolblk[c] = nbc;
      start += p->counts[c];
    }
    e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(s);
  } else {
    // blocks are consecutive runs of the element order
    if (p->n > 0) {
      hipLaunchKernelGGL(iota, blocks(p->n), dim3(256), 0, s, elems, p->n);
      e = hipGetLastError();
      std::vector<int> bc(p->nblocks), nt(p->nblocks);
      if (e == hipSuccess && p->nblocks)
        e = hipMemcpyAsync(bc.data(), colour, sizeof(int) * p->nblocks, hipMemcpyDeviceToHost, s);
      if (e == hipSuccess && p->nblocks)
        e = hipMemcpyAsync(nt.data(), cnt, sizeof(int) * p->nblocks, hipMemcpyDeviceToHost, s);
      if (e == hipSuccess) e = hipStreamSynchronize(s);
      if (e == hipSuccess) {
        std::vector<int> cc(p->ncolors, 0);
        for (int b = 0; b < p->nblocks; ++b) {
          offset[b] = b * p->ps;
          nelem[b] = p->n - b * p->ps < p->ps ? p->n - b * p->ps : p->ps;
          nthrcol[b] = nt[b];
          blkcol[b] = bc[b];
          cc[bc[b]]++;
        }
        std::vector<int> at_c(p->ncolors + 1, 0);
        for (int c = 0; c < p->ncolors; ++c) {
          ncolblk[c] = cc[c];
          at_c[c + 1] = at_c[c] + cc[c];
        }
        for (int b = 0; b < p->nblocks; ++b) blkmap[at_c[bc[b]]++] = b;
      }
    }
  }
  delete p;
  if (e != hipSuccess) {
    op2hip::set_error("device plan fill failed: %s", hipGetErrorString(e));
    return OP2HIP_ERR_HIP;
  }
  return OP2HIP_OK;
}

int op2hip_plan_ddiscard(void* handle) {
  delete static_cast<ColourPlan*>(handle);
  return OP2HIP_OK;
}

}  // extern "C"
