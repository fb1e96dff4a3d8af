// Execution plan for indirect loops with conflicting writes.
//
// Replaces the reference's Cython plan (pyop2/plan.pyx:68-456).  The
// reference cuts the iteration set into contiguous blocks of `part_size`
// elements (plan.pyx:107-110) and colours blocks, then threads within a
// block, with greedy 32-colour bit-mask rounds (plan.pyx:214-376).  On meshes
// whose numbering is permuted, contiguous blocks touch targets all over the
// mesh and need as many colours as there are blocks, so this plan first
// orders the iteration set into compact regions grown breadth-first through
// the conflict maps, then colours with 64-bit masks:
//
//   elems[n]            execution order (blocks are consecutive runs of it)
//   offset/nelem[b]     the run of block b inside elems
//   thrcol[n]           thread colour of elems[i]; nthrcol[b] colours in b
//   blkcol[b]           block colour; blocks of one colour share no target
//   blkmap[nblocks]     block ids sorted (stably) by colour; ncolblk[c]
//
// Invariants (checked by op2hip_plan_verify): no two elements of one
// (block, thread colour) share a race target, and no two blocks of one
// colour share a race target.
#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <thread>
#include <vector>

#include "common.h"

namespace {

struct RaceMap {
  const int* values;
  int dim;
  int idx;    // -1: every column of the map
  int space;  // id of the target index space (a Dat, or a Mat's rows)
};

struct Plan {
  int nelems = 0;
  int nblocks = 0;
  int ncolors = 0;
  int max_nelem = 0;
  int max_thrcol = 0;
  std::vector<int> elems, offset, nelem, thrcol, nthrcol, blkcol, blkmap, ncolblk;
};

template <class F>
inline void for_targets(const std::vector<RaceMap>& maps, int e, F&& f) {
  for (size_t m = 0; m < maps.size(); ++m) {
    const RaceMap& rm = maps[m];
    const int* row = rm.values + static_cast<int64_t>(e) * rm.dim;
    if (rm.idx >= 0) {
      f(rm.space, row[rm.idx]);
    } else {
      for (int k = 0; k < rm.dim; ++k) f(rm.space, row[k]);
    }
  }
}

// Greedy colouring of `items` (each a list of (space, target) touches) with
// 64-bit masks per target, in rounds of 64 colours.  `touch(item, cb)` calls
// cb(space, target) for every target of an item.  Returns colours per item.
template <class Touch>
int greedy_colour(int nitems, std::vector<std::vector<uint64_t>>& masks, Touch&& touch,
                  std::vector<int>& colour) {
  colour.assign(nitems, -1);
  int base = 0;
  int remaining = nitems;
  int ncolours = 0;
  while (remaining > 0) {
    std::vector<std::pair<int, int>> touched;  // (space, target) to reset
    for (int it = 0; it < nitems; ++it) {
      if (colour[it] >= 0) continue;
      uint64_t mask = 0;
      touch(it, [&](int s, int t) { mask |= masks[s][t]; });
      if (mask == ~0ull) continue;  // no free colour in this round
      int c = __builtin_ctzll(~mask);
      colour[it] = base + c;
      ncolours = std::max(ncolours, base + c + 1);
      --remaining;
      touch(it, [&](int s, int t) {
        if (masks[s][t] == 0) touched.emplace_back(s, t);
        masks[s][t] |= (1ull << c);
      });
    }
    for (auto& st : touched) masks[st.first][st.second] = 0;
    base += 64;
  }
  return ncolours;
}

// Breadth-first region growing over the element graph (elements adjacent
// when they share a target of a race map).  Regions hold `ps` elements; the
// seed of the next region is the oldest unassigned element on the frontier
// of earlier regions, so regions tile the mesh in a sweep.
void grow_regions(int n, int ps, const std::vector<RaceMap>& maps,
                  const std::vector<int>& space_size, std::vector<int>& order,
                  std::vector<int>& region_start) {
  // target -> element incidence (CSR) per space
  size_t nspace = space_size.size();
  std::vector<std::vector<int>> ptr(nspace), adj(nspace);
  for (size_t s = 0; s < nspace; ++s) ptr[s].assign(space_size[s] + 1, 0);
  for (int e = 0; e < n; ++e)
    for_targets(maps, e, [&](int s, int t) { ptr[s][t + 1]++; });
  for (size_t s = 0; s < nspace; ++s) {
    for (int t = 0; t < space_size[s]; ++t) ptr[s][t + 1] += ptr[s][t];
    adj[s].resize(ptr[s][space_size[s]]);
  }
  {
    std::vector<std::vector<int>> cur(nspace);
    for (size_t s = 0; s < nspace; ++s) cur[s].assign(ptr[s].begin(), ptr[s].end() - 1);
    for (int e = 0; e < n; ++e)
      for_targets(maps, e, [&](int s, int t) { adj[s][cur[s][t]++] = e; });
  }
  std::vector<int> region(n, -1);
  std::deque<int> frontier;
  std::vector<int> q;
  q.reserve(ps);
  order.clear();
  order.reserve(n);
  region_start.clear();
  int next_scan = 0;
  int r = 0;
  while (static_cast<int>(order.size()) < n) {
    int seed = -1;
    while (!frontier.empty()) {
      int x = frontier.front();
      frontier.pop_front();
      if (region[x] < 0) {
        seed = x;
        break;
      }
    }
    if (seed < 0) {
      while (region[next_scan] >= 0) ++next_scan;
      seed = next_scan;
    }
    region_start.push_back(static_cast<int>(order.size()));
    q.clear();
    q.push_back(seed);
    region[seed] = r;
    size_t head = 0;
    while (true) {
      while (head < q.size()) {
        int x = q[head++];
        for_targets(maps, x, [&](int s, int t) {
          for (int k = ptr[s][t]; k < ptr[s][t + 1]; ++k) {
            int y = adj[s][k];
            if (region[y] >= 0) continue;
            if (static_cast<int>(q.size()) < ps) {
              region[y] = r;
              q.push_back(y);
            } else {
              frontier.push_back(y);
            }
          }
        });
      }
      // the region's component is exhausted before it is full: keep filling
      // it from the next unassigned element, so blocks stay ps elements
      if (static_cast<int>(q.size()) >= ps ||
          static_cast<int>(order.size() + q.size()) >= n)
        break;
      int next = -1;
      while (!frontier.empty()) {
        int x = frontier.front();
        frontier.pop_front();
        if (region[x] < 0) {
          next = x;
          break;
        }
      }
      if (next < 0) {
        while (next_scan < n && region[next_scan] >= 0) ++next_scan;
        if (next_scan >= n) break;
        next = next_scan;
      }
      region[next] = r;
      q.push_back(next);
    }
    std::sort(q.begin(), q.end());
    order.insert(order.end(), q.begin(), q.end());
    ++r;
  }
  region_start.push_back(n);
}

}  // namespace

extern "C" {

// Build a plan over the iteration range [0, nelems).
//   maps[i]       host int32 map values, row-major (rows >= nelems, dims[i] cols)
//   idxs[i]       column used (-1: all columns)
//   spaces[i]     target index space of map i (0..nspaces-1)
//   space_size[s] number of targets in space s
//   reorder       1: locality-ordered regions, 0: contiguous blocks
int op2hip_plan_build(int nelems, int part_size, int reorder, int nmaps, const int* const* maps,
                      const int* dims, const int* idxs, const int* spaces, int nspaces,
                      const int* space_size, void** plan_out) {
  OP2HIP_REQUIRE(plan_out, "null plan_out");
  OP2HIP_REQUIRE(nelems >= 0 && part_size > 0, "bad plan sizes");
  OP2HIP_REQUIRE(nmaps >= 0 && nspaces >= 0, "bad map counts");
  std::vector<RaceMap> rmaps;
  for (int i = 0; i < nmaps; ++i) {
    OP2HIP_REQUIRE(maps[i] && dims[i] > 0, "bad race map %d", i);
    OP2HIP_REQUIRE(spaces[i] >= 0 && spaces[i] < nspaces, "bad space of map %d", i);
    OP2HIP_REQUIRE(idxs[i] < dims[i], "bad index of map %d", i);
    rmaps.push_back({maps[i], dims[i], idxs[i], spaces[i]});
  }
  std::vector<int> ssize(space_size, space_size + nspaces);
  // every target must lie inside its space
  for (int e = 0; e < nelems; ++e) {
    bool ok = true;
    int bad_s = 0, bad_t = 0;
    for_targets(rmaps, e, [&](int s, int t) {
      if (t < 0 || t >= ssize[s]) {
        ok = false;
        bad_s = s;
        bad_t = t;
      }
    });
    OP2HIP_REQUIRE(ok, "map value %d out of range [0,%d) at element %d", bad_t, ssize[bad_s], e);
  }
  Plan* p = new Plan();
  p->nelems = nelems;
  if (reorder == 2 && !rmaps.empty() && nelems > 0) {
    // element colouring: every element gets the first colour none of its
    // targets carries yet (elements in order), elements are grouped by
    // colour (ascending element order within a colour) and each colour's
    // run is cut into blocks of part_size.  All blocks of a colour are
    // conflict-free together and inside, so each block has one thread
    // colour and a launch per colour runs without intra-block rounds.
    std::vector<std::vector<uint64_t>> emasks(nspaces);
    for (int s = 0; s < nspaces; ++s) emasks[s].assign(ssize[s], 0);
    std::vector<int> ecol;
    const int nc = greedy_colour(
        nelems, emasks, [&](int it, auto&& cb) { for_targets(rmaps, it, cb); }, ecol);
    std::vector<int> cstart(nc + 1, 0);
    for (int e = 0; e < nelems; ++e) cstart[ecol[e] + 1]++;
    for (int c = 0; c < nc; ++c) cstart[c + 1] += cstart[c];
    std::vector<int> cur(cstart.begin(), cstart.end() - 1);
    p->elems.resize(nelems);
    for (int e = 0; e < nelems; ++e) p->elems[cur[ecol[e]]++] = e;
    p->ncolors = nc;
    p->ncolblk.assign(nc, 0);
    for (int c = 0; c < nc; ++c)
      for (int s0 = cstart[c]; s0 < cstart[c + 1]; s0 += part_size) {
        const int len = std::min(part_size, cstart[c + 1] - s0);
        p->offset.push_back(s0);
        p->nelem.push_back(len);
        p->blkcol.push_back(c);
        p->ncolblk[c]++;
        p->max_nelem = std::max(p->max_nelem, len);
      }
    p->nblocks = static_cast<int>(p->offset.size());
    p->thrcol.assign(nelems, 0);
    p->nthrcol.assign(p->nblocks, 1);
    p->max_thrcol = 1;
    p->blkmap.resize(p->nblocks);
    for (int b = 0; b < p->nblocks; ++b) p->blkmap[b] = b;
    *plan_out = p;
    return OP2HIP_OK;
  }
  std::vector<int> starts;
  if (reorder && !rmaps.empty() && nelems > 0) {
    grow_regions(nelems, part_size, rmaps, ssize, p->elems, starts);
  } else {
    p->elems.resize(nelems);
    for (int e = 0; e < nelems; ++e) p->elems[e] = e;
    for (int s = 0; s < nelems; s += part_size) starts.push_back(s);
    starts.push_back(nelems);
  }
  p->nblocks = static_cast<int>(starts.size()) - 1;
  p->offset.resize(p->nblocks);
  p->nelem.resize(p->nblocks);
  for (int b = 0; b < p->nblocks; ++b) {
    p->offset[b] = starts[b];
    p->nelem[b] = starts[b + 1] - starts[b];
    p->max_nelem = std::max(p->max_nelem, p->nelem[b]);
  }
  std::vector<std::vector<uint64_t>> masks(nspaces);
  for (int s = 0; s < nspaces; ++s) masks[s].assign(ssize[s], 0);
  // thread colours, block by block
  p->thrcol.assign(nelems, 0);
  p->nthrcol.assign(p->nblocks, rmaps.empty() ? 1 : 0);
  if (!rmaps.empty()) {
    std::vector<int> col;
    for (int b = 0; b < p->nblocks; ++b) {
      const int* el = p->elems.data() + p->offset[b];
      int nb = p->nelem[b];
      int nc = greedy_colour(
          nb, masks,
          [&](int it, auto&& cb) { for_targets(rmaps, el[it], cb); }, col);
      for (int k = 0; k < nb; ++k) p->thrcol[p->offset[b] + k] = col[k];
      p->nthrcol[b] = nc;
      p->max_thrcol = std::max(p->max_thrcol, nc);
    }
  } else {
    p->max_thrcol = 1;
  }
  // block colours
  std::vector<int> bcol;
  if (!rmaps.empty()) {
    p->ncolors = greedy_colour(
        p->nblocks, masks,
        [&](int b, auto&& cb) {
          const int* el = p->elems.data() + p->offset[b];
          for (int k = 0; k < p->nelem[b]; ++k) for_targets(rmaps, el[k], cb);
        },
        bcol);
  } else {
    bcol.assign(p->nblocks, 0);
    p->ncolors = p->nblocks > 0 ? 1 : 0;
  }
  p->blkcol = bcol;
  p->ncolblk.assign(p->ncolors, 0);
  for (int b = 0; b < p->nblocks; ++b) p->ncolblk[bcol[b]]++;
  p->blkmap.resize(p->nblocks);
  std::vector<int> cur(p->ncolors + 1, 0);
  for (int c = 0; c < p->ncolors; ++c) cur[c + 1] = cur[c] + p->ncolblk[c];
  for (int b = 0; b < p->nblocks; ++b) p->blkmap[cur[bcol[b]]++] = b;
  *plan_out = p;
  return OP2HIP_OK;
}

enum {
  PLAN_NELEMS = 0,
  PLAN_NBLOCKS = 1,
  PLAN_NCOLORS = 2,
  PLAN_MAX_NELEM = 3,
  PLAN_MAX_THRCOL = 4,
};

int op2hip_plan_scalar(void* plan, int field, int64_t* out) {
  Plan* p = static_cast<Plan*>(plan);
  OP2HIP_REQUIRE(p && out, "null plan");
  switch (field) {
    case PLAN_NELEMS: *out = p->nelems; break;
    case PLAN_NBLOCKS: *out = p->nblocks; break;
    case PLAN_NCOLORS: *out = p->ncolors; break;
    case PLAN_MAX_NELEM: *out = p->max_nelem; break;
    case PLAN_MAX_THRCOL: *out = p->max_thrcol; break;
    default: OP2HIP_REQUIRE(false, "unknown plan field %d", field);
  }
  return OP2HIP_OK;
}

// Copy one plan array into `dst` (host int32 of the array's length):
// 0 elems, 1 offset, 2 nelem, 3 thrcol, 4 nthrcol, 5 blkcol, 6 blkmap, 7 ncolblk.
int op2hip_plan_array(void* plan, int which, int* dst) {
  Plan* p = static_cast<Plan*>(plan);
  OP2HIP_REQUIRE(p && dst, "null plan");
  const std::vector<int>* v = nullptr;
  switch (which) {
    case 0: v = &p->elems; break;
    case 1: v = &p->offset; break;
    case 2: v = &p->nelem; break;
    case 3: v = &p->thrcol; break;
    case 4: v = &p->nthrcol; break;
    case 5: v = &p->blkcol; break;
    case 6: v = &p->blkmap; break;
    case 7: v = &p->ncolblk; break;
    default: OP2HIP_REQUIRE(false, "unknown plan array %d", which);
  }
  if (!v->empty()) memcpy(dst, v->data(), v->size() * sizeof(int));
  return OP2HIP_OK;
}

int op2hip_plan_destroy(void* plan) {
  delete static_cast<Plan*>(plan);
  return OP2HIP_OK;
}

// Check the conflict-freedom invariants of a plan (given as arrays, so a
// plan from any source can be checked) against the race maps.  Sets *ok to
// 1 when every element appears once, no (block, thread colour) touches a
// target from two elements, and no two blocks of one colour share a target.
// One element naming a target twice is one kernel call, not a conflict.
int op2hip_plan_verify(int nelems, int nblocks, const int* elems, const int* offset,
                       const int* nelem, const int* thrcol, const int* blkcol, int nmaps,
                       const int* const* maps, const int* dims, const int* idxs,
                       const int* spaces, int nspaces, const int* space_size, int* ok) {
  OP2HIP_REQUIRE(ok, "null ok");
  std::vector<RaceMap> rmaps;
  for (int i = 0; i < nmaps; ++i) rmaps.push_back({maps[i], dims[i], idxs[i], spaces[i]});
  *ok = 1;
  std::vector<char> seen(nelems, 0);
  int64_t total = 0;
  for (int b = 0; b < nblocks; ++b) {
    for (int k = 0; k < nelem[b]; ++k) {
      int e = elems[offset[b] + k];
      if (e < 0 || e >= nelems || seen[e]) {
        *ok = 0;
        return OP2HIP_OK;
      }
      seen[e] = 1;
      ++total;
    }
  }
  if (total != nelems) {
    *ok = 0;
    return OP2HIP_OK;
  }
  // stamp[s][t]: last (block, colour) tag; bstamp[s][t]: last block;
  // bcolour[s][t]: colours of the blocks that touched t, as a set of
  // (colour, block) pairs kept in a flat per-target list
  std::vector<std::vector<int64_t>> stamp(nspaces);
  for (int s = 0; s < nspaces; ++s) stamp[s].assign(space_size[s], -1);
  std::vector<std::pair<int, int>> ts;
  for (int b = 0; b < nblocks && *ok; ++b) {
    for (int k = 0; k < nelem[b] && *ok; ++k) {
      int pos = offset[b] + k;
      int64_t tag = static_cast<int64_t>(b) * 1000003 + thrcol[pos];
      ts.clear();
      for_targets(rmaps, elems[pos], [&](int s, int t) { ts.emplace_back(s, t); });
      std::sort(ts.begin(), ts.end());
      ts.erase(std::unique(ts.begin(), ts.end()), ts.end());
      for (auto& st : ts) {
        if (stamp[st.first][st.second] == tag) *ok = 0;
        stamp[st.first][st.second] = tag;
      }
    }
  }
  if (!*ok) return OP2HIP_OK;
  std::vector<std::vector<std::vector<std::pair<int, int>>>> touches(nspaces);
  for (int s = 0; s < nspaces; ++s) touches[s].resize(space_size[s]);
  for (int b = 0; b < nblocks && *ok; ++b) {
    for (int k = 0; k < nelem[b]; ++k) {
      for_targets(rmaps, elems[offset[b] + k], [&](int s, int t) {
        auto& lst = touches[s][t];
        bool have = false;
        for (auto& cb : lst) {
          if (cb.second == b) have = true;
          else if (cb.first == blkcol[b]) *ok = 0;
        }
        if (!have) lst.emplace_back(blkcol[b], b);
      });
    }
  }
  return OP2HIP_OK;
}

// ---- owner-computes plan -------------------------------------------------------

namespace {

struct OwnerPlan {
  int nelems = 0, ntargets = 0, nac = 0, k = 0;
  int nactive = 0, max_exec = 0;
  std::vector<int> blk_t0, blk_t1, ex_off, exec, tp, contrib, tg_off, tgt, rng_lo, rng_n;
};

}  // namespace

// Owner-computes plan for a loop whose conflicting writes are increments into
// one target set.  Targets are owned by blocks in contiguous ranges of k
// (contiguous positions are mesh-local under the device layout).  A block
// executes every element that names one of its targets, in ascending element
// order, and sums the increments of each owned target in that order (the
// reference's sequential order), so no two blocks write one target and no
// colouring or atomic is needed.  Elements named by several blocks' targets
// run once per block; the block owning the element's first target is its
// primary (the one whose Global contributions count).
//
// elem_order (may be NULL: index order) lists the elements in the order
// their increments are summed: the user numbering when the iteration set is
// laid out in another order, which reproduces the sequential order exactly.
//
// Race columns ("ac"): race map i contributes columns idxs[i] (or all of its
// dims[i] columns when idxs[i] < 0), numbered in order across maps.
//   blk_t0/blk_t1[nactive]  owned target range of active block b
//   rng_lo/rng_n[nactive]    when block b's primary elements are one run of
//                            consecutive elements, that run: slots
//                            [0, rng_n) are elements rng_lo + slot (else
//                            rng_n = 0)
//   ex_off[nactive+1], exec  the block's other executed elements, slots
//                            rng_n onward; bit 31 set when not primary
//   tp[ntargets+1], contrib  per target, its (slot * nac + ac) increments in
//                            ascending element order (slot: block-local)
//   tg_off[nactive+1], tgt   the targets of each block that receive increments
int op2hip_owner_plan_build(int nelems, int ntargets, int nmaps, const int* const* maps,
                            const int* dims, const int* idxs, int k, const int* elem_order,
                            void** plan_out) {
  OP2HIP_REQUIRE(plan_out && nelems >= 0 && ntargets >= 0 && k > 0 && nmaps > 0,
                 "bad owner plan arguments");
  struct Col {
    const int* v;
    int dim, col;
  };
  std::vector<Col> cols;
  for (int m = 0; m < nmaps; ++m) {
    OP2HIP_REQUIRE(maps[m] && dims[m] > 0 && idxs[m] < dims[m], "bad race map %d", m);
    if (idxs[m] >= 0) {
      cols.push_back({maps[m], dims[m], idxs[m]});
    } else {
      for (int c = 0; c < dims[m]; ++c) cols.push_back({maps[m], dims[m], c});
    }
  }
  const int nac = static_cast<int>(cols.size());
  auto target = [&](int e, int ac) {
    const Col& c = cols[ac];
    return c.v[static_cast<long long>(e) * c.dim + c.col];
  };
  // incidences per target, ascending element then column
  std::vector<int> cnt(ntargets + 1, 0);
  for (int e = 0; e < nelems; ++e)
    for (int ac = 0; ac < nac; ++ac) {
      int t = target(e, ac);
      OP2HIP_REQUIRE(t >= 0 && t < ntargets, "map value %d out of range [0,%d)", t, ntargets);
      cnt[t + 1]++;
    }
  for (int t = 0; t < ntargets; ++t) cnt[t + 1] += cnt[t];
  std::vector<long long> inc(static_cast<size_t>(cnt[ntargets]));
  {
    std::vector<int> cur(cnt.begin(), cnt.end() - 1);
    for (int i = 0; i < nelems; ++i) {
      int e = elem_order ? elem_order[i] : i;
      OP2HIP_REQUIRE(e >= 0 && e < nelems, "bad element order entry %d", e);
      for (int ac = 0; ac < nac; ++ac) {
        int t = target(e, ac);
        inc[cur[t]++] = static_cast<long long>(e) * nac + ac;
      }
    }
  }
  OwnerPlan* p = new OwnerPlan();
  p->nelems = nelems;
  p->ntargets = ntargets;
  p->nac = nac;
  p->k = k;
  p->tp.assign(ntargets + 1, 0);
  p->contrib.reserve(inc.size());
  p->ex_off.push_back(0);
  p->tg_off.push_back(0);
  std::vector<int> els, prim, nonp;
  for (int t0 = 0; t0 < ntargets; t0 += k) {
    int t1 = std::min(ntargets, t0 + k);
    els.clear();
    for (int t = t0; t < t1; ++t)
      for (int q = cnt[t]; q < cnt[t + 1]; ++q) els.push_back(static_cast<int>(inc[q] / nac));
    std::sort(els.begin(), els.end());
    els.erase(std::unique(els.begin(), els.end()), els.end());
    // primary elements (first race column owned here) as one run when they
    // are consecutive, then the others; slots number them in that order
    prim.clear();
    nonp.clear();
    for (int e : els) (target(e, 0) / k == t0 / k ? prim : nonp).push_back(e);
    const bool run = !prim.empty() && prim.back() - prim.front() + 1 == static_cast<int>(prim.size());
    const int rn = run ? static_cast<int>(prim.size()) : 0;
    auto slot_of = [&](int e) -> int {
      if (!run) return static_cast<int>(std::lower_bound(els.begin(), els.end(), e) - els.begin());
      if (e >= prim.front() && e <= prim.back()) return e - prim.front();
      return rn + static_cast<int>(std::lower_bound(nonp.begin(), nonp.end(), e) - nonp.begin());
    };
    for (int t = t0; t < t1; ++t) {
      for (int q = cnt[t]; q < cnt[t + 1]; ++q) {
        int e = static_cast<int>(inc[q] / nac);
        int ac = static_cast<int>(inc[q] % nac);
        p->contrib.push_back(slot_of(e) * nac + ac);
      }
      p->tp[t + 1] = static_cast<int>(p->contrib.size());
    }
    if (els.empty()) continue;
    if (run) {
      for (int e : nonp) p->exec.push_back(static_cast<int>(static_cast<unsigned>(e) | 0x80000000u));
    } else {
      for (int e : els) {
        bool primary = target(e, 0) / k == t0 / k;
        p->exec.push_back(primary ? e : static_cast<int>(static_cast<unsigned>(e) | 0x80000000u));
      }
    }
    p->rng_lo.push_back(run ? prim.front() : 0);
    p->rng_n.push_back(rn);
    p->blk_t0.push_back(t0);
    p->blk_t1.push_back(t1);
    p->ex_off.push_back(static_cast<int>(p->exec.size()));
    for (int t = t0; t < t1; ++t)
      if (cnt[t + 1] > cnt[t]) p->tgt.push_back(t);
    p->tg_off.push_back(static_cast<int>(p->tgt.size()));
    p->max_exec = std::max(p->max_exec, static_cast<int>(els.size()));
  }
  p->nactive = static_cast<int>(p->blk_t0.size());
  *plan_out = p;
  return OP2HIP_OK;
}

// 0 nactive, 1 max_exec, 2 nac, 3 exec length, 4 contrib length, 5 tgt length
int op2hip_owner_plan_scalar(void* plan, int field, int64_t* out) {
  OwnerPlan* p = static_cast<OwnerPlan*>(plan);
  OP2HIP_REQUIRE(p && out, "null plan");
  switch (field) {
    case 0: *out = p->nactive; break;
    case 1: *out = p->max_exec; break;
    case 2: *out = p->nac; break;
    case 3: *out = static_cast<int64_t>(p->exec.size()); break;
    case 4: *out = static_cast<int64_t>(p->contrib.size()); break;
    case 5: *out = static_cast<int64_t>(p->tgt.size()); break;
    default: OP2HIP_REQUIRE(false, "unknown owner plan field %d", field);
  }
  return OP2HIP_OK;
}

// 0 blk_t0, 1 blk_t1, 2 ex_off, 3 exec, 4 tp, 5 contrib, 6 tg_off, 7 tgt,
// 8 rng_lo, 9 rng_n
int op2hip_owner_plan_array(void* plan, int which, int* dst) {
  OwnerPlan* p = static_cast<OwnerPlan*>(plan);
  OP2HIP_REQUIRE(p && dst, "null plan");
  const std::vector<int>* v = nullptr;
  switch (which) {
    case 0: v = &p->blk_t0; break;
    case 1: v = &p->blk_t1; break;
    case 2: v = &p->ex_off; break;
    case 3: v = &p->exec; break;
    case 4: v = &p->tp; break;
    case 5: v = &p->contrib; break;
    case 6: v = &p->tg_off; break;
    case 7: v = &p->tgt; break;
    case 8: v = &p->rng_lo; break;
    case 9: v = &p->rng_n; break;
    default: OP2HIP_REQUIRE(false, "unknown owner plan array %d", which);
  }
  if (!v->empty()) memcpy(dst, v->data(), v->size() * sizeof(int));
  return OP2HIP_OK;
}

int op2hip_owner_plan_destroy(void* plan) {
  delete static_cast<OwnerPlan*>(plan);
  return OP2HIP_OK;
}

// ---- data layout orderings ---------------------------------------------------

// Breadth-first (Cuthill-McKee style) ordering of the targets of a map: two
// targets are adjacent when one iteration element names both.  Starts from a
// pseudo-peripheral target (the last one reached by a first sweep), restarts
// at the lowest unvisited target for every further component.  perm[pos] is
// the target placed at position pos.  Gives mesh-local data contiguous
// positions, so gathers through the map reuse cache lines.
int op2hip_order_bfs(int ntarget, int nelem, const int* map, int dim, int* perm) {
  OP2HIP_REQUIRE(ntarget >= 0 && nelem >= 0 && dim > 0 && perm, "bad ordering arguments");
  std::vector<int> ptr(ntarget + 1, 0), adj(static_cast<size_t>(nelem) * dim);
  for (long long k = 0; k < static_cast<long long>(nelem) * dim; ++k) {
    int t = map[k];
    OP2HIP_REQUIRE(t >= 0 && t < ntarget, "map value %d out of range [0,%d)", t, ntarget);
    ptr[t + 1]++;
  }
  for (int t = 0; t < ntarget; ++t) ptr[t + 1] += ptr[t];
  {
    std::vector<int> cur(ptr.begin(), ptr.end() - 1);
    for (int e = 0; e < nelem; ++e)
      for (int d = 0; d < dim; ++d) adj[cur[map[static_cast<long long>(e) * dim + d]]++] = e;
  }
  std::vector<int> pos(ntarget, -1);
  std::vector<int> q;
  q.reserve(ntarget);
  auto bfs = [&](int start, std::vector<int>& order, std::vector<int>& mark, int stamp) {
    size_t head = order.size();
    order.push_back(start);
    mark[start] = stamp;
    while (head < order.size()) {
      int t = order[head++];
      for (int k = ptr[t]; k < ptr[t + 1]; ++k) {
        const int* row = map + static_cast<long long>(adj[k]) * dim;
        for (int d = 0; d < dim; ++d) {
          int u = row[d];
          if (mark[u] != stamp) {
            mark[u] = stamp;
            order.push_back(u);
          }
        }
      }
    }
  };
  std::vector<int> mark(ntarget, -1);
  int next = 0;
  int comp = 0;
  while (static_cast<int>(q.size()) < ntarget) {
    while (next < ntarget && pos[next] >= 0) ++next;
    // first sweep from `next` to find a far end of its component
    std::vector<int> sweep;
    bfs(next, sweep, mark, 2 * comp);
    int start = sweep.back();
    size_t before = q.size();
    bfs(start, q, mark, 2 * comp + 1);
    for (size_t k = before; k < q.size(); ++k) pos[q[k]] = static_cast<int>(k);
    ++comp;
  }
  memcpy(perm, q.data(), sizeof(int) * ntarget);
  return OP2HIP_OK;
}

}  // extern "C"

namespace {

// Runs f(thread, lo, hi) over [0, n) split into nt contiguous pieces.
template <class F>
void parallel_ranges(int n, int nt, F f) {
  if (nt <= 1) {
    f(0, 0, n);
    return;
  }
  std::vector<std::thread> th;
  for (int i = 0; i < nt; ++i) {
    int lo = static_cast<int>(static_cast<long long>(n) * i / nt);
    int hi = static_cast<int>(static_cast<long long>(n) * (i + 1) / nt);
    th.emplace_back(f, i, lo, hi);
  }
  for (auto& t : th) t.join();
}

// Host threads for plan and ordering work: OP2HIP_HOST_THREADS if set, else
// the hardware's, at most 16; one for small inputs.
int host_threads(long long work) {
  if (const char* env = getenv("OP2HIP_HOST_THREADS")) {
    int v = atoi(env);
    if (v > 0) return std::min(v, 64);
  }
  if (work < (1 << 20)) return 1;
  unsigned h = std::thread::hardware_concurrency();
  return static_cast<int>(std::max(1u, std::min(h, 16u)));
}

// Distinct-neighbour adjacency of the targets of a map (CSR): two targets
// are adjacent when one element names both.  Map values are checked by the
// caller.
struct Adjacency {
  std::vector<long long> ptr;
  std::vector<int> adj;
};

void build_adjacency(int ntarget, int nelem, const int* map, int dim, int nt, Adjacency& g) {
  std::vector<long long> eptr(ntarget + 1, 0);
  std::vector<int> eadj(static_cast<size_t>(nelem) * dim);
  for (long long k = 0; k < static_cast<long long>(nelem) * dim; ++k) eptr[map[k] + 1]++;
  for (int t = 0; t < ntarget; ++t) eptr[t + 1] += eptr[t];
  {
    std::vector<long long> cur(eptr.begin(), eptr.end() - 1);
    for (int e = 0; e < nelem; ++e)
      for (int d = 0; d < dim; ++d) eadj[cur[map[static_cast<long long>(e) * dim + d]]++] = e;
  }
  // each thread collects the neighbour lists of a contiguous target range
  std::vector<std::vector<int>> part(nt);
  std::vector<long long> deg(ntarget + 1, 0);
  parallel_ranges(ntarget, nt, [&](int th, int lo, int hi) {
    std::vector<int> seen(ntarget, -1);
    std::vector<int>& out = part[th];
    for (int t = lo; t < hi; ++t) {
      seen[t] = t;
      size_t before = out.size();
      for (long long k = eptr[t]; k < eptr[t + 1]; ++k) {
        const int* row = map + static_cast<long long>(eadj[k]) * dim;
        for (int d = 0; d < dim; ++d) {
          int u = row[d];
          if (seen[u] != t) {
            seen[u] = t;
            out.push_back(u);
          }
        }
      }
      deg[t + 1] = static_cast<long long>(out.size() - before);
    }
  });
  g.ptr.assign(ntarget + 1, 0);
  for (int t = 0; t < ntarget; ++t) g.ptr[t + 1] = g.ptr[t] + deg[t + 1];
  g.adj.resize(static_cast<size_t>(g.ptr[ntarget]));
  parallel_ranges(ntarget, nt, [&](int th, int lo, int hi) {
    if (lo < hi) std::copy(part[th].begin(), part[th].end(), g.adj.begin() + g.ptr[lo]);
  });
}

// Breadth-first order of a whole graph: every component from a
// pseudo-peripheral vertex (the last reached by a sweep from its lowest
// vertex), components in order of their lowest vertex.  comp receives the
// first position of every component and n.
void bfs_order(const Adjacency& g, int n, std::vector<int>& ord, std::vector<int>& comp) {
  std::vector<int> mark(n, 0), sweep;
  ord.clear();
  ord.reserve(n);
  int next = 0;
  auto run = [&](int s, std::vector<int>& out, int stamp) {
    size_t head = out.size();
    out.push_back(s);
    mark[s] = stamp;
    while (head < out.size()) {
      int t = out[head++];
      for (long long k = g.ptr[t]; k < g.ptr[t + 1]; ++k) {
        int u = g.adj[k];
        if (mark[u] < stamp) {
          mark[u] = stamp;
          out.push_back(u);
        }
      }
    }
  };
  comp.clear();
  while (static_cast<int>(ord.size()) < n) {
    while (mark[next] == 2) ++next;
    sweep.clear();
    run(next, sweep, 1);
    comp.push_back(static_cast<int>(ord.size()));
    run(sweep.back(), ord, 2);
  }
  comp.push_back(n);
}

}  // namespace

extern "C" {

// Recursive bisection ordering of the targets of a map (adjacency as for
// op2hip_order_bfs).  A range of targets is ordered breadth-first from a
// pseudo-peripheral target of the subgraph it induces (components one after
// another) and cut at its middle; each half is ordered the same way, down to
// ranges of at most `leaf` targets, which keep their breadth-first order.
// Every aligned range of the result is a compact patch of the mesh at every
// scale (a graph analogue of a space-filling curve), so a block that owns a
// range of targets shares few elements with its neighbours.  Connected
// components are ordered one after another (in order of their lowest target)
// and each exactly as it would be on its own.  The graph is
// first renumbered breadth-first for cache locality; independent ranges are
// ordered on up to 16 host threads (the result does not depend on how many).
int op2hip_order_bisect(int ntarget, int nelem, const int* map, int dim, int leaf, int* perm) {
  OP2HIP_REQUIRE(ntarget >= 0 && nelem >= 0 && dim > 0 && perm && leaf > 0,
                 "bad ordering arguments");
  if (ntarget == 0) return OP2HIP_OK;
  for (long long k = 0; k < static_cast<long long>(nelem) * dim; ++k)
    OP2HIP_REQUIRE(map[k] >= 0 && map[k] < ntarget, "map value %d out of range [0,%d)", map[k],
                   ntarget);
  const int nt = host_threads(static_cast<long long>(nelem) * dim * dim);
  Adjacency g0;
  build_adjacency(ntarget, nelem, map, dim, nt, g0);
  // renumber breadth-first: vertex v of the working graph is target ord[v]
  std::vector<int> ord, iord(ntarget), comp;
  bfs_order(g0, ntarget, ord, comp);
  for (int v = 0; v < ntarget; ++v) iord[ord[v]] = v;
  Adjacency g;
  g.ptr.assign(ntarget + 1, 0);
  for (int v = 0; v < ntarget; ++v) g.ptr[v + 1] = g.ptr[v] + (g0.ptr[ord[v] + 1] - g0.ptr[ord[v]]);
  g.adj.resize(g0.adj.size());
  parallel_ranges(ntarget, nt, [&](int, int lo, int hi) {
    for (int v = lo; v < hi; ++v) {
      long long o = g0.ptr[ord[v]], k = g.ptr[v], e = g.ptr[v + 1];
      for (; k < e; ++k, ++o) g.adj[k] = iord[g0.adj[o]];
    }
  });
  Adjacency().adj.swap(g0.adj);

  const int n = ntarget;
  // part[v]: the range v currently belongs to, as lo * (n + 1) + hi
  std::vector<long long> part(n, static_cast<long long>(n));
  std::vector<int> order(n);
  std::vector<long long> stamp(n, -1), placed(n, -1);
  for (int v = 0; v < n; ++v) order[v] = v;
  struct Range { int lo, hi; };
  // stamps are unique across threads: thread th uses th, th + nt, ...
  auto order_range = [&](Range r, int th, long long& counter, std::vector<int>& seq,
                         std::vector<int>& sweep, std::vector<Range>& out) {
    const long long pid = static_cast<long long>(r.lo) * (n + 1) + r.hi;
    const int len = r.hi - r.lo;
    auto bfs = [&](int s, std::vector<int>& o, long long st) {
      size_t head = o.size();
      o.push_back(s);
      stamp[s] = st;
      while (head < o.size()) {
        int t = o[head++];
        for (long long k = g.ptr[t]; k < g.ptr[t + 1]; ++k) {
          int u = g.adj[k];
          if (part[u] == pid && stamp[u] != st) {
            stamp[u] = st;
            o.push_back(u);
          }
        }
      }
    };
    // breadth-first order of the range, component by component (a
    // component never holds targets placed before, so sweeps need no check)
    seq.clear();
    int scan = r.lo;
    const long long mark = th + nt * counter++;
    while (static_cast<int>(seq.size()) < len) {
      while (placed[order[scan]] == mark) ++scan;
      sweep.clear();
      bfs(order[scan], sweep, th + nt * counter++);
      const size_t b = seq.size();
      bfs(sweep.back(), seq, th + nt * counter++);
      for (size_t k = b; k < seq.size(); ++k) placed[seq[k]] = mark;
    }
    std::copy(seq.begin(), seq.end(), order.begin() + r.lo);
    if (len <= leaf) return;
    const int mid = r.lo + len / 2;
    const long long left = static_cast<long long>(r.lo) * (n + 1) + mid;
    const long long right = static_cast<long long>(mid) * (n + 1) + r.hi;
    for (int k = r.lo; k < mid; ++k) part[order[k]] = left;
    for (int k = mid; k < r.hi; ++k) part[order[k]] = right;
    out.push_back({mid, r.hi});
    out.push_back({r.lo, mid});
  };
  // Components are bisected separately (a component is ordered as it
  // would be alone).  The working numbering is already a breadth-first
  // order of each component from a pseudo-peripheral vertex: it is the
  // first level's order.
  std::vector<Range> level, next;
  for (size_t c = 0; c + 1 < comp.size(); ++c) {
    const int lo = comp[c], hi = comp[c + 1];
    if (hi - lo <= leaf) continue;
    const int mid = lo + (hi - lo) / 2;
    const long long left = static_cast<long long>(lo) * (n + 1) + mid;
    const long long right = static_cast<long long>(mid) * (n + 1) + hi;
    for (int k = lo; k < mid; ++k) part[k] = left;
    for (int k = mid; k < hi; ++k) part[k] = right;
    level.push_back({lo, mid});
    level.push_back({mid, hi});
  }
  // further levels range by range on the threads, until every thread has
  // several ranges; then each range with its whole subtree on one thread
  std::vector<long long> counters(nt, 0);
  std::vector<std::vector<int>> seqs(nt), sweeps(nt);
  std::vector<std::vector<Range>> outs(nt);
  while (!level.empty() && static_cast<int>(level.size()) < 4 * nt) {
    std::atomic<int> cursor(0);
    const int nr = static_cast<int>(level.size());
    std::vector<std::vector<Range>> kids(nr);
    parallel_ranges(std::min(nt, nr), std::min(nt, nr), [&](int th, int, int) {
      for (int i = cursor++; i < nr; i = cursor++)
        order_range(level[i], th, counters[th], seqs[th], sweeps[th], kids[i]);
    });
    next.clear();
    for (int i = 0; i < nr; ++i)
      for (auto it = kids[i].rbegin(); it != kids[i].rend(); ++it) next.push_back(*it);
    level.swap(next);
  }
  std::sort(level.begin(), level.end(),
            [](const Range& a, const Range& b) { return a.hi - a.lo > b.hi - b.lo; });
  std::atomic<int> cursor(0);
  const int nr = static_cast<int>(level.size());
  parallel_ranges(nt, nt, [&](int th, int, int) {
    std::vector<Range> stack;
    for (int i = cursor++; i < nr; i = cursor++) {
      stack.assign(1, level[i]);
      while (!stack.empty()) {
        Range r = stack.back();
        stack.pop_back();
        order_range(r, th, counters[th], seqs[th], sweeps[th], stack);
      }
    }
  });
  for (int pos = 0; pos < n; ++pos) perm[pos] = ord[order[pos]];
  return OP2HIP_OK;
}

// First-touch ordering of the targets of a map from an ordered source set:
// walking the source elements in `src_order`, each target takes the next
// position when first named; targets never named follow in index order.
int op2hip_order_first_touch(int ntarget, int nsrc, const int* map, int dim,
                             const int* src_order, int* perm) {
  OP2HIP_REQUIRE(ntarget >= 0 && nsrc >= 0 && dim > 0 && perm, "bad ordering arguments");
  std::vector<char> seen(ntarget, 0);
  int k = 0;
  for (int i = 0; i < nsrc; ++i) {
    int e = src_order ? src_order[i] : i;
    const int* row = map + static_cast<long long>(e) * dim;
    for (int d = 0; d < dim; ++d) {
      int t = row[d];
      OP2HIP_REQUIRE(t >= 0 && t < ntarget, "map value %d out of range [0,%d)", t, ntarget);
      if (!seen[t]) {
        seen[t] = 1;
        perm[k++] = t;
      }
    }
  }
  for (int t = 0; t < ntarget; ++t)
    if (!seen[t]) perm[k++] = t;
  return OP2HIP_OK;
}

}  // extern "C"

// ---- owner Mat assembly program ---------------------------------------------

extern "C" {

// Scatter program of an owner-computes Mat assembly: for every node pair
// (row node n, column node m) of the Mat's node-level pattern, in pattern
// order, the list of (block-local contribution index c, column position j)
// whose local-matrix entry [i0][j] adds into that pair, in the order the
// sequential backend adds them (ascending user element, then j).  A
// contribution is an (element, row position i0) pair of an owned row node;
// the owner plan lists them per node in ascending user element order.
//
// The node-level pattern comes from the scalar CSR (rowptr/colidx, device
// numbering) with rd x cd blocks: node n's pairs are the columns of scalar
// row n*rd taken every cd-th entry, and pair p of n starts at
// rowptr[n*rd] / (rd*cd).  Outputs (caller-allocated): poff[npairs+1]
// (npairs = rowptr[nnodes*rd] / (rd*cd)) and prog[ncontrib * e1] holding
// c * e1 + j as uint16 (wide == 0) or int32 (wide == 1).  Fails if an
// element names a column the pattern lacks or a code does not fit.
int op2hip_mat_program(int nactive, const int* blk_t0, const int* blk_t1, const int* ex_off,
                       const int* exec, const int* rng_lo, const int* rng_n, const int* tp,
                       const int* contrib, int nac,
                       const int* cmap, int carity, int e1, int nnodes, const int* rowptr,
                       const int* colidx, int rd, int cd, int wide, int* poff, void* prog) {
  OP2HIP_REQUIRE(nactive >= 0 && nac > 0 && carity > 0 && e1 > 0 && e1 <= carity && rd > 0 &&
                     cd > 0 && poff && prog,
                 "bad Mat program arguments");
  const long long npairs = static_cast<long long>(rowptr[static_cast<long long>(nnodes) * rd]) /
                           (static_cast<long long>(rd) * cd);
  // entries per pair: each contribution of node n adds e1 entries to n's pairs
  std::vector<int> cnt(static_cast<size_t>(npairs) + 1, 0);
  std::atomic<int> bad(0);
  const int nt = host_threads(static_cast<long long>(tp[nnodes]) * e1);
  auto pair_of = [&](int n, int m) -> long long {
    // columns of scalar row n*rd, one per cd: binary search for m
    const int s0 = rowptr[static_cast<long long>(n) * rd], s1 = rowptr[static_cast<long long>(n) * rd + 1];
    int lo = 0, hi = (s1 - s0) / cd - 1;
    while (lo <= hi) {
      const int mid = (lo + hi) >> 1;
      const int v = colidx[s0 + mid * cd] / cd;
      if (v == m) return static_cast<long long>(s0) / (static_cast<long long>(rd) * cd) + mid;
      if (v < m) lo = mid + 1; else hi = mid - 1;
    }
    return -1;
  };
  // pass 1: count (pairs of a node are touched only by its own block)
  parallel_ranges(nactive, nt, [&](int, int b0, int b1) {
    for (int b = b0; b < b1; ++b)
      for (int n = blk_t0[b]; n < blk_t1[b]; ++n)
        for (int c = tp[n]; c < tp[n + 1]; ++c) {
          const int slot = contrib[c] / nac;
          const int e = slot < rng_n[b] ? rng_lo[b] + slot
              : static_cast<int>(static_cast<unsigned>(exec[ex_off[b] + slot - rng_n[b]]) & 0x7fffffffu);
          for (int j = 0; j < e1; ++j) {
            const long long p = pair_of(n, cmap[static_cast<long long>(e) * carity + j]);
            if (p < 0) { bad = 1; continue; }
            cnt[p + 1]++;
          }
        }
  });
  OP2HIP_REQUIRE(!bad, "an element names a matrix entry outside the sparsity pattern");
  for (long long p = 0; p < npairs; ++p) cnt[p + 1] += cnt[p];
  OP2HIP_REQUIRE(static_cast<long long>(cnt[npairs]) ==
                     static_cast<long long>(tp[nnodes]) * e1 || nactive == 0,
                 "Mat program size mismatch");
  for (long long p = 0; p <= npairs; ++p) poff[p] = cnt[p];
  // pass 2: fill, contributions of each node in their (element) order
  parallel_ranges(nactive, nt, [&](int, int b0, int b1) {
    std::vector<int> cur;
    for (int b = b0; b < b1; ++b) {
      const int cbase = tp[blk_t0[b]];
      for (int n = blk_t0[b]; n < blk_t1[b]; ++n)
        for (int c = tp[n]; c < tp[n + 1]; ++c) {
          const int slot = contrib[c] / nac;
          const int e = slot < rng_n[b] ? rng_lo[b] + slot
              : static_cast<int>(static_cast<unsigned>(exec[ex_off[b] + slot - rng_n[b]]) & 0x7fffffffu);
          const long long code0 = static_cast<long long>(c - cbase) * e1;
          for (int j = 0; j < e1; ++j) {
            const long long p = pair_of(n, cmap[static_cast<long long>(e) * carity + j]);
            const long long code = code0 + j;
            const int at = cnt[p]++;
            if (wide) {
              static_cast<int*>(prog)[at] = static_cast<int>(code);
            } else {
              if (code > 0xffff) bad = 1;
              static_cast<uint16_t*>(prog)[at] = static_cast<uint16_t>(code);
            }
          }
        }
    }
  });
  OP2HIP_REQUIRE(!bad, "Mat program codes do not fit 16 bits");
  return OP2HIP_OK;
}

}  // extern "C"
