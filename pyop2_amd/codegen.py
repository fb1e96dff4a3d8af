"""Generation of HIP loop sources around op2 user kernels.

A ``par_loop`` is compiled (by hiprtc, in :mod:`pyop2_amd.jit`) from a user
kernel string placed inside one of two hand-written loop templates:

``stream``
    One thread per iteration element in a grid-stride loop over a range of
    the iteration set, natural order.  Used when no argument needs conflict
    handling: direct loops and indirect loops that only read through maps
    (or write without conflicts, WRITE/RW, which op2 leaves to the user).

``coloured``
    Driven by a plan: each workgroup runs one plan block; the increments of
    indirect INC arguments and the entries of Mat arguments are accumulated
    in registers and applied thread colour by thread colour, a barrier
    between colours, and one launch per block colour.  No atomics, and the
    order in which increments reach a target is fixed by the plan, so runs
    are bit-reproducible.

The kernel-argument contract is that of the reference's sequential backend
(pyop2/host.py:108-134, sequential.py:55-72): a direct Dat is ``dat + e*cdim``;
an indirect Dat ``dat + map[e*mdim+idx]*cdim``; a vector map an array of
``mdim`` such pointers; an iteration-space Dat uses ``map[e*mdim+i_n]``; a
Global is a pointer to its value (a private accumulator for reductions); a
scalar Mat is ``(T (*)[1])&local[i_0][i_1]`` and a vector Mat
``local[rdim][cdim]``; Consts are global variables of the module; the
iteration-space indices ``i_0, i_1`` are trailing int arguments.

Global reductions are reduced per workgroup (wavefront shuffles, then LDS
across wavefronts) into per-block partials, then a one-workgroup pass in
fixed order combines them with the Global's value: INC adds, MIN/MAX take
the min/max (reference semantics: pyop2/base.py:216-248).
"""

import re

import numpy as np

from .base import IterationIndex, READ, WRITE, RW, INC, MIN, MAX, canonical


# contribution codes an owner write-back loads together per target
OWNER_CMAX = 4


class ArgInfo(object):
    """How one ``par_loop`` argument is lowered into the loop template."""

    def __init__(self, pos, arg, extents):
        self.pos = pos
        self.arg = arg
        self.access = arg.access
        self.ctype = arg.data.ctype
        d = arg.data
        if arg._is_global:
            self.kind = 'gbl_red' if arg._is_global_reduction else 'gbl_read'
            self.cdim = d.cdim
        elif arg._is_mat:
            self.kind = 'mat'
            self.rdim, self.cdim_m = d.dims
            self.rarity = arg.map[0].dim
            self.carity = arg.map[1].dim
            self.ridx = arg.idx[0].index
            self.cidx = arg.idx[1].index
            self.vector = d._is_vector_field
        elif arg._is_direct:
            self.kind = 'direct'
            self.cdim = d.cdim
        else:
            self.cdim = d.cdim
            self.mdim = arg.map.dim
            if arg._is_vec_map:
                self.kind = 'vec'
            elif isinstance(arg.idx, IterationIndex):
                self.kind = 'itspace'
                self.iidx = arg.idx.index
                self.extent = extents[self.iidx]
            else:
                self.kind = 'ind'
                self.idx = arg.idx
        self.race = (self.kind == 'mat') or \
            (self.kind in ('ind', 'vec', 'itspace') and self.access is INC)
        # structure-of-arrays device storage (reference: device.py:153-174,
        # cuda_direct_loop.jinja2:80-82): component k of row r at
        # k * stride + r, stride the Dat's rows on the device.  The kernel
        # sees a private copy of each row (OP2_STRIDE(a, k) is a[k]), so
        # the loads and stores of consecutive elements are contiguous per
        # component.
        self.soa = bool(self.kind in ('direct', 'ind', 'vec', 'itspace') and d.soa and d.cdim > 1)
        self.stride = d.dataset.total_size if self.soa else 0
        if self.soa and self.kind in ('vec', 'itspace') and self.access not in (READ, INC):
            raise NotImplementedError("structure-of-arrays Dats are written through vector "
                                      "and iteration-space maps only by INC")

    def at(self, row, k):
        """Index expression of component ``k`` of device row ``row``."""
        if self.soa:
            return '(op2_i64)(%s) * %d + (op2_i64)(%s)' % (k, self.stride, row)
        return '(op2_i64)(%s) * %d + (%s)' % (row, self.cdim, k)

    # parameter names
    @property
    def d(self):
        return 'd%d' % self.pos

    @property
    def m(self):
        return 'm%d' % self.pos

    def params(self):
        """(C declaration, role) of the kernel parameters of this argument."""
        t, p = self.ctype, self.pos
        if self.kind == 'gbl_read':
            return [('%s* g%d' % (t, p), 'gbl')]
        if self.kind == 'gbl_red':
            return [('const %s* g%d' % (t, p), 'gbl'),
                    ('%s* r%d' % (t, p), 'partial')]
        if self.kind == 'direct':
            return [('%s* d%d' % (t, p), 'data')]
        if self.kind == 'mat':
            return [('%s* d%d' % (t, p), 'mat_vals'),
                    ('const int* __restrict__ rp%d' % p, 'mat_rowptr'),
                    ('const int* __restrict__ ci%d' % p, 'mat_colidx'),
                    ('const int* __restrict__ m%d' % p, 'mat_rmap'),
                    ('const int* __restrict__ mc%d' % p, 'mat_cmap')]
        # data pointers carry no __restrict__: one Dat may be passed through
        # several arguments (e.g. res via pevcell[0] and pevcell[1])
        return [('%s* d%d' % (t, p), 'data'),
                ('const int* __restrict__ m%d' % p, 'map')]


_REDUCE_OPS = {'INC': ('a + b', '0'), 'MIN': ('(b < a ? b : a)', None),
               'MAX': ('(b > a ? b : a)', None)}


def _mode(access):
    """Mode name of an access descriptor (this module's or, matched by
    name, another op2 implementation's)."""
    return {id(READ): 'READ', id(WRITE): 'WRITE', id(RW): 'RW', id(INC): 'INC',
            id(MIN): 'MIN', id(MAX): 'MAX'}[id(canonical(access))]


class LoopSource(object):
    """HIP source of one loop kernel plus its reduction pass."""

    PRELUDE = r'''
#define OP2_STRIDE(a, idx) a[idx]
// owner-plan contribution codes: 16-bit when they fit (own_c16), else 32-bit
#define OP2_CONTRIB(i) (own_c16 ? (int)((const unsigned short*)p_contrib)[i] : ((const int*)p_contrib)[i])
typedef long long op2_i64;
template <typename T> __device__ inline T op2_shfl_xor(T v, int m) { return __shfl_xor(v, m, 64); }
__device__ inline unsigned int op2_shfl_xor(unsigned int v, int m) {
  return (unsigned int)__shfl_xor((int)v, m, 64);
}
__device__ inline unsigned long op2_shfl_xor(unsigned long v, int m) {
  return (unsigned long)__shfl_xor((long long)v, m, 64);
}
__device__ inline unsigned long long op2_shfl_xor(unsigned long long v, int m) {
  return (unsigned long long)__shfl_xor((long long)v, m, 64);
}
__device__ inline int op2_csr_find(const int* __restrict__ rowptr, const int* __restrict__ colidx,
                                   int row, int col) {
  int lo = rowptr[row], hi = rowptr[row + 1] - 1;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (colidx[mid] < col) lo = mid + 1; else hi = mid;
  }
  return (lo <= hi && colidx[lo] == col) ? lo : -1;
}
// A deferred FP64 Global reduction (the combine pass of an earlier loop)
// run by one extra workgroup of this loop's launch, its first: op 0 INC,
// 1 MIN, 2 MAX.  The workgroup's waves stand in for the 16 wavefronts of
// the 1024-thread combine kernel (its thread t combines partials t,
// t + 1024, t + 2048, t + 3072, then t + 4096, ... in that order; then a
// shuffle tree per wavefront and the 16 wavefront values in order), so the
// Global gets that kernel's value bit for bit.  A wave takes two of its
// stand-in wavefronts at a time, their loads issued together (more would
// raise the register count of every loop kernel).
__device__ inline double op2_pend_op(int op, double a, double b) {
  return op == 0 ? a + b : op == 1 ? (b < a ? b : a) : (b > a ? b : a);
}
__device__ inline void op2_pending_reduce(op2_i64 nparts, const double* __restrict__ r,
                                          double* g, op2_i64 info) {
  __shared__ double op2_pend_s[16];
  const int cdim = (int)(info & 0xffff), op = (int)(info >> 16);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
  for (int k = 0; k < cdim; ++k) {
    for (int vg = wv; vg < 16; vg += 2 * nw) {
      double v[2];
#pragma unroll
      for (int j = 0; j < 2; ++j) v[j] = op == 0 ? 0.0 : g[k];
      // partial counts are workgroup counts: 32-bit offsets
      const int np = (int)nparts;
      for (int q00 = 0; q00 < np; q00 += 4096) {
        double x[2][4];
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const int vw = vg + j * nw;
            const int q = q00 + vw * 64 + lane + u * 1024;
            x[j][u] = (vw < 16 && q < np) ? r[q * cdim + k] : 0.0;
          }
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const int vw = vg + j * nw;
            const int q = q00 + vw * 64 + lane + u * 1024;
            if (vw < 16 && q < np) v[j] = op2_pend_op(op, v[j], x[j][u]);
          }
      }
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int vw = vg + j * nw;
        if (vw < 16) {
          double t = v[j];
          for (int off = 32; off > 0; off >>= 1) t = op2_pend_op(op, t, op2_shfl_xor(t, off));
          if (lane == 0) op2_pend_s[vw] = t;
        }
      }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      double v = op2_pend_s[0];
      for (int w = 1; w < 16; ++w) v = op2_pend_op(op, v, op2_pend_s[w]);
      g[k] = op2_pend_op(op, g[k], v);
    }
    __syncthreads();
  }
}
'''

    def __init__(self, kernel, extents, args, consts, mode, block_size, prefetch=1,
                 stage_reads=True, stage_writes=True, unroll=1, owner_stage=(),
                 rows_stage=False):
        assert mode in ('stream', 'coloured', 'owner', 'rows')
        # rows template: the rows its elements read through maps are staged
        # in LDS per block (distinct targets once) and read through local
        # positions, instead of gathered per element
        self.rows_stage = bool(rows_stage) and mode == 'rows'
        # owner write-back: 1 issues the loads locating a thread's first
        # target before the elements run (they overlap), 0 after them; 2
        # also loads that target's old value early
        self.prefetch = prefetch
        # occupancy request (waves per SIMD) for the register budget, 0: none
        self.min_waves = 0
        # argument position -> position whose map parameter it reads (args
        # whose maps share one device array read one parameter, so the
        # element's map entries are loaded once)
        self.map_of = {}
        self.stage_reads = stage_reads
        self.stage_writes = stage_writes
        # elements per thread and chunk of a direct loop (their loads overlap)
        self.unroll = unroll
        # owner loops: positions of read-only FP64 Dats on the increment
        # target's layout whose owned rows a block stages in LDS
        self.owner_stage = tuple(owner_stage) if mode == 'owner' else ()
        self.kernel = kernel
        self.extents = tuple(extents)
        self.infos = [ArgInfo(k, a, self.extents) for k, a in enumerate(args)]
        # arguments naming one Dat share its staged rows (the first one's)
        self._stage_of = {}
        for q in self.owner_stage:
            d = self.infos[q].arg.data
            self._stage_of[q] = next(r for r in self.owner_stage if self.infos[r].arg.data is d)
        self.consts = consts
        self.mode = mode
        self.bs = block_size
        # kernel names carry the user kernel's name so profiles tell loops apart
        self.entry = 'op2_%s_%s' % (kernel.name, mode)
        self.reduce_entry = ('op2_%s_reduce' % kernel.name) if self.reductions else None

    @property
    def reductions(self):
        return [a for a in self.infos if a.kind == 'gbl_red']

    @property
    def has_mats(self):
        return any(a.kind == 'mat' for a in self.infos)

    # -- parameter layout --------------------------------------------------
    def param_roles(self):
        """Roles of the loop kernel's parameters, in order, for the launcher."""
        roles = [('start', None), ('count', None)]
        if self.mode == 'coloured':
            roles += [('plan_elems', None), ('plan_blkmap', None), ('plan_offset', None),
                      ('plan_nelem', None), ('plan_nthrcol', None), ('plan_thrcol', None)]
        elif self.mode == 'owner':
            roles += [('own_t0', None), ('own_t1', None), ('own_rlo', None), ('own_rn', None),
                      ('own_tgoff', None),
                      ('own_tgt', None), ('own_tgcb', None), ('own_exoff', None), ('own_exec', None),
                      ('own_tp', None), ('own_contrib', None), ('own_c16', None), ('own_pitch', None),
                      ('own_k', None),
                      ('own_cpitch', None), ('own_poff', None), ('own_prog', None),
                      ('own_wide', None)]
        elif self.mode == 'rows':
            roles += [('own_tp', None), ('own_xoff', None), ('own_exec', None),
                      ('own_contrib', None), ('own_c16', None), ('own_pos', None),
                      ('own_k', None), ('own_nnodes', None), ('own_slice', None)]
            if self.rows_stage:
                roles.append(('own_stage', None))
                for m in self.rows_maps():
                    roles += [('rows_sn', m), ('rows_so', m), ('rows_xl', m)]
            else:
                roles += [('rows_xmap', m) for m in self.rows_maps()]
        for a in self.infos:
            for _, role in a.params():
                roles.append((role, a.pos))
        if self.has_mats:
            roles.append(('err_flag', None))
        # a deferred Global reduction of an earlier loop (zero when none)
        roles += [('pend_n', None), ('pend_part', None), ('pend_out', None), ('pend_info', None)]
        return roles

    def reduce_roles(self):
        roles = [('nparts', None)]
        for a in self.reductions:
            roles += [('gbl_out', a.pos), ('partial', a.pos)]
        return roles

    # -- code pieces -------------------------------------------------------
    def _const_decls(self):
        out = []
        for c in self.consts:
            if c.cdim == 1:
                out.append('__constant__ %s %s;' % (c.ctype, c.name))
            else:
                out.append('__constant__ %s %s[%d];' % (c.ctype, c.name, c.cdim))
        return '\n'.join(out)

    def _red_init(self):
        lines = []
        for a in self.reductions:
            op = _mode(a.access)
            lines.append('%s red%d[%d];' % (a.ctype, a.pos, a.cdim))
            for k in range(a.cdim):
                init = '(%s)0' % a.ctype if op == 'INC' else 'g%d[%d]' % (a.pos, k)
                lines.append('red%d[%d] = %s;' % (a.pos, k, init))
        return lines

    def _red_finish(self, part_index):
        """Block reduction of the private accumulators into partials."""
        lines = []
        if not self.reductions:
            return lines
        nw = self.bs // 64
        for a in self.reductions:
            op = _mode(a.access)
            expr = _REDUCE_OPS[op][0]
            lines.append('{')
            lines.append('  __shared__ %s sred%d[%d][%d];' % (a.ctype, a.pos, nw, a.cdim))
            lines.append('  for (int k = 0; k < %d; ++k) {' % a.cdim)
            lines.append('    %s v = red%d[k];' % (a.ctype, a.pos))
            lines.append('    for (int off = 32; off > 0; off >>= 1) {')
            lines.append('      %s a = v, b = op2_shfl_xor(v, off); v = %s;' % (a.ctype, expr))
            lines.append('    }')
            lines.append('    if ((threadIdx.x & 63) == 0) sred%d[threadIdx.x >> 6][k] = v;'
                         % a.pos)
            lines.append('  }')
            lines.append('  __syncthreads();')
            lines.append('  if (threadIdx.x == 0) {')
            lines.append('    for (int k = 0; k < %d; ++k) {' % a.cdim)
            lines.append('      %s v = sred%d[0][k];' % (a.ctype, a.pos))
            lines.append('      for (int w = 1; w < %d; ++w) { %s a = v, b = sred%d[w][k]; v = %s; }'
                         % (nw, a.ctype, a.pos, expr))
            lines.append('      r%d[(op2_i64)(%s) * %d + k] = v;' % (a.pos, part_index, a.cdim))
            lines.append('    }')
            lines.append('  }')
            lines.append('}')
        return lines

    def _element_body(self, sfx='', ev='e', split=False, infos=None, kernel=None, mslot=None):
        """Code that runs the kernel for element ``ev`` (all itspace points),
        leaving INC increments and Mat entries in private arrays; private
        names carry the suffix ``sfx``.

        Returns (decls, compute): ``decls`` declares and zeroes the private
        arrays and touches no memory, so it can run for idle threads of a
        workgroup; ``compute`` reads maps and data and calls the kernel.
        With ``split`` it returns (decls, loads, compute): ``loads`` copies
        the rows of read-only Dats into registers (local rows), so the
        template can issue the loads of several elements before their
        kernels run.  With ``split='maps'`` it returns (decls, mloads, loads,
        compute): ``mloads`` copies the map entries of the local rows of
        ind/vec arguments into private ints (``mi<pos>``, whatever ``sfx``),
        which ``loads`` then reads, so a template can load an element's map
        entries one step ahead of its rows; with ``mslot`` (an expression)
        ``mloads`` reads them from the exec-ordered map rows ``xm<pos>`` at
        that slot instead of the map at the element."""
        decls, loads, compute, call_args, zeros, post = [], [], [], [], [], []
        mloads, mseen = [], set()
        by_maps = split == 'maps'
        ext = self.extents
        infos = self.infos if infos is None else infos
        kernel = self.kernel if kernel is None else kernel
        written = set(id(a.arg.data) for a in infos
                      if a.kind not in ('gbl_read', 'gbl_red') and a.access is not READ)
        E = '(op2_i64)%s' % ev
        for a in infos:
            p, t = a.pos, a.ctype
            n = lambda base: '%s%d%s' % (base, p, sfx)
            # read-only rows of Dats no argument writes go through registers
            local = (a.access is READ and a.kind in ('direct', 'ind', 'vec') and
                     id(a.arg.data) not in written and
                     a.cdim * (a.mdim if a.kind == 'vec' else 1) <= 16 and
                     not (a.kind == 'direct' and self.mode == 'stream' and self._staged(a)))
            if a.kind == 'gbl_read':
                call_args.append('g%d' % p)
            elif a.kind == 'gbl_red':
                if self.mode == 'owner':
                    decls.append('%s %s[%d];' % (t, n('junk'), a.cdim))
                    decls.append('for (int k = 0; k < %d; ++k) %s[k] = red%d[k];'
                                 % (a.cdim, n('junk'), p))
                    call_args.append('(prim ? red%d : %s)' % (p, n('junk')))
                else:
                    call_args.append('red%d' % p)
            elif a.kind == 'direct' and a.soa:
                # a private row, loaded (unless written only) and stored back
                decls.append('%s %s[%d];' % (t, n('rl'), a.cdim))
                if a.access is not WRITE:
                    (loads if local else compute).append(
                        'for (int k = 0; k < %d; ++k) %s[k] = d%d[%s];'
                        % (a.cdim, n('rl'), p, a.at(E, 'k')))
                if a.access is not READ:
                    post.append('for (int k = 0; k < %d; ++k) d%d[%s] = %s[k];'
                                % (a.cdim, p, a.at(E, 'k'), n('rl')))
                call_args.append(n('rl'))
            elif a.kind == 'direct':
                if self.mode == 'stream' and self._staged(a):
                    call_args.append('st%d + threadIdx.x * %d' % (p, self._pitch(a)))
                elif self._row_vector(a):
                    decls.append('op2_row%d %s;' % (p, n('pr')))
                    if a.access is not WRITE:
                        compute.append('%s = *(const op2_row%d*)(d%d + %s * %d);'
                                       % (n('pr'), p, p, E, a.cdim))
                    call_args.append('%s.v' % n('pr'))
                    post.append('*(op2_row%d*)(d%d + %s * %d) = %s;'
                                % (p, p, E, a.cdim, n('pr')))
                elif local:
                    decls.append('%s %s[%d];' % (t, n('rl'), a.cdim))
                    loads.append('for (int k = 0; k < %d; ++k) %s[k] = d%d[%s * %d + k];'
                                 % (a.cdim, n('rl'), p, E, a.cdim))
                    call_args.append(n('rl'))
                else:
                    call_args.append('d%d + %s * %d' % (p, E, a.cdim))
            elif a.kind == 'ind':
                row = '(op2_i64)m%d[%s * %d + %d]' % (self._m(p), E, a.mdim, a.idx)
                if a.race:
                    decls.append('%s %s[%d];' % (t, n('inc'), a.cdim))
                    decls.append('for (int k = 0; k < %d; ++k) %s[k] = (%s)0;' % (a.cdim, n('inc'), t))
                    call_args.append(n('inc'))
                elif a.soa:
                    decls.append('%s %s[%d];' % (t, n('rl'), a.cdim))
                    if a.access is not READ:
                        decls.append('op2_i64 %s = 0;' % n('rw'))
                        compute.append('%s = %s;' % (n('rw'), row))
                        row = n('rw')
                    if a.access is not WRITE:
                        (loads if local else compute).append(
                            '{ const op2_i64 tt = %s; for (int k = 0; k < %d; ++k) %s[k] = d%d[%s]; }'
                            % (row, a.cdim, n('rl'), p, a.at('tt', 'k')))
                    if a.access is not READ:
                        post.append('for (int k = 0; k < %d; ++k) d%d[%s] = %s[k];'
                                    % (a.cdim, p, a.at(row, 'k'), n('rl')))
                    call_args.append(n('rl'))
                elif local and p in self.owner_stage:
                    # rows of the block's owned range come from LDS
                    decls.append('%s %s[%d];' % (t, n('rl'), a.cdim))
                    loads.append('{ const int tt = m%d[%s * %d + %d]; const %s* src = '
                                 '(unsigned)(tt - t0) < (unsigned)nn_own ? s_st%d + (tt - t0) * %d : '
                                 'd%d + (op2_i64)tt * %d; for (int k = 0; k < %d; ++k) %s[k] = src[k]; }'
                                 % (self._m(p), E, a.mdim, a.idx, t, self._stage_of[p], a.cdim, p,
                                    a.cdim, a.cdim, n('rl')))
                    call_args.append(n('rl'))
                elif local and by_maps:
                    mi = 'mi%d_%d' % (self._m(p), a.idx)
                    if mi not in mseen:
                        mseen.add(mi)
                        decls.append('int %s = 0;' % mi)
                        mloads.append('%s = %s[%s * %d + %d];'
                                      % (mi, *self._mrow(p, E, mslot), a.mdim, a.idx))
                    decls.append('%s %s[%d];' % (t, n('rl'), a.cdim))
                    loads.append('{ const %s* src = %s + (op2_i64)%s * %d; '
                                 'for (int k = 0; k < %d; ++k) %s[k] = src[k]; }'
                                 % (t, self._rows_src(p), mi, a.cdim, a.cdim, n('rl')))
                    call_args.append(n('rl'))
                elif local:
                    decls.append('%s %s[%d];' % (t, n('rl'), a.cdim))
                    loads.append('{ const %s* src = d%d + (op2_i64)m%d[%s * %d + %d] * %d; '
                                 'for (int k = 0; k < %d; ++k) %s[k] = src[k]; }'
                                 % (t, p, self._m(p), E, a.mdim, a.idx, a.cdim, a.cdim, n('rl')))
                    call_args.append(n('rl'))
                else:
                    call_args.append('d%d + (op2_i64)m%d[%s * %d + %d] * %d'
                                     % (p, self._m(p), E, a.mdim, a.idx, a.cdim))
            elif a.kind == 'vec':
                decls.append('%s* %s[%d];' % (t, n('v'), a.mdim))
                if a.race:
                    decls.append('%s %s[%d][%d];' % (t, n('inc'), a.mdim, a.cdim))
                    decls.append('for (int j = 0; j < %d; ++j) { %s[j] = %s[j]; '
                                 'for (int k = 0; k < %d; ++k) %s[j][k] = (%s)0; }'
                                 % (a.mdim, n('v'), n('inc'), a.cdim, n('inc'), t))
                elif a.soa:
                    # the row pointers are set next to the call, so private
                    # rows loaded one loop iteration ahead stay in registers
                    decls.append('%s %s[%d][%d];' % (t, n('rl'), a.mdim, a.cdim))
                    (loads if local else compute).append(
                        'for (int j = 0; j < %d; ++j) { const op2_i64 tt = m%d[%s * %d + j]; '
                        'for (int k = 0; k < %d; ++k) %s[j][k] = d%d[%s]; }'
                        % (a.mdim, self._m(p), E, a.mdim, a.cdim, n('rl'), p, a.at('tt', 'k')))
                    compute.append('for (int j = 0; j < %d; ++j) %s[j] = %s[j];'
                                   % (a.mdim, n('v'), n('rl')))
                elif local and by_maps:
                    mi = 'mi%d' % self._m(p)
                    if mi not in mseen:
                        mseen.add(mi)
                        decls.append('int %s[%d] = {0};' % (mi, a.mdim))
                        mloads.append('for (int j = 0; j < %d; ++j) %s[j] = %s[%s * %d + j];'
                                      % (a.mdim, mi, *self._mrow(p, E, mslot), a.mdim))
                    decls.append('%s %s[%d][%d];' % (t, n('rl'), a.mdim, a.cdim))
                    loads.append('for (int j = 0; j < %d; ++j) { const %s* src = %s + '
                                 '(op2_i64)%s[j] * %d; for (int k = 0; k < %d; ++k) '
                                 '%s[j][k] = src[k]; }'
                                 % (a.mdim, t, self._rows_src(p), mi, a.cdim, a.cdim, n('rl')))
                    compute.append('for (int j = 0; j < %d; ++j) %s[j] = %s[j];'
                                   % (a.mdim, n('v'), n('rl')))
                elif local:
                    decls.append('%s %s[%d][%d];' % (t, n('rl'), a.mdim, a.cdim))
                    loads.append('for (int j = 0; j < %d; ++j) { const %s* src = d%d + '
                                 '(op2_i64)m%d[%s * %d + j] * %d; '
                                 'for (int k = 0; k < %d; ++k) %s[j][k] = src[k]; }'
                                 % (a.mdim, t, p, self._m(p), E, a.mdim, a.cdim, a.cdim, n('rl')))
                    compute.append('for (int j = 0; j < %d; ++j) %s[j] = %s[j];'
                                   % (a.mdim, n('v'), n('rl')))
                else:
                    compute.append('for (int j = 0; j < %d; ++j) %s[j] = d%d + '
                                   '(op2_i64)m%d[%s * %d + j] * %d;'
                                   % (a.mdim, n('v'), p, self._m(p), E, a.mdim, a.cdim))
                call_args.append(n('v'))
            elif a.kind == 'itspace':
                ivar = 'i_%d' % a.iidx
                if a.race:
                    decls.append('%s %s[%d][%d];' % (t, n('inc'), a.extent, a.cdim))
                    decls.append('for (int j = 0; j < %d; ++j) for (int k = 0; k < %d; ++k) '
                                 '%s[j][k] = (%s)0;' % (a.extent, a.cdim, n('inc'), t))
                    call_args.append('%s[%s]' % (n('inc'), ivar))
                elif a.soa:
                    decls.append('%s %s[%d][%d];' % (t, n('rl'), a.extent, a.cdim))
                    compute.append('for (int j = 0; j < %d; ++j) { const op2_i64 tt = '
                                   'm%d[%s * %d + j]; for (int k = 0; k < %d; ++k) '
                                   '%s[j][k] = d%d[%s]; }'
                                   % (a.extent, self._m(p), E, a.mdim, a.cdim, n('rl'), p,
                                      a.at('tt', 'k')))
                    call_args.append('%s[%s]' % (n('rl'), ivar))
                else:
                    call_args.append('d%d + (op2_i64)m%d[%s * %d + %s] * %d'
                                     % (p, self._m(p), E, a.mdim, ivar, a.cdim))
            elif a.kind == 'mat':
                e0 = ext[0] if len(ext) > 0 else 1
                e1 = ext[1] if len(ext) > 1 else 1
                if a.vector:
                    decls.append('%s %s[%d][%d][%d][%d];' % (t, n('lm'), e0, e1, a.rdim, a.cdim_m))
                    zeros.append('for (int r = 0; r < %d; ++r) for (int c = 0; c < %d; ++c) '
                                 '%s[i_0][i_1][r][c] = (%s)0;' % (a.rdim, a.cdim_m, n('lm'), t))
                    call_args.append('%s[i_0][i_1]' % n('lm'))
                else:
                    decls.append('%s %s[%d][%d];' % (t, n('lm'), e0, e1))
                    zeros.append('%s[i_0][i_1] = (%s)0;' % (n('lm'), t))
                    call_args.append('(%s (*)[1])&%s[i_0][i_1]' % (t, n('lm')))
        call_args += ['i_%d' % k for k in range(len(ext))]
        opens = ['for (int i_%d = 0; i_%d < %d; ++i_%d) {' % (k, k, e, k)
                 for k, e in enumerate(ext)]
        mats = [a for a in infos if a.kind == 'mat']
        if mats and len(ext) < 2:
            # Mat locals are indexed by both itspace indices
            if len(ext) == 0:
                compute.append('const int i_0 = 0;')
            compute.append('const int i_1 = 0;')
            compute.append('{')
            closing = ['}']
        else:
            closing = []
        compute += opens
        compute += zeros
        compute.append('%s(%s);' % (kernel.name, ', '.join(call_args)))
        compute += ['}'] * len(opens)
        compute += closing
        compute += post
        if by_maps:
            return decls, mloads, loads, compute
        if split:
            return decls, loads, compute
        return decls, loads + compute

    def _mrow(self, p, E, mslot):
        """(array, row) of argument ``p``'s map entries: the map at element
        ``E``, or the exec-ordered map rows at slot ``mslot``."""
        if mslot is None:
            return 'm%d' % self._m(p), E
        return ('xl%d' if self.rows_stage else 'xm%d') % self._m(p), '(op2_i64)(%s)' % mslot

    def _rows_src(self, p):
        """Where argument ``p``'s rows are read through its map entries: the
        Dat, or (staged rows template) the block's LDS region, the entries
        then being local positions."""
        return 's_g%d' % self._stage_region(p) if self.rows_stage else 'd%d' % p

    def rows_maps(self):
        """Map parameter positions whose entries the rows template reads
        from exec-ordered map rows (the maps of register rows)."""
        mloads = self._element_body(split='maps', mslot='0')[1]
        out = []
        for l in mloads:
            m = int(re.search(r'= x[ml](\d+)\[', l).group(1))
            if m not in out:
                out.append(m)
        return out

    def rows_staged(self):
        """The LDS regions of a staged rows template: (argument position,
        map parameter, values per row, C type) per distinct (Dat, map) its
        register rows come through, in argument order."""
        out, seen = [], {}
        for a in self._by_map_rows():
            k = (id(a.arg.data), self._m(a.pos))
            if k not in seen:
                seen[k] = a.pos
                out.append((a.pos, self._m(a.pos), a.cdim, a.ctype))
        return out

    def _stage_region(self, p):
        """The staged region argument ``p`` reads (the first argument's of
        its Dat and map)."""
        a = self.infos[p]
        for a2 in self._by_map_rows():
            if a2.arg.data is a.arg.data and self._m(a2.pos) == self._m(p):
                return a2.pos
        raise KeyError(p)

    def _by_map_rows(self):
        """Arguments whose rows the element body of a template with
        ``split='maps'`` loads into registers through map entries."""
        written = set(id(a.arg.data) for a in self.infos
                      if a.kind not in ('gbl_read', 'gbl_red') and a.access is not READ)
        return [a for a in self.infos
                if a.kind in ('ind', 'vec') and a.access is READ and not a.soa and
                id(a.arg.data) not in written and
                a.cdim * (a.mdim if a.kind == 'vec' else 1) <= 16 and
                not (a.kind == 'ind' and a.pos in self.owner_stage)]

    def _m(self, p):
        """Position of the map parameter argument ``p`` reads."""
        return self.map_of.get(p, p)

    def _apply_increments(self):
        """Apply private INC increments and Mat entries of element ``e``."""
        lines = []
        ext = self.extents
        for a in self.infos:
            p = a.pos
            if not a.race:
                continue
            if a.kind == 'ind':
                lines.append('{ const op2_i64 tt = m%d[(op2_i64)e * %d + %d];'
                             % (p, a.mdim, a.idx))
                lines.append('  for (int k = 0; k < %d; ++k) d%d[%s] += inc%d[k]; }'
                             % (a.cdim, p, a.at('tt', 'k'), p))
            elif a.kind in ('vec', 'itspace'):
                lines.append('for (int j = 0; j < %d; ++j) { const op2_i64 tt = '
                             'm%d[(op2_i64)e * %d + j];'
                             % (a.mdim if a.kind == 'vec' else a.extent, p, a.mdim))
                lines.append('  for (int k = 0; k < %d; ++k) d%d[%s] += inc%d[j][k]; }'
                             % (a.cdim, p, a.at('tt', 'k'), p))
            elif a.kind == 'mat':
                e0 = ext[0] if len(ext) > 0 else 1
                e1 = ext[1] if len(ext) > 1 else 1
                insert = a.access is WRITE
                lines.append('for (int j0 = 0; j0 < %d; ++j0) for (int j1 = 0; j1 < %d; ++j1) {'
                             % (e0, e1))
                lines.append('  const int nr = m%d[(op2_i64)e * %d + j0];' % (p, a.rarity))
                lines.append('  const int nc = mc%d[(op2_i64)e * %d + j1];' % (p, a.carity))
                if a.vector:
                    lines.append('  for (int r = 0; r < %d; ++r) for (int c = 0; c < %d; ++c) {'
                                 % (a.rdim, a.cdim_m))
                    lines.append('    %s v = lm%d[j0][j1][r][c];' % (a.ctype, p))
                    if not insert:
                        lines.append('    if (v == (%s)0) continue;' % a.ctype)
                    lines.append('    int s = op2_csr_find(rp%d, ci%d, %d * nr + r, %d * nc + c);'
                                 % (p, p, a.rdim, a.cdim_m))
                    lines.append('    if (s < 0) { *err_flag = 1; continue; }')
                    lines.append('    d%d[s] %s v;' % (p, '=' if insert else '+='))
                    lines.append('  }')
                else:
                    lines.append('  %s v = lm%d[j0][j1];' % (a.ctype, p))
                    lines.append('  int s = op2_csr_find(rp%d, ci%d, nr, nc);' % (p, p))
                    lines.append('  if (s < 0) { *err_flag = 1; }')
                    lines.append('  else d%d[s] %s v;' % (p, '=' if insert else '+='))
                lines.append('}')
        return lines

    def _staged(self, a):
        """Direct Dats with several components go through LDS in the
        stream template, so a workgroup's loads and stores of its rows are
        contiguous instead of one strided row per thread (strided row
        stores cost partial-line writes).  Read-only ones are staged only
        when ``stage_reads`` is set."""
        if a.kind != 'direct' or a.cdim == 1 or a.soa:
            return False
        if a.access is READ:
            return self.stage_reads
        return self.stage_writes

    def _row_vector(self, a):
        """Direct Dats with several components that the kernel writes,
        when not staged through LDS: the row lives in registers and moves
        with the widest aligned vector loads and stores (whole rows)."""
        return (self.mode == 'stream' and a.kind == 'direct' and a.cdim > 1 and not a.soa and
                a.access is not READ and not self._staged(a))

    def _row_types(self):
        out = []
        for a in self.infos:
            if self._row_vector(a):
                nbytes = a.cdim * np.dtype(a.arg.data.dtype).itemsize
                w = 16
                while nbytes % w:
                    w //= 2
                out.append('struct __attribute__((aligned(%d))) op2_row%d { %s v[%d]; };'
                           % (w, a.pos, a.ctype, a.cdim))
        return '\n'.join(out)

    @staticmethod
    def _pitch(a):
        # odd row pitch (in elements) spreads a wavefront's rows over LDS banks
        return a.cdim | 1

    def _stream_body(self, decls, compute):
        ind = lambda lines, n: '\n'.join(' ' * n + l for l in lines)
        bs = self.bs
        staged = [a for a in self.infos if self._staged(a)]
        out = [ind(self._red_init(), 2)]
        for a in staged:
            out.append('  __shared__ %s st%d[%d];' % (a.ctype, a.pos, bs * self._pitch(a)))
        u = self.stream_unroll()
        # workgroups are dealt to the 8 XCDs round-robin: give each XCD a
        # contiguous run of chunks, so neighbouring chunks share its L2
        out.append('  const int nbk = (int)gridDim.x - (pend_n > 0 ? 1 : 0), bid = op2_bid;')
        out.append('  const int xcd = bid & 7, qx = nbk >> 3, rx = nbk & 7;')
        out.append('  const op2_i64 lb = xcd * qx + (xcd < rx ? xcd : rx) + (bid >> 3);')
        out.append('  const op2_i64 stride = (op2_i64)nbk * %d;' % (bs * u))
        out.append('  for (op2_i64 base = lb * %d; base < count; base += stride) {' % (bs * u))
        out.append('    const op2_i64 e = start + base + threadIdx.x;')
        out.append('    const bool active = base + threadIdx.x < count;')
        if u == 2:
            # two elements per thread, bs apart: both elements' register
            # rows are loaded before either kernel runs
            da, la, ca = self._element_body(split=True)
            db, lb_, cb = self._element_body(sfx='_b', ev='e_b', split=True)
            out.append('    const op2_i64 e_b = e + %d;' % bs)
            out.append('    const bool active_b = base + %d + threadIdx.x < count;' % bs)
            out.append(ind(da + db, 4))
            out.append('    if (active_b) {')
            out.append(ind(la + lb_ + ca + cb, 6))
            out.append('    } else if (active) {')
            out.append(ind(la + ca, 6))
            out.append('    }')
            out.append('  }')
            return out
        out.append('    const int nrow = count - base < %d ? (int)(count - base) : %d;' % (bs, bs))
        out.append('    (void)nrow;')
        loads = [a for a in staged if a.access is not WRITE]
        stores = [a for a in staged if a.access is not READ]
        for a in loads:
            out.append('    for (int w = threadIdx.x; w < nrow * %d; w += %d) '
                       'st%d[(w / %d) * %d + w %% %d] = d%d[(start + base) * %d + w];'
                       % (a.cdim, bs, a.pos, a.cdim, self._pitch(a), a.cdim, a.pos, a.cdim))
        if staged:
            out.append('    __syncthreads();')
        out.append('    if (active) {')
        out.append(ind(decls + compute, 6))
        out.append('    }')
        if stores:
            out.append('    __syncthreads();')
            for a in stores:
                out.append('    for (int w = threadIdx.x; w < nrow * %d; w += %d) '
                           'd%d[(start + base) * %d + w] = st%d[(w / %d) * %d + w %% %d];'
                           % (a.cdim, bs, a.pos, a.cdim, a.pos, a.cdim, self._pitch(a), a.cdim))
        if staged:
            out.append('    __syncthreads();')
        out.append('  }')
        return out

    def stream_unroll(self):
        """Elements per thread and chunk of the direct-loop template: 2 when
        asked for and the loop stages nothing through LDS and has no
        iteration space, else 1."""
        if self.unroll != 2 or self.mode != 'stream' or self.extents:
            return 1
        if any(self._staged(a) for a in self.infos):
            return 1
        return 2

    def owner_columns(self):
        """(arg info, first race column, number of columns) of the race
        arguments, in the race-column order of the owner plan."""
        out, ac = [], 0
        for a in self.infos:
            if not a.race:
                continue
            if a.kind == 'mat':
                n = a.rarity
            else:
                n = 1 if a.kind == 'ind' else (a.mdim if a.kind == 'vec' else a.extent)
            out.append((a, ac, n))
            ac += n
        return out

    def owner_slot_bytes(self):
        """LDS bytes per executed element in the owner template."""
        cols = self.owner_columns()
        a = cols[0][0]
        if a.kind == 'mat':
            e0, e1 = self.extents[0], self.extents[1]
            return e0 * e1 * a.rdim * a.cdim_m * 8 + a.carity * 4
        nac = sum(n for _, _, n in cols)
        return nac * a.cdim * np.dtype(a.arg.data.dtype).itemsize

    def _owner_body(self, decls, compute):
        if self.owner_columns()[0][0].kind == 'mat':
            return self._owner_mat_body(decls, compute)
        ind = lambda lines, n: '\n'.join(' ' * n + l for l in lines)
        bs = self.bs
        cols = self.owner_columns()
        nac = sum(n for _, _, n in cols)
        tgt = cols[0][0]
        t, cdim = tgt.ctype, tgt.cdim
        pf = self.prefetch
        store = []
        for a, ac0, n in cols:
            # LDS holds increments column-major, [ac][k][slot] with a row
            # pitch of lds_pitch slots: a wavefront's stores are contiguous
            if a.kind == 'ind':
                store.append('for (int k = 0; k < %d; ++k) op2_lds[(%d + k) * pitch + l] = inc%d[k];'
                             % (cdim, ac0 * cdim, a.pos))
            else:
                store.append('for (int j = 0; j < %d; ++j) for (int k = 0; k < %d; ++k) '
                             'op2_lds[((%d + j) * %d + k) * pitch + l] = inc%d[j][k];'
                             % (n, cdim, ac0, cdim, a.pos))
        sub = {'cdim': cdim, 'bs': bs, 't': t, 'pos': tgt.pos, 'nac': nac, 'cmax': OWNER_CMAX,
               'at_f': tgt.at('f_row', 'k'), 'at_w': tgt.at('w_t', 'k')}
        codes = '''#pragma unroll
    for (int u = 0; u < %(cmax)d; ++u) w_code[u] = u < w_nc ? OP2_CONTRIB(w_cb + u) : 0;''' % sub
        out = ['  extern __shared__ %s op2_lds[];' % t,
               ind(self._red_init(), 2),
               '''  // consecutive owner blocks (neighbouring target ranges) go to one XCD
  const int nb = (int)count, bid = op2_bid;
  const int xcd = bid & 7, q = nb >> 3, rr = nb & 7;
  const int b = (int)start + xcd * q + (xcd < rr ? xcd : rr) + (bid >> 3);
  // slots [0, rn) run elements rlo + slot (primary); the rest are listed
  const int rlo = p_rlo[b], rn = p_rn[b];
  const int ex0 = p_exoff[b], nex = rn + p_exoff[b + 1] - ex0;
  const int g0 = p_tgoff[b], ng = p_tgoff[b + 1] - g0;
  // a block whose every owned target receives increments needs no list
  const int t0 = p_t0[b];
  const bool dense = ng == p_t1[b] - t0;
  const int pitch = (int)lds_pitch;''' % sub]
        if self.owner_stage:
            out.append('  // the owned rows of read-only Dats on the target layout, staged')
            out.append('  const int nn_own = p_t1[b] - t0, kmax = (int)own_k;')
            off = '(op2_i64)%d * pitch' % (nac * cdim)
            for pos in self.owner_stage:
                if self._stage_of[pos] != pos:
                    continue
                a = self.infos[pos]
                out.append('  %s* s_st%d = (%s*)(op2_lds + %s) + (op2_i64)kmax * %d;'
                           % (a.ctype, pos, a.ctype, off, self._stage_offset(pos)))
                out.append('  for (int i = threadIdx.x; i < nn_own * %d; i += %d) '
                           's_st%d[i] = d%d[(op2_i64)t0 * %d + i];'
                           % (a.cdim, bs, pos, pos, a.cdim))
            out.append('  __syncthreads();')
        if pf >= 1:
            out.append('''  // write-back work of this thread's first owned target (its list entry
  // and contribution range), loaded while the elements run; the index is
  // clamped to the padding entry the device lists carry
  const int jf = (int)threadIdx.x < ng ? (int)threadIdx.x : ng;
  const int f_t = p_tgt[g0 + jf], f_cb = p_tgcb[g0 + jf], f_ce = p_tgcb[g0 + jf + 1];''')
        if pf >= 2:
            out.append('''  // and its old value (row 0 stands in when the thread owns none)
  const int f_row = dense ? (jf < ng ? t0 + jf : 0) : f_t;
  %(t)s f_old[%(cdim)d];
#pragma unroll
  for (int k = 0; k < %(cdim)d; ++k) f_old[k] = d%(pos)d[%(at_f)s];''' % sub)
        out += ['''  for (int l = threadIdx.x; l < nex; l += %d) {
    const unsigned code = l < rn ? (unsigned)(rlo + l) : (unsigned)p_exec[ex0 + l - rn];
    const int e = (int)(code & 0x7fffffffu);
    const bool prim = (code >> 31) == 0u;
    (void)prim;''' % bs,
                ind(decls + compute + store, 4),
                '  }',
                '  __syncthreads();',
                '''  // each owned target that receives increments, one per thread: its old
  // value plus its increments in the order of the plan (ascending user
  // element); contribution codes are loaded %(cmax)d at a time
  for (int j = threadIdx.x; j < ng; j += %(bs)d) {
    int w_t, w_cb, w_nc;''' % sub]
        if pf >= 1:
            out.append('''    if (j == (int)threadIdx.x) {
      w_t = dense ? t0 + j : f_t;
      w_cb = f_cb;
      w_nc = f_ce - f_cb;
    } else {
      w_t = dense ? t0 + j : p_tgt[g0 + j];
      w_cb = p_tgcb[g0 + j];
      w_nc = p_tgcb[g0 + j + 1] - w_cb;
    }''')
        else:
            out.append('''    w_t = dense ? t0 + j : p_tgt[g0 + j];
    w_cb = p_tgcb[g0 + j];
    w_nc = p_tgcb[g0 + j + 1] - w_cb;''')
        out.append('    int w_code[%(cmax)d];' % sub)
        out.append(codes)
        if pf >= 2:
            out.append('''    %(t)s acc[%(cdim)d];
#pragma unroll
    for (int k = 0; k < %(cdim)d; ++k)
      acc[k] = j == (int)threadIdx.x ? f_old[k] : d%(pos)d[%(at_w)s];''' % sub)
        else:
            out.append('''    %(t)s acc[%(cdim)d];
#pragma unroll
    for (int k = 0; k < %(cdim)d; ++k) acc[k] = d%(pos)d[%(at_w)s];''' % sub)
        out.append('''#pragma unroll
    for (int u = 0; u < %(cmax)d; ++u) {
      if (u < w_nc) {
        const int slot = w_code[u] / %(nac)d, ac = w_code[u] - slot * %(nac)d;
#pragma unroll
        for (int k = 0; k < %(cdim)d; ++k) acc[k] += op2_lds[(ac * %(cdim)d + k) * pitch + slot];
      }
    }
    for (int c = %(cmax)d; c < w_nc; ++c) {
      const int cc = OP2_CONTRIB(w_cb + c);
      const int slot = cc / %(nac)d, ac = cc - slot * %(nac)d;
#pragma unroll
      for (int k = 0; k < %(cdim)d; ++k) acc[k] += op2_lds[(ac * %(cdim)d + k) * pitch + slot];
    }
#pragma unroll
    for (int k = 0; k < %(cdim)d; ++k) d%(pos)d[%(at_w)s] = acc[k];
  }''' % sub)
        out.append(ind(self._red_finish('b'), 2))
        return out

    def _owner_mat_body(self, decls, compute):
        """Owner-computes Mat assembly: a block owns the rows of a range of
        row-set nodes.  Its contributions are the (element, row position)
        pairs of its nodes, listed per node in ascending user element order.
        Phase 1 runs the elements naming its nodes and keeps, per
        contribution, the element's local-matrix row for that node in LDS.
        Phase 2 gives each node pair (row node, column node) of the owned
        rows to a thread, which adds into the pair's CSR entries the
        local-matrix entries its scatter program lists, in sequential order
        (the order addto_vector/addto_scalar reach them)."""
        ind = lambda lines, n: '\n'.join(' ' * n + l for l in lines)
        bs = self.bs
        a = self.owner_columns()[0][0]
        p = a.pos
        e0, e1 = self.extents[0], self.extents[1]
        rd, cd = a.rdim, a.cdim_m
        ent = e1 * rd * cd          # doubles kept per contribution
        insert = a.access is WRITE
        sub = {'bs': bs, 'p': p, 'rd': rd, 'cd': cd, 'nac': a.rarity, 'e0': e0, 'e1': e1,
               'ent': ent, 'rdcd': rd * cd, 'op': '=' if insert else '+='}
        val = ('lm%(p)d[i0][i1][r][c]' if a.vector else 'lm%(p)d[i0][i1]') % sub
        store = ['''for (int i0 = 0; i0 < %(e0)d; ++i0) {
  const int cx = s_inv[l * %(nac)d + i0];
  if (cx >= 0)
    for (int i1 = 0; i1 < %(e1)d; ++i1) for (int r = 0; r < %(rd)d; ++r) for (int c = 0; c < %(cd)d; ++c)
      s_val[((i1 * %(rd)d + r) * %(cd)d + c) * cpitch + cx] = %(val)s;
}''' % dict(sub, val=val)]
        out = ['  extern __shared__ double op2_lds[];',
               ind(self._red_init(), 2),
               '''  const int nb = (int)count, bid = op2_bid;
  const int xcd = bid & 7, q = nb >> 3, rr = nb & 7;
  const int b = (int)start + xcd * q + (xcd < rr ? xcd : rr) + (bid >> 3);
  // slots [0, rn) run elements rlo + slot (primary); the rest are listed
  const int rlo = p_rlo[b], rn = p_rn[b];
  const int ex0 = p_exoff[b], nex = rn + p_exoff[b + 1] - ex0;
  const int t0 = p_t0[b], nn = p_t1[b] - t0, kmax = (int)own_k;
  const int pitch = (int)lds_pitch, cpitch = (int)own_cpitch;
  // LDS: per contribution its local-matrix row [%(ent)d][cpitch]; per element
  // the contribution index of each row position [pitch * %(nac)d] (-1: the
  // row is not owned here); the owned rows' CSR offsets
  double* s_val = op2_lds;
  int* s_inv = (int*)(s_val + (op2_i64)%(ent)d * cpitch);
  int* s_rp = s_inv + %(nac)d * pitch;
  (void)kmax;
  const int c0 = p_tp[t0], ncode = p_tp[t0 + nn] - c0;
  for (int i = threadIdx.x; i < nex * %(nac)d; i += %(bs)d) s_inv[i] = -1;
  for (int i = threadIdx.x; i <= nn * %(rd)d; i += %(bs)d) s_rp[i] = rp%(p)d[t0 * %(rd)d + i];
  __syncthreads();
  // the rows' CSR slots of pair pp: last owned node whose first pair <= pp
#define OP2_PAIR_ROWS(pp, sl) { \\
    int lo = 0, hi = nn - 1; \\
    while (lo < hi) { \\
      const int mid = (lo + hi + 1) >> 1; \\
      if (s_rp[mid * %(rd)d] / %(rdcd)d <= (pp)) lo = mid; else hi = mid - 1; \\
    } \\
    const int kcol = (pp) - s_rp[lo * %(rd)d] / %(rdcd)d; \\
    for (int r = 0; r < %(rd)d; ++r) sl[r] = s_rp[lo * %(rd)d + r] + kcol * %(cd)d; }
  const int pp0 = s_rp[0] / %(rdcd)d, pp1 = s_rp[nn * %(rd)d] / %(rdcd)d;
  // this thread's first pair: slots, program range and old values, loaded
  // now so they arrive while the elements run
  int f_sl[%(rd)d], f_q0 = 0, f_q1 = 0;
  double f_acc[%(rd)d][%(cd)d];
  {
    const int pp = pp0 + (int)threadIdx.x < pp1 ? pp0 + (int)threadIdx.x : pp0;
    OP2_PAIR_ROWS(pp, f_sl);
    if (pp0 < pp1) {
      f_q0 = p_poff[pp];
      f_q1 = p_poff[pp + 1];
#pragma unroll
      for (int r = 0; r < %(rd)d; ++r)
#pragma unroll
        for (int c = 0; c < %(cd)d; ++c) f_acc[r][c] = d%(p)d[f_sl[r] + c];
    }
  }
  // a contribution code is slot * %(nac)d + row position: it indexes s_inv
  for (int i = threadIdx.x; i < ncode; i += %(bs)d) s_inv[OP2_CONTRIB(c0 + i)] = i;
  __syncthreads();
  for (int l = threadIdx.x; l < nex; l += %(bs)d) {
    const unsigned code = l < rn ? (unsigned)(rlo + l) : (unsigned)p_exec[ex0 + l - rn];
    const int e = (int)(code & 0x7fffffffu);
    const bool prim = (code >> 31) == 0u;
    (void)prim;''' % sub,
               ind(decls + compute + store, 4),
               '  }',
               '  __syncthreads();',
               '''  // one thread per node pair of the owned rows (pair pp of node n
  // starts at rowptr[n * %(rd)d] / %(rdcd)d); the first pair's lookups and old
  // values were loaded before the elements ran
  for (int pp = pp0 + (int)threadIdx.x, it = 0; pp < pp1; pp += %(bs)d, ++it) {
    int sl[%(rd)d];
    double acc[%(rd)d][%(cd)d];
    int q0, q1;
    if (it == 0) {
      q0 = f_q0;
      q1 = f_q1;
#pragma unroll
      for (int r = 0; r < %(rd)d; ++r) {
        sl[r] = f_sl[r];
#pragma unroll
        for (int c = 0; c < %(cd)d; ++c) acc[r][c] = f_acc[r][c];
      }
    } else {
      OP2_PAIR_ROWS(pp, sl);
      q0 = p_poff[pp];
      q1 = p_poff[pp + 1];
#pragma unroll
      for (int r = 0; r < %(rd)d; ++r)
#pragma unroll
        for (int c = 0; c < %(cd)d; ++c) acc[r][c] = d%(p)d[sl[r] + c];
    }
    // scatter-program codes, %(pmax)d loads at a time
    for (int qb = q0; qb < q1; qb += %(pmax)d) {
      int code[%(pmax)d];
#pragma unroll
      for (int u = 0; u < %(pmax)d; ++u)
        code[u] = qb + u < q1 ? (own_wide ? ((const int*)p_prog)[qb + u]
                                          : (int)((const unsigned short*)p_prog)[qb + u]) : 0;
#pragma unroll
      for (int u = 0; u < %(pmax)d; ++u) {
        if (qb + u < q1) {
          const int cl = code[u] / %(e1)d, j = code[u] - cl * %(e1)d;
#pragma unroll
          for (int r = 0; r < %(rd)d; ++r)
#pragma unroll
            for (int c = 0; c < %(cd)d; ++c)
              acc[r][c] %(op)s s_val[((j * %(rd)d + r) * %(cd)d + c) * cpitch + cl];
        }
      }
    }
#pragma unroll
    for (int r = 0; r < %(rd)d; ++r)
#pragma unroll
      for (int c = 0; c < %(cd)d; ++c) d%(p)d[sl[r] + c] = acc[r][c];
  }''' % dict(sub, pmax=8),
               ind(self._red_finish('b'), 2)]
        return out

    # records a row thread holds in registers per reload
    ROWS_RECORDS = 8

    def _rows_body(self):
        """Owner-computes Mat assembly by row blocks: block b owns the row
        nodes [b*k, (b+1)*k) and runs the elements naming one of them in
        ascending user order (its exec list), a chunk of one element per
        thread at a time, keeping the chunk's local matrices in LDS.  After
        each chunk, one thread per owned node adds the chunk's entries of its
        row into the block's CSR slice, which sits in LDS from the first
        chunk to the last: its records (element slot and row position, and
        the column positions of the element's columns in the row) come in
        ascending (user element, row position) order and each record's
        columns in ascending order, so every entry receives its terms in the
        order of the reference's sequential loop (addto_vector /
        addto_scalar).  Element loads run as a pipeline over the chunks:
        map entries two chunks ahead (from the plan's exec-ordered copy of
        the map rows: coalesced, no element number needed first) and rows
        one ahead, into a second set of registers (the chunk loop is
        unrolled by two, the sets trading roles).  No atomics, no colours;
        the slice is read and written once."""
        bs = self.bs
        ml = lambda slot: self._element_body(split='maps', mslot=slot)[1]
        decls, _, loads, compute = self._element_body(split='maps', mslot='0')
        decls_n, _, loads_n, compute_n = self._element_body(sfx='_n', ev='e_n', split='maps',
                                                            mslot='0')
        decls_n = [l for l in decls_n if not l.startswith('int mi')]
        ahead = ml('ex0 + c0 + 2 * %d + threadIdx.x' % bs)
        ind = lambda lines, n: '\n'.join(' ' * n + l for l in lines)
        if self.reductions:
            raise NotImplementedError("row-block Mat assembly with Global reductions")
        a = [x for x in self.infos if x.kind == 'mat'][0]
        p, bs, t = a.pos, self.bs, a.ctype
        e0, e1 = self.extents[0], self.extents[1]
        rd, cd = a.rdim, a.cdim_m
        W = (e1 + 3) // 4
        sub = {'p': p, 'bs': bs, 't': t, 'e0': e0, 'e1': e1, 'rd': rd, 'cd': cd, 'W': W,
               'ent': e0 * e1 * rd * cd, 'P': self.ROWS_RECORDS,
               'pitch': self.rows_pitch(),
               'op': '=' if a.access is WRITE else '+='}
        head = """  extern __shared__ double op2_lds[];
  // consecutive blocks (neighbouring row ranges) go to one XCD
  const int nb = (int)count, bid = op2_bid;
  const int xcd = bid & 7, q = nb >> 3, rr = nb & 7;
  const int b = (int)start + xcd * q + (xcd < rr ? xcd : rr) + (bid >> 3);
  const int kk = (int)own_k, t0 = b * kk;
  const int t1 = t0 + kk < (int)own_nnodes ? t0 + kk : (int)own_nnodes, nr = t1 - t0;
  const int ex0 = p_xoff[b], nex = p_xoff[b + 1] - ex0;
  const int base = rp%(p)d[(op2_i64)t0 * %(rd)d];
  const int nsl = rp%(p)d[(op2_i64)t1 * %(rd)d] - base;
  // LDS: the block's CSR slice [own_slice], the chunk's local matrices
  // [%(bs)d][%(pitch)d] (an odd pitch: a row thread reads the entries of one
  // local-matrix row together, and the slots its wavefront reads spread
  // over the banks) and the owned scalar rows' offsets in the slice
  // [k * %(rd)d + 1]
  %(t)s* s_acc = (%(t)s*)op2_lds;
  %(t)s* s_val = s_acc + (op2_i64)own_slice;
  int* s_rp = (int*)(s_val + %(pitch)d * %(bs)d);
  // a row thread's records: [rec, rend)
  int rec = 0, rend = 0;
  if ((int)threadIdx.x < nr) {
    rec = p_tp[t0 + threadIdx.x];
    rend = p_tp[t0 + threadIdx.x + 1];
  }
  // the element pipeline's loads take no guard (the exec list and the
  // exec-ordered map rows carry padding naming element 0, whose map
  // entries and rows exist), so they stay straight-line code whose
  // outstanding loads the compiler counts exactly; elements past nex are
  // loaded but never run
  int e = p_exec[ex0 + threadIdx.x];
  int e_n = p_exec[ex0 + %(bs)d + threadIdx.x];""" % sub
        # a record (code cv: slot * e0 + i0; column positions in words qv)
        # added into the row
        apply = """{
  const int sl = CV / %(e0)d, i0 = CV - sl * %(e0)d, cl = sl - c0;
  int kc[%(e1)d];
  %(t)s v[%(e1)d][%(rd)d][%(cd)d];
  bool dup = false;
#pragma unroll
  for (int i1 = 0; i1 < %(e1)d; ++i1) {
    kc[i1] = (int)((QV[i1 >> 2] >> (8 * (i1 & 3))) & 255u);
#pragma unroll
    for (int j1 = 0; j1 < i1; ++j1) dup |= kc[j1] == kc[i1];
#pragma unroll
    for (int r = 0; r < %(rd)d; ++r)
#pragma unroll
      for (int c = 0; c < %(cd)d; ++c)
        v[i1][r][c] = s_val[cl * %(pitch)d + ((i0 * %(e1)d + i1) * %(rd)d + r) * %(cd)d + c];
  }
  if (!dup) {
    // distinct entries: their loads issue together
    %(t)s w[%(e1)d][%(rd)d][%(cd)d];
#pragma unroll
    for (int i1 = 0; i1 < %(e1)d; ++i1)
#pragma unroll
      for (int r = 0; r < %(rd)d; ++r)
#pragma unroll
        for (int c = 0; c < %(cd)d; ++c) w[i1][r][c] = s_acc[rows[r] + kc[i1] * %(cd)d + c];
#pragma unroll
    for (int i1 = 0; i1 < %(e1)d; ++i1)
#pragma unroll
      for (int r = 0; r < %(rd)d; ++r)
#pragma unroll
        for (int c = 0; c < %(cd)d; ++c) {
          w[i1][r][c] %(op)s v[i1][r][c];
          s_acc[rows[r] + kc[i1] * %(cd)d + c] = w[i1][r][c];
        }
  } else {
    // an entry named twice (a node repeated among the columns): one
    // column after the other
#pragma unroll
    for (int i1 = 0; i1 < %(e1)d; ++i1)
#pragma unroll
      for (int r = 0; r < %(rd)d; ++r)
#pragma unroll
        for (int c = 0; c < %(cd)d; ++c)
          s_acc[rows[r] + kc[i1] * %(cd)d + c] %(op)s v[i1][r][c];
  }
}""" % sub
        fast = apply.replace('CV', 'rc[u]').replace('QV', 'rq[u]').split('\n')
        slow = apply.replace('CV', 'cv').replace('QV', 'qv').split('\n')
        fetch = """    // the row thread's next %(P)d records (the record arrays carry padding
    // entries: these loads need no guard and issue together; past rend a
    // code reads as INT_MAX), in flight while the chunk's elements run
    int rc[%(P)d];
    unsigned rq[%(P)d][%(W)d];
#pragma unroll
    for (int u = 0; u < %(P)d; ++u) {
      const int cv = (int)((const unsigned short*)p_contrib)[rec + u];
      rc[u] = rec + u < rend ? cv : 0x7fffffff;
#pragma unroll
      for (int w = 0; w < %(W)d; ++w) rq[u][w] = p_pos[(op2_i64)(rec + u) * %(W)d + w];
    }""" % sub
        rows = """    __syncthreads();
    if ((int)threadIdx.x < nr) {
      int rows[%(rd)d];
#pragma unroll
      for (int r = 0; r < %(rd)d; ++r) rows[r] = s_rp[threadIdx.x * %(rd)d + r];
      // this chunk's records have codes below cend
      const int cend = (c0 + %(bs)d) * %(e0)d;
      int dn = 0;
#pragma unroll
      for (int u = 0; u < %(P)d; ++u) {
        if (dn == u && rc[u] < cend) {""" % sub
        rows2 = """          ++dn;
        }
      }
      rec += dn;
      // more than %(P)d records in this chunk: the rest one at a time
      while (dn == %(P)d && rec < rend) {
        const int cv = (int)((const unsigned short*)p_contrib)[rec];
        if (cv >= cend) break;
        unsigned qv[%(W)d];
#pragma unroll
        for (int w = 0; w < %(W)d; ++w) qv[w] = p_pos[(op2_i64)rec * %(W)d + w];""" % sub
        rows3 = """        ++rec;
      }
    }
    __syncthreads();""" % sub

        def step(ev, ldn, cmp, lm):
            # one chunk: compute from one register set while the next chunk's
            # rows load into the other and the map entries of the one after
            # into mi
            return ['    {', fetch, ind(ldn, 4), ind(ahead, 4),
                    '    if (c0 + (int)threadIdx.x < nex) {', ind(cmp, 6),
                    '#pragma unroll',
                    '      for (int j = 0; j < %(ent)d; ++j)' % sub,
                    '        s_val[threadIdx.x * %d + j] = ((const %s*)%s)[j];'
                    % (sub['pitch'], t, lm),
                    '    }',
                    '    %s = p_exec[ex0 + c0 + 2 * %d + threadIdx.x];' % (ev, bs),
                    rows, ind(fast, 10), rows2, ind(slow, 8), rows3, '    }']
        slice_pre = """  // the first %(U)d * %(bs)d entries of the block's CSR slice: their loads go
  // out first (indices past the end read the last entry again) and their
  // LDS stores wait until the first element loads are issued
  %(t)s sv[%(U)d];
  if (nsl > 0) {
#pragma unroll
    for (int u = 0; u < %(U)d; ++u) {
      const int i = u * %(bs)d + (int)threadIdx.x;
      sv[u] = d%(p)d[(op2_i64)base + (i < nsl ? i : nsl - 1)];
    }
  }""" % dict(sub, U=8)
        slice_in = """#pragma unroll
  for (int u = 0; u < %(U)d; ++u) {
    const int i = u * %(bs)d + (int)threadIdx.x;
    if (i < nsl) s_acc[i] = sv[u];
  }
  // the rest of the slice, %(U)d loads per thread before their stores
  for (int i0 = %(U)d * %(bs)d; i0 < nsl; i0 += %(U)d * %(bs)d) {
#pragma unroll
    for (int u = 0; u < %(U)d; ++u) {
      const int i = i0 + u * %(bs)d + (int)threadIdx.x;
      sv[u] = d%(p)d[(op2_i64)base + (i < nsl ? i : nsl - 1)];
    }
#pragma unroll
    for (int u = 0; u < %(U)d; ++u) {
      const int i = i0 + u * %(bs)d + (int)threadIdx.x;
      if (i < nsl) s_acc[i] = sv[u];
    }
  }
  for (int i = threadIdx.x; i <= nr * %(rd)d; i += %(bs)d)
    s_rp[i] = rp%(p)d[(op2_i64)t0 * %(rd)d + i] - base;""" % dict(sub, U=8)
        slice_out = """  for (int i0 = 0; i0 < nsl; i0 += %(U)d * %(bs)d) {
    %(t)s sv[%(U)d];
#pragma unroll
    for (int u = 0; u < %(U)d; ++u) {
      const int i = i0 + u * %(bs)d + (int)threadIdx.x;
      sv[u] = s_acc[i < nsl ? i : nsl - 1];
    }
#pragma unroll
    for (int u = 0; u < %(U)d; ++u) {
      const int i = i0 + u * %(bs)d + (int)threadIdx.x;
      if (i < nsl) d%(p)d[(op2_i64)base + i] = sv[u];
    }
  }""" % dict(sub, U=8)
        if self.rows_stage:
            # the staged rows go to LDS before the first chunk's rows are read
            out = [head, self._rows_stage_fill(sub), slice_pre, ind(decls, 2), ind(decls_n, 2),
                   ind(ml('ex0 + threadIdx.x'), 2), self._rows_stage_copy(sub), slice_in,
                   '  __syncthreads();', ind(loads, 2),
                   ind(ml('ex0 + %d + threadIdx.x' % bs), 2),
                   '  for (int c0 = 0; c0 < nex; c0 += %(bs)d) {' % sub]
        else:
            out = [head, slice_pre, ind(decls, 2), ind(decls_n, 2),
                   ind(ml('ex0 + threadIdx.x'), 2), ind(loads, 2),
                   ind(ml('ex0 + %d + threadIdx.x' % bs), 2), slice_in,
                   '  for (int c0 = 0; c0 < nex; c0 += %(bs)d) {' % sub]
        out += step('e', loads_n, compute, 'lm%d' % p)
        out += ['    c0 += %(bs)d;' % sub, '    if (c0 >= nex) break;']
        out += step('e_n', loads, compute_n, 'lm%d_n' % p)
        out += ['  }', slice_out]
        return out

    # staged rows one thread loads per pass
    ROWS_STAGE_UNROLL = 4

    def _rows_stage_fill(self, sub):
        """Staged rows template, block start: the LDS regions' pointers and
        the first pass of the staged rows' loads (node numbers, then rows,
        into registers; the LDS stores come in ``_rows_stage_copy``)."""
        regs = self.rows_staged()
        U, bs = self.ROWS_STAGE_UNROLL, self.bs
        lines = ['  // staged rows: per (Dat, map) the rows of the targets the block\'s',
                 '  // elements name, at 16-byte aligned LDS offsets after the slice,',
                 '  // the local matrices and the row offsets (own_stage rows each)',
                 '  char* s_gb = (char*)op2_lds + (((((op2_i64)own_slice + %d * %d) * (op2_i64)sizeof(%s))'
                 ' + 4 * ((op2_i64)kk * %d + 1) + 15) & ~(op2_i64)15);'
                 % (sub['pitch'], bs, sub['t'], sub['rd'])]
        off = '0'
        for p, m, c, t in regs:
            lines.append('  %s* s_g%d = (%s*)(s_gb + %s);' % (t, p, t, off))
            off = '%s + ((own_stage * %d * (op2_i64)sizeof(%s) + 15) & ~(op2_i64)15)' % (off, c, t)
        for m in sorted(set(r[1] for r in regs)):
            lines += ['  const int sg%d0 = so%d[b], sg%dn = so%d[b + 1] - sg%d0;' % (m, m, m, m, m),
                      '  int sgn%d[%d];' % (m, U),
                      '#pragma unroll',
                      '  for (int u = 0; u < %d; ++u) {' % U,
                      '    const int j = u * %d + (int)threadIdx.x;' % bs,
                      '    sgn%d[u] = sg%dn > 0 ? sn%d[sg%d0 + (j < sg%dn ? j : sg%dn - 1)] : 0;'
                      % (m, m, m, m, m, m),
                      '  }']
        for p, m, c, t in regs:
            lines += ['  %s sgv%d[%d][%d];' % (t, p, U, c),
                      '#pragma unroll',
                      '  for (int u = 0; u < %d; ++u)' % U,
                      '#pragma unroll',
                      '    for (int k = 0; k < %d; ++k) sgv%d[u][k] = d%d[(op2_i64)sgn%d[u] * %d + k];'
                      % (c, p, p, m, c)]
        return '\n'.join(lines)

    def _rows_stage_copy(self, sub):
        """Staged rows template: the first pass's rows stored to LDS, then
        the further passes (node numbers, rows, stores)."""
        regs = self.rows_staged()
        U, bs = self.ROWS_STAGE_UNROLL, self.bs
        lines = []

        def store(j0):
            out = []
            for p, m, c, t in regs:
                out += ['#pragma unroll',
                        '  for (int u = 0; u < %d; ++u) {' % U,
                        '    const int j = %s + u * %d + (int)threadIdx.x;' % (j0, bs),
                        '    if (j < sg%dn)' % m,
                        '#pragma unroll',
                        '      for (int k = 0; k < %d; ++k) s_g%d[j * %d + k] = sgv%d[u][k];'
                        % (c, p, c, p),
                        '  }']
            return out
        lines += store('0')
        maps = sorted(set(r[1] for r in regs))
        cond = ' || '.join('j0 < sg%dn' % m for m in maps)
        lines += ['  for (int j0 = %d; %s; j0 += %d) {' % (U * bs, cond, U * bs)]
        for m in maps:
            lines += ['#pragma unroll',
                      '    for (int u = 0; u < %d; ++u) {' % U,
                      '      const int j = j0 + u * %d + (int)threadIdx.x;' % bs,
                      '      sgn%d[u] = sg%dn > 0 ? sn%d[sg%d0 + (j < sg%dn ? j : sg%dn - 1)] : 0;'
                      % (m, m, m, m, m, m),
                      '    }']
        for p, m, c, t in regs:
            lines += ['#pragma unroll',
                      '    for (int u = 0; u < %d; ++u)' % U,
                      '#pragma unroll',
                      '      for (int k = 0; k < %d; ++k) sgv%d[u][k] = d%d[(op2_i64)sgn%d[u] * %d + k];'
                      % (c, p, p, m, c)]
        lines += ['  ' + l for l in store('j0')]
        lines.append('  }')
        return '\n'.join(lines)

    def rows_pitch(self):
        """Values per slot of a row block's local matrices (odd)."""
        a = [x for x in self.infos if x.kind == 'mat'][0]
        ent = self.extents[0] * self.extents[1] * a.rdim * a.cdim_m
        return ent | 1

    def rows_lds_bytes(self, k, slice_max, stage_max=0):
        """LDS bytes of a row block: its CSR slice, the chunk's local
        matrices and its rows' slice offsets; staged, also ``stage_max``
        rows per staged region."""
        a = [x for x in self.infos if x.kind == 'mat'][0]
        isz = np.dtype(a.arg.data.dtype).itemsize
        n = isz * (slice_max + self.rows_pitch() * self.bs) + 4 * (k * a.rdim + 1)
        if not self.rows_stage:
            return n
        r16 = lambda x: (x + 15) & ~15
        return r16(n) + sum(r16(stage_max * c * np.dtype(self.infos[q].arg.data.dtype).itemsize)
                            for q, _, c, _ in self.rows_staged())

    def _stage_offset(self, pos):
        """Values per owned row staged before argument ``pos`` (all of
        them for None)."""
        n = 0
        for q in self.owner_stage:
            if q == pos:
                break
            if self._stage_of[q] == q:
                n += self.infos[q].cdim
        return n

    def owner_lds_bytes(self, k, max_exec, max_codes):
        """LDS bytes of an owner block: its elements' increments (Dat), or
        (Mat) its contributions' local-matrix rows, its elements'
        row-position table and its rows' CSR offsets."""
        a = self.owner_columns()[0][0]
        if a.kind != 'mat':
            return max_exec * self.owner_slot_bytes() + 8 * k * self._stage_offset(None)
        ent = self.extents[1] * a.rdim * a.cdim_m
        return 8 * ent * max(max_codes, 1) + 4 * a.rarity * max_exec + 4 * (k * a.rdim + 1)

    def _params_decl(self):
        decl = ['op2_i64 start', 'op2_i64 count']
        if self.mode == 'coloured':
            decl += ['const int* __restrict__ p_elems', 'const int* __restrict__ p_blkmap',
                     'const int* __restrict__ p_offset', 'const int* __restrict__ p_nelem',
                     'const int* __restrict__ p_nthrcol', 'const int* __restrict__ p_thrcol']
        elif self.mode == 'owner':
            decl += ['const int* __restrict__ p_t0', 'const int* __restrict__ p_t1',
                     'const int* __restrict__ p_rlo', 'const int* __restrict__ p_rn',
                     'const int* __restrict__ p_tgoff', 'const int* __restrict__ p_tgt',
                     'const int* __restrict__ p_tgcb',
                     'const int* __restrict__ p_exoff', 'const int* __restrict__ p_exec',
                     'const int* __restrict__ p_tp', 'const void* __restrict__ p_contrib',
                     'op2_i64 own_c16', 'op2_i64 lds_pitch', 'op2_i64 own_k', 'op2_i64 own_cpitch',
                     'const int* __restrict__ p_poff', 'const void* __restrict__ p_prog',
                     'op2_i64 own_wide']
        elif self.mode == 'rows':
            decl += ['const int* __restrict__ p_tp', 'const int* __restrict__ p_xoff',
                     'const int* __restrict__ p_exec', 'const void* __restrict__ p_contrib',
                     'op2_i64 own_c16', 'const unsigned* __restrict__ p_pos', 'op2_i64 own_k',
                     'op2_i64 own_nnodes', 'op2_i64 own_slice']
            if self.rows_stage:
                decl.append('op2_i64 own_stage')
                for m in self.rows_maps():
                    decl += ['const int* __restrict__ sn%d' % m, 'const int* __restrict__ so%d' % m,
                             'const unsigned short* __restrict__ xl%d' % m]
            else:
                decl += ['const int* __restrict__ xm%d' % m for m in self.rows_maps()]
        for a in self.infos:
            decl += [d for d, _ in a.params()]
        if self.has_mats:
            decl.append('int* __restrict__ err_flag')
        decl += ['op2_i64 pend_n', 'const double* __restrict__ pend_part', 'double* pend_out',
                 'op2_i64 pend_info']
        return decl

    def source(self):
        bs = self.bs
        ind = lambda lines, n: '\n'.join(' ' * n + l for l in lines)
        out = [self.PRELUDE, self._const_decls(), self._row_types(),
               '#pragma clang force_cuda_host_device begin',
               'inline',
               self.kernel.code,
               '#pragma clang force_cuda_host_device end', '']
        bounds = '%d, %d' % (bs, self.min_waves) if self.min_waves else '%d' % bs
        out.append('extern "C" __global__ void __launch_bounds__(%s) %s(\n    %s) {'
                   % (bounds, self.entry, ',\n    '.join(self._params_decl())))
        out.append('''  // with a deferred reduction the launch has one more workgroup, the
  // first (dispatched first, it overlaps the loop's), which runs it; the
  // loop's own workgroups count from the next
  if (pend_n > 0 && blockIdx.x == 0) {
    op2_pending_reduce(pend_n, pend_part, pend_out, pend_info);
    return;
  }
  const int op2_bid = (int)blockIdx.x - (pend_n > 0 ? 1 : 0);''')
        if self.mode == 'rows':
            out += self._rows_body()
            out.append('}')
            return '\n'.join(out)
        decls, compute = self._element_body()
        if self.mode == 'stream':
            out += self._stream_body(decls, compute)
            out.append(ind(self._red_finish('op2_bid'), 2))
        elif self.mode == 'owner':
            out += self._owner_body(decls, compute)
        else:
            out.append(ind(self._red_init(), 2))
            out.append('''  // blocks of one colour are spread over the 8 XCDs in contiguous
  // chunks, so neighbouring plan blocks share an L2
  const int nb = (int)count, bid = op2_bid;
  const int xcd = bid & 7, q = nb >> 3, rr = nb & 7;
  const int lb = xcd * q + (xcd < rr ? xcd : rr) + (bid >> 3);
  const int b = p_blkmap[start + lb];
  const int off = p_offset[b], nel = p_nelem[b], ncol = p_nthrcol[b];
  for (int base = 0; base < nel; base += %d) {
    const int l = base + (int)threadIdx.x;
    const bool active = l < nel;
    const int e = active ? p_elems[off + l] : 0;''' % bs)
            out.append(ind(decls, 4))
            out.append(ind(['if (active) {'], 4))
            out.append(ind(compute, 6))
            out.append(ind(['}'], 4))
            out.append('    const int col = active ? p_thrcol[off + l] : -1;')
            out.append('    for (int c = 0; c < ncol; ++c) {')
            out.append('      if (col == c) {')
            out.append(ind(self._apply_increments(), 8))
            out.append('      }')
            out.append('      __syncthreads();')
            out.append('    }')
            out.append('  }')
            out.append(ind(self._red_finish('b'), 2))
        out.append('}')
        if self.reductions:
            out.append(self._reduce_source())
        return '\n'.join(out)

    def _reduce_source(self):
        """One block of 1024 threads combines the per-block partials: each
        thread loads 4 partials at a time (the loads overlap), then a
        wavefront shuffle tree and the 16 wavefront sums in order."""
        nt, unroll = 1024, 4
        decl = ['op2_i64 nparts']
        # the Globals are not __restrict__: one Global may be reduced into at
        # two positions, each section then reads the one before it wrote
        for a in self.reductions:
            decl += ['%s* g%d' % (a.ctype, a.pos),
                     'const %s* __restrict__ r%d' % (a.ctype, a.pos)]
        lines = ['extern "C" __global__ void __launch_bounds__(%d) %s(%s) {'
                 % (nt, self.reduce_entry, ', '.join(decl))]
        for a in self.reductions:
            op = _mode(a.access)
            expr = _REDUCE_OPS[op][0]
            init = '(%s)0' % a.ctype if op == 'INC' else 'g%d[k]' % a.pos
            lines.append('  {')
            lines.append('    __shared__ %s s%d[%d];' % (a.ctype, a.pos, nt // 64))
            lines.append('    for (int k = 0; k < %d; ++k) {' % a.cdim)
            lines.append('      %s v = %s;' % (a.ctype, init))
            lines.append('      for (op2_i64 q0 = threadIdx.x; q0 < nparts; q0 += %d) {'
                         % (nt * unroll))
            lines.append('        %s x[%d];' % (a.ctype, unroll))
            lines.append('#pragma unroll')
            lines.append('        for (int u = 0; u < %d; ++u) {' % unroll)
            lines.append('          const op2_i64 q = q0 + (op2_i64)u * %d;' % nt)
            lines.append('          x[u] = q < nparts ? r%d[q * %d + k] : (%s)0;'
                         % (a.pos, a.cdim, a.ctype))
            lines.append('        }')
            lines.append('#pragma unroll')
            lines.append('        for (int u = 0; u < %d; ++u)' % unroll)
            lines.append('          if (q0 + (op2_i64)u * %d < nparts) { %s a = v, b = x[u]; v = %s; }'
                         % (nt, a.ctype, expr))
            lines.append('      }')
            lines.append('      for (int off = 32; off > 0; off >>= 1) '
                         '{ %s a = v, b = op2_shfl_xor(v, off); v = %s; }' % (a.ctype, expr))
            lines.append('      if ((threadIdx.x & 63) == 0) s%d[threadIdx.x >> 6] = v;' % a.pos)
            lines.append('      __syncthreads();')
            lines.append('      if (threadIdx.x == 0) {')
            lines.append('        v = s%d[0];' % a.pos)
            lines.append('        for (int w = 1; w < %d; ++w) { %s a = v, b = s%d[w]; v = %s; }'
                         % (nt // 64, a.ctype, a.pos, expr))
            lines.append('        %s a = g%d[k], b = v; g%d[k] = %s;' % (a.ctype, a.pos, a.pos, expr))
            lines.append('      }')
            lines.append('      __syncthreads();')
            lines.append('    }')
            lines.append('  }')
        lines.append('}')
        return '\n'.join(lines)
