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

import numpy as np

from .base import IterationIndex, READ, WRITE, RW, INC, MIN, MAX


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
    return {id(READ): 'READ', id(WRITE): 'WRITE', id(RW): 'RW', id(INC): 'INC',
            id(MIN): 'MIN', id(MAX): 'MAX'}[id(access)]


class LoopSource(object):
    """HIP source of one loop kernel plus its reduction pass."""

    PRELUDE = r'''
#define OP2_STRIDE(a, idx) a[idx]
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
'''

    def __init__(self, kernel, extents, args, consts, mode, block_size):
        assert mode in ('stream', 'coloured', 'owner')
        self.kernel = kernel
        self.extents = tuple(extents)
        self.infos = [ArgInfo(k, a, self.extents) for k, a in enumerate(args)]
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
            roles += [('own_tgoff', None), ('own_tgt', None), ('own_exoff', None),
                      ('own_exec', None), ('own_tp', None), ('own_contrib', None),
                      ('own_pitch', None)]
        for a in self.infos:
            for _, role in a.params():
                roles.append((role, a.pos))
        if self.has_mats:
            roles.append(('err_flag', None))
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

    def _element_body(self):
        """Code that runs the kernel for element ``e`` (all itspace points),
        leaving INC increments and Mat entries in private arrays.

        Returns (decls, compute): ``decls`` declares and zeroes the private
        arrays and touches no memory, so it can run for idle threads of a
        workgroup; ``compute`` reads maps and data and calls the kernel."""
        decls, compute, call_args, zeros = [], [], [], []
        ext = self.extents
        for a in self.infos:
            p, t = a.pos, a.ctype
            if a.kind == 'gbl_read':
                call_args.append('g%d' % p)
            elif a.kind == 'gbl_red':
                if self.mode == 'owner':
                    decls.append('%s junk%d[%d];' % (t, p, a.cdim))
                    decls.append('for (int k = 0; k < %d; ++k) junk%d[k] = red%d[k];'
                                 % (a.cdim, p, p))
                    call_args.append('(prim ? red%d : junk%d)' % (p, p))
                else:
                    call_args.append('red%d' % p)
            elif a.kind == 'direct':
                if self.mode == 'stream' and self._staged(a):
                    call_args.append('st%d + threadIdx.x * %d' % (p, self._pitch(a)))
                else:
                    call_args.append('d%d + (op2_i64)e * %d' % (p, a.cdim))
            elif a.kind == 'ind':
                if a.race:
                    decls.append('%s inc%d[%d];' % (t, p, a.cdim))
                    decls.append('for (int k = 0; k < %d; ++k) inc%d[k] = (%s)0;' % (a.cdim, p, t))
                    call_args.append('inc%d' % p)
                else:
                    call_args.append('d%d + (op2_i64)m%d[(op2_i64)e * %d + %d] * %d'
                                     % (p, p, a.mdim, a.idx, a.cdim))
            elif a.kind == 'vec':
                decls.append('%s* v%d[%d];' % (t, p, a.mdim))
                if a.race:
                    decls.append('%s inc%d[%d][%d];' % (t, p, a.mdim, a.cdim))
                    decls.append('for (int j = 0; j < %d; ++j) { v%d[j] = inc%d[j]; '
                                 'for (int k = 0; k < %d; ++k) inc%d[j][k] = (%s)0; }'
                                 % (a.mdim, p, p, a.cdim, p, t))
                else:
                    compute.append('for (int j = 0; j < %d; ++j) v%d[j] = d%d + '
                                   '(op2_i64)m%d[(op2_i64)e * %d + j] * %d;'
                                   % (a.mdim, p, p, p, a.mdim, a.cdim))
                call_args.append('v%d' % p)
            elif a.kind == 'itspace':
                ivar = 'i_%d' % a.iidx
                if a.race:
                    decls.append('%s inc%d[%d][%d];' % (t, p, a.extent, a.cdim))
                    decls.append('for (int j = 0; j < %d; ++j) for (int k = 0; k < %d; ++k) '
                                 'inc%d[j][k] = (%s)0;' % (a.extent, a.cdim, p, t))
                    call_args.append('inc%d[%s]' % (p, ivar))
                else:
                    call_args.append('d%d + (op2_i64)m%d[(op2_i64)e * %d + %s] * %d'
                                     % (p, p, a.mdim, ivar, a.cdim))
            elif a.kind == 'mat':
                e0 = ext[0] if len(ext) > 0 else 1
                e1 = ext[1] if len(ext) > 1 else 1
                if a.vector:
                    decls.append('%s lm%d[%d][%d][%d][%d];' % (t, p, e0, e1, a.rdim, a.cdim_m))
                    zeros.append('for (int r = 0; r < %d; ++r) for (int c = 0; c < %d; ++c) '
                                 'lm%d[i_0][i_1][r][c] = (%s)0;' % (a.rdim, a.cdim_m, p, t))
                    call_args.append('lm%d[i_0][i_1]' % p)
                else:
                    decls.append('%s lm%d[%d][%d];' % (t, p, e0, e1))
                    zeros.append('lm%d[i_0][i_1] = (%s)0;' % (p, t))
                    call_args.append('(%s (*)[1])&lm%d[i_0][i_1]' % (t, p))
        call_args += ['i_%d' % k for k in range(len(ext))]
        opens = ['for (int i_%d = 0; i_%d < %d; ++i_%d) {' % (k, k, e, k)
                 for k, e in enumerate(ext)]
        mats = [a for a in self.infos if a.kind == 'mat']
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
        compute.append('%s(%s);' % (self.kernel.name, ', '.join(call_args)))
        compute += ['}'] * len(opens)
        compute += closing
        return decls, compute

    def _apply_increments(self):
        """Apply private INC increments and Mat entries of element ``e``."""
        lines = []
        ext = self.extents
        for a in self.infos:
            p = a.pos
            if not a.race:
                continue
            if a.kind == 'ind':
                lines.append('{ %s* t = d%d + (op2_i64)m%d[(op2_i64)e * %d + %d] * %d;'
                             % (a.ctype, p, p, a.mdim, a.idx, a.cdim))
                lines.append('  for (int k = 0; k < %d; ++k) t[k] += inc%d[k]; }' % (a.cdim, p))
            elif a.kind == 'vec':
                lines.append('for (int j = 0; j < %d; ++j) { %s* t = d%d + '
                             '(op2_i64)m%d[(op2_i64)e * %d + j] * %d;'
                             % (a.mdim, a.ctype, p, p, a.mdim, a.cdim))
                lines.append('  for (int k = 0; k < %d; ++k) t[k] += inc%d[j][k]; }' % (a.cdim, p))
            elif a.kind == 'itspace':
                lines.append('for (int j = 0; j < %d; ++j) { %s* t = d%d + '
                             '(op2_i64)m%d[(op2_i64)e * %d + j] * %d;'
                             % (a.extent, a.ctype, p, p, a.mdim, a.cdim))
                lines.append('  for (int k = 0; k < %d; ++k) t[k] += inc%d[j][k]; }' % (a.cdim, p))
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

    @staticmethod
    def _staged(a):
        """Direct Dats with several components go through LDS in the
        stream template, so a workgroup's loads and stores of its rows are
        contiguous instead of one strided row per thread."""
        return a.kind == 'direct' and a.cdim > 1

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
        out.append('  const op2_i64 stride = (op2_i64)gridDim.x * %d;' % bs)
        out.append('  for (op2_i64 base = (op2_i64)blockIdx.x * %d; base < count; base += stride) {'
                   % bs)
        out.append('    const op2_i64 e = start + base + threadIdx.x;')
        out.append('    const bool active = base + threadIdx.x < count;')
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
        out = ['  extern __shared__ %s op2_lds[];' % t,
               ind(self._red_init(), 2),
               '''  // consecutive owner blocks (neighbouring target ranges) go to one XCD
  const int nb = (int)count, bid = blockIdx.x;
  const int xcd = bid & 7, q = nb >> 3, rr = nb & 7;
  const int b = (int)start + xcd * q + (xcd < rr ? xcd : rr) + (bid >> 3);
  const int ex0 = p_exoff[b], nex = p_exoff[b + 1] - ex0;
  const int pitch = (int)lds_pitch;
  for (int l = threadIdx.x; l < nex; l += %d) {
    const unsigned code = (unsigned)p_exec[ex0 + l];
    const int e = (int)(code & 0x7fffffffu);
    const bool prim = (code >> 31) == 0u;
    (void)prim;''' % bs,
               ind(decls + compute + store, 4),
               '  }',
               '  __syncthreads();',
               '''  // each owned target that receives increments: its old value plus its
  // increments in the order of the plan (ascending user element)
  const int g0 = p_tgoff[b], ng = p_tgoff[b + 1] - g0;
  for (int j = threadIdx.x; j < ng * %(cdim)d; j += %(bs)d) {
    const int t = p_tgt[g0 + j / %(cdim)d], k = j - (j / %(cdim)d) * %(cdim)d;
    const int cb = p_tp[t], ce = p_tp[t + 1];
    %(t)s acc = d%(pos)d[(op2_i64)t * %(cdim)d + k];
    for (int c = cb; c < ce; ++c) {
      const int code = p_contrib[c];
      const int slot = code / %(nac)d, ac = code - slot * %(nac)d;
      acc += op2_lds[(ac * %(cdim)d + k) * pitch + slot];
    }
    d%(pos)d[(op2_i64)t * %(cdim)d + k] = acc;
  }''' % {'cdim': cdim, 'bs': bs, 't': t, 'pos': tgt.pos, 'nac': nac},
               ind(self._red_finish('b'), 2)]
        return out

    def _owner_mat_body(self, decls, compute):
        """Owner-computes Mat assembly: a block owns the rows of a range of
        row-set nodes.  Phase 1 runs the elements naming its nodes and keeps
        their local matrices (and column-map rows) in LDS; phase 2 gives each
        owned node's CSR slots to a wavefront, one slot per lane, and sums the
        node's element contributions to that slot in ascending user element
        order (the order addto_vector/addto_scalar reach it sequentially)."""
        ind = lambda lines, n: '\n'.join(' ' * n + l for l in lines)
        bs = self.bs
        a = self.owner_columns()[0][0]
        p = a.pos
        e0, e1 = self.extents[0], self.extents[1]
        rd, cd = a.rdim, a.cdim_m
        nent = e0 * e1 * rd * cd
        insert = a.access is WRITE
        store = []
        if a.vector:
            store.append('for (int i0 = 0; i0 < %d; ++i0) for (int i1 = 0; i1 < %d; ++i1) '
                         'for (int r = 0; r < %d; ++r) for (int c = 0; c < %d; ++c) '
                         'op2_lds[(((i0 * %d + i1) * %d + r) * %d + c) * pitch + l] = '
                         'lm%d[i0][i1][r][c];' % (e0, e1, rd, cd, e1, rd, cd, p))
        else:
            store.append('for (int i0 = 0; i0 < %d; ++i0) for (int i1 = 0; i1 < %d; ++i1) '
                         'op2_lds[(i0 * %d + i1) * pitch + l] = lm%d[i0][i1];'
                         % (e0, e1, e1, p))
        store.append('for (int j = 0; j < %d; ++j) op2_cmap[j * pitch + l] = '
                     'mc%d[(op2_i64)e * %d + j];' % (a.carity, p, a.carity))
        op = '=' if insert else '+='
        out = ['  extern __shared__ double op2_lds[];',
               ind(self._red_init(), 2),
               '''  const int nb = (int)count, bid = blockIdx.x;
  const int xcd = bid & 7, q = nb >> 3, rr = nb & 7;
  const int b = (int)start + xcd * q + (xcd < rr ? xcd : rr) + (bid >> 3);
  const int ex0 = p_exoff[b], nex = p_exoff[b + 1] - ex0;
  const int pitch = (int)lds_pitch;
  int* op2_cmap = (int*)(op2_lds + (op2_i64)%(nent)d * pitch);
  for (int l = threadIdx.x; l < nex; l += %(bs)d) {
    const unsigned code = (unsigned)p_exec[ex0 + l];
    const int e = (int)(code & 0x7fffffffu);
    const bool prim = (code >> 31) == 0u;
    (void)prim;''' % {'nent': nent, 'bs': bs},
               ind(decls + compute + store, 4),
               '  }',
               '  __syncthreads();',
               '''  // one wavefront per owned node; lanes over the CSR slots of its rows
  const int g0 = p_tgoff[b], ng = p_tgoff[b + 1] - g0;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int g = wv; g < ng; g += %(nw)d) {
    const int node = p_tgt[g0 + g];
    const int cb = p_tp[node], ce = p_tp[node + 1];
    const int s0 = rp%(p)d[node * %(rd)d], s1 = rp%(p)d[node * %(rd)d + %(rd)d];
    for (int sl = s0 + lane; sl < s1; sl += 64) {
      int row = node * %(rd)d;
      while (row + 1 < node * %(rd)d + %(rd)d && rp%(p)d[row + 1] <= sl) ++row;
      const int rr2 = row - node * %(rd)d;
      const int col = ci%(p)d[sl];
      const int m = col / %(cd)d, cc = col - m * %(cd)d;
      double acc = d%(p)d[sl];
      for (int c = cb; c < ce; ++c) {
        const int code = p_contrib[c];
        const int slot = code / %(nac)d, i0 = code - slot * %(nac)d;
        for (int j = 0; j < %(e1)d; ++j) {
          if (op2_cmap[j * pitch + slot] == m)
            acc %(op)s op2_lds[(((i0 * %(e1)d + j) * %(rd)d + rr2) * %(cd)d + cc) * pitch + slot];
        }
      }
      d%(p)d[sl] = acc;
    }
  }''' % {'nw': bs // 64, 'p': p, 'rd': rd, 'cd': cd, 'nac': a.rarity, 'e1': e1, 'op': op},
               ind(self._red_finish('b'), 2)]
        return out

    def _params_decl(self):
        decl = ['op2_i64 start', 'op2_i64 count']
        if self.mode == 'coloured':
            decl += ['const int* __restrict__ p_elems', 'const int* __restrict__ p_blkmap',
                     'const int* __restrict__ p_offset', 'const int* __restrict__ p_nelem',
                     'const int* __restrict__ p_nthrcol', 'const int* __restrict__ p_thrcol']
        elif self.mode == 'owner':
            decl += ['const int* __restrict__ p_tgoff', 'const int* __restrict__ p_tgt',
                     'const int* __restrict__ p_exoff', 'const int* __restrict__ p_exec',
                     'const int* __restrict__ p_tp', 'const int* __restrict__ p_contrib',
                     'op2_i64 lds_pitch']
        for a in self.infos:
            decl += [d for d, _ in a.params()]
        if self.has_mats:
            decl.append('int* __restrict__ err_flag')
        return decl

    def source(self):
        bs = self.bs
        ind = lambda lines, n: '\n'.join(' ' * n + l for l in lines)
        out = [self.PRELUDE, self._const_decls(),
               '#pragma clang force_cuda_host_device begin',
               'inline',
               self.kernel.code,
               '#pragma clang force_cuda_host_device end', '']
        out.append('extern "C" __global__ void __launch_bounds__(%d) %s(\n    %s) {'
                   % (bs, self.entry, ',\n    '.join(self._params_decl())))
        decls, compute = self._element_body()
        if self.mode == 'stream':
            out += self._stream_body(decls, compute)
            out.append(ind(self._red_finish('blockIdx.x'), 2))
        elif self.mode == 'owner':
            out += self._owner_body(decls, compute)
        else:
            out.append(ind(self._red_init(), 2))
            out.append('''  // blocks of one colour are spread over the 8 XCDs in contiguous
  // chunks, so neighbouring plan blocks share an L2
  const int nb = (int)count, bid = blockIdx.x;
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
        decl = ['op2_i64 nparts']
        for a in self.reductions:
            decl += ['%s* __restrict__ g%d' % (a.ctype, a.pos),
                     'const %s* __restrict__ r%d' % (a.ctype, a.pos)]
        lines = ['extern "C" __global__ void __launch_bounds__(256) %s(%s) {'
                 % (self.reduce_entry, ', '.join(decl))]
        for a in self.reductions:
            op = _mode(a.access)
            expr = _REDUCE_OPS[op][0]
            ident = '(%s)0' % a.ctype if op == 'INC' else None
            lines.append('  {')
            lines.append('    __shared__ %s s%d[4];' % (a.ctype, a.pos))
            lines.append('    for (int k = 0; k < %d; ++k) {' % a.cdim)
            init = ident if ident else 'g%d[k]' % a.pos
            lines.append('      %s v = %s;' % (a.ctype, init))
            lines.append('      for (op2_i64 q = threadIdx.x; q < nparts; q += 256) '
                         '{ %s a = v, b = r%d[q * %d + k]; v = %s; }' % (a.ctype, a.pos, a.cdim, expr))
            lines.append('      for (int off = 32; off > 0; off >>= 1) '
                         '{ %s a = v, b = op2_shfl_xor(v, off); v = %s; }' % (a.ctype, expr))
            lines.append('      if ((threadIdx.x & 63) == 0) s%d[threadIdx.x >> 6] = v;' % a.pos)
            lines.append('      __syncthreads();')
            lines.append('      if (threadIdx.x == 0) {')
            lines.append('        %s a = s%d[0], b = s%d[1]; v = %s;' % (a.ctype, a.pos, a.pos, expr))
            lines.append('        a = v; b = s%d[2]; v = %s; a = v; b = s%d[3]; v = %s;'
                         % (a.pos, expr, a.pos, expr))
            lines.append('        a = g%d[k]; b = v; g%d[k] = %s;' % (a.pos, a.pos, expr))
            lines.append('      }')
            lines.append('      __syncthreads();')
            lines.append('    }')
            lines.append('  }')
        lines.append('}')
        return '\n'.join(lines)
