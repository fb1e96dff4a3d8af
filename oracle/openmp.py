"""CPU OpenMP baseline: the reference's OpenMP backend restated (TEST
INFRASTRUCTURE, used only as ``bench.py``'s CPU baseline and by its tests).

Follows pyop2/openmp.py:130-261: the iteration set is cut into contiguous
blocks of ``part_size`` elements (1024, as the reference's default), the
blocks are coloured so that no two blocks of one colour share a target of
an INC Dat or a Mat row (greedy 32-colour rounds, pyop2/plan.pyx:214-376,
restated as ``oracle_plan_colour``), and each colour's blocks run in a
``#pragma omp parallel for``; a block runs its elements in order with the
sequential wrapper.  Global reductions accumulate into per-thread copies
that are combined after each colour, in thread order (pyop2/openmp.py:84-115
keeps per-thread scratch likewise).  The reference's generated code runs
serially unless ``debug > 0`` (host.py:249); this restatement runs the
intended parallel loop.

The data classes are the sequential oracle's; only ``par_loop`` differs.
"""

import ctypes
import hashlib
import os
import subprocess

import numpy as np

from pyop2_amd import base
from pyop2_amd.base import INC, MIN, MAX, READ
from pyop2_amd.utils import as_tuple, maybe_setflags

from . import sequential as seq
from .sequential import (Set, Map, Dat, Global, Const, Kernel, Sparsity, Mat, Solver,  # noqa
                         IterationSpace, Halo, Arg, i, IdentityMap, solve)

PART_SIZE = 1024
CFLAGS = seq.CFLAGS + ['-fopenmp']


def generate(kernel, extents, args, consts):
    """The sequential wrapper plus ``wrap_omp`` running coloured blocks."""
    src = seq.generate(kernel, extents, args, consts)
    wargs, pargs, red = [], [], []
    for k, a in enumerate(args):
        name, mname = 'arg%d' % k, 'arg%d_map' % k
        if a._is_mat:
            wargs.append('void* %s, int* %s, int* %s2' % (name, mname, mname))
            pargs += [name, mname, mname + '2']
        else:
            wargs.append('void* %s' % name + (', int* %s' % mname if a._is_indirect else ''))
            if a._is_global_reduction:
                red.append((k, a))
                pargs.append('(void*)%s_loc' % name)
            else:
                pargs.append(name)
            if a._is_indirect:
                pargs.append(mname)
    cargs = ['void* c_%s' % c.name for c in consts]
    pargs += ['c_%s' % c.name for c in consts]
    body = ['void wrap_omp(int n, int part, int ncol, const int* col_off, const int* col_blk%s%s) {'
            % (''.join(', ' + w for w in wargs), ''.join(', ' + c for c in cargs))]
    body.append('  for (int col = 0; col < ncol; ++col) {')
    body.append('    #pragma omp parallel')
    body.append('    {')
    for k, a in red:
        t, d = a.data.ctype, a.data.cdim
        op = a.access
        body.append('      %s arg%d_loc[%d];' % (t, k, d))
        init = '0' if op is INC else '((%s*)arg%d)[j]' % (t, k)
        body.append('      for (int j = 0; j < %d; ++j) arg%d_loc[j] = %s;' % (d, k, init))
    body.append('      #pragma omp for schedule(static)')
    body.append('      for (int q = col_off[col]; q < col_off[col + 1]; ++q) {')
    body.append('        int b = col_blk[q];')
    body.append('        int e1 = (b + 1) * part < n ? (b + 1) * part : n;')
    body.append('        wrap(b * part, e1, %s);' % ', '.join(pargs))
    body.append('      }')
    for k, a in red:
        t, d = a.data.ctype, a.data.cdim
        body.append('      #pragma omp for ordered schedule(static, 1)')
        body.append('      for (int th = 0; th < omp_get_num_threads(); ++th) {')
        body.append('        #pragma omp ordered')
        body.append('        for (int j = 0; j < %d; ++j) {' % d)
        g = '((%s*)arg%d)[j]' % (t, k)
        v = 'arg%d_loc[j]' % k
        if a.access is INC:
            body.append('          %s += %s;' % (g, v))
        elif a.access is MIN:
            body.append('          if (%s < %s) %s = %s;' % (v, g, g, v))
        else:
            body.append('          if (%s > %s) %s = %s;' % (v, g, g, v))
        body.append('        }')
        body.append('      }')
    body.append('    }')
    body.append('  }')
    body.append('}')
    return '#include <omp.h>\n' + src + '\n' + '\n'.join(body)


_compiled = {}


def compile_wrapper(src):
    h = hashlib.md5((src + ' '.join(CFLAGS)).encode()).hexdigest()
    f = _compiled.get(h)
    if f is not None:
        return f
    os.makedirs(seq.CACHE, exist_ok=True)
    so = os.path.join(seq.CACHE, 'omp_%s.so' % h)
    if not os.path.exists(so):
        cpath = os.path.join(seq.CACHE, 'omp_%s.c' % h)
        with open(cpath, 'w') as fh:
            fh.write(src)
        tmp = so + '.%d.tmp' % os.getpid()
        subprocess.check_call(['gcc'] + CFLAGS + ['-I', os.path.join(seq.HERE, 'csrc'), cpath,
                                                  '-o', tmp, '-lm'])
        os.replace(tmp, so)
    lib = ctypes.CDLL(so)
    f = lib.wrap_omp
    f.restype = None
    _compiled[h] = (f, lib)
    return _compiled[h]


_plans = {}


def plan(iterset, n, args, part=PART_SIZE):
    """(ncol, col_off, col_blk): blocks of ``part`` elements grouped by
    colour; race targets are INC Dats' maps and Mat row maps."""
    race = []
    for a in args:
        if a._is_mat:
            race.append(a.map[0])
        elif a._is_indirect and a.access is INC:
            race += list(as_tuple(a.map, base.Map))
    key = (id(iterset), n, part, tuple(id(m) for m in race))
    p = _plans.get(key)
    if p is not None:
        return p
    nblocks = (n + part - 1) // part
    if race:
        keep = [np.ascontiguousarray(m.values[:n], dtype=np.int32) for m in race]
        ptrs = (ctypes.c_void_p * len(keep))(*[k.ctypes.data for k in keep])
        dims = (ctypes.c_int * len(keep))(*[m.dim for m in race])
        sizes = (ctypes.c_int * len(keep))(*[m.dataset.total_size for m in race])
        blkcol = np.zeros(nblocks, dtype=np.int32)
        f = seq.rt().oracle_plan_colour
        f.restype = ctypes.c_int
        ncol = f(n, part, len(keep), ptrs, dims, sizes,
                 blkcol.ctypes.data_as(ctypes.c_void_p))
    else:
        blkcol = np.zeros(nblocks, dtype=np.int32)
        ncol = 1 if nblocks else 0
    order = np.argsort(blkcol, kind='stable').astype(np.int32)
    off = np.searchsorted(blkcol[order], np.arange(ncol + 1)).astype(np.int32)
    p = (int(ncol), off, order)
    _plans[key] = p
    return p


class ParLoop(base.ParLoop):

    def compute(self):
        consts = base.Const._definitions()
        src = generate(self.kernel, self._it_space.extents, self.args, consts)
        fun, _lib = compile_wrapper(src)
        n = self._it_space.exec_size if self.needs_exec_halo else self._it_space.size
        ncol, off, blk = plan(self._it_space.iterset, n, self.args)
        cargs, keep = [], [off, blk]
        for a in self.args:
            if a._is_mat:
                cargs.append(ctypes.addressof(a.data._csr))
            else:
                arr = a.data._data
                if a._is_dat:
                    arr = a.data.data_ro if a.data.dataset.total_size else arr
                cargs.append(arr.ctypes.data)
            if a._is_indirect or a._is_mat:
                for m in as_tuple(a.map, base.Map):
                    v = np.ascontiguousarray(m.values, dtype=np.int32)
                    keep.append(v)
                    cargs.append(v.ctypes.data)
        for c in consts:
            v = np.ascontiguousarray(c.data)
            keep.append(v)
            cargs.append(v.ctypes.data)
        fun(ctypes.c_int(n), ctypes.c_int(PART_SIZE), ctypes.c_int(ncol),
            ctypes.c_void_p(off.ctypes.data), ctypes.c_void_p(blk.ctypes.data),
            *[ctypes.c_void_p(x) for x in cargs])
        for a in self.args:
            if a._is_dat:
                maybe_setflags(a.data._data, write=False)
        self.reduction_begin()
        self.reduction_end()
        self.maybe_set_halo_update_needed()
        for a in self.args:
            if a._is_mat:
                a.data._assemble()


def par_loop(kernel, it_space, *args):
    ParLoop(kernel, it_space, *args).compute()


def threads():
    """OpenMP threads the baseline runs on."""
    lib = ctypes.CDLL('libgomp.so.1')
    lib.omp_get_max_threads.restype = ctypes.c_int
    return lib.omp_get_max_threads()
