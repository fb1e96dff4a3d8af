"""CPU oracle: the reference's sequential backend restated (TEST INFRASTRUCTURE).

Only the tests, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline
leg use this module, as the checker of the hip backend; the product never
imports it.

Each ``par_loop`` is lowered to a C wrapper following the reference's code
generation (pyop2/sequential.py:55-72, pyop2/host.py:45-322):

    void wrap(int start, int end, <data, map[, map2]>..., <consts>...)
      for (i = start; i < end; i++) {
        <vector-map pointer arrays>
        for (i_0 ...) for (i_1 ...) {
          <zero local tensor>; kernel(args..., i_0, i_1); <addto_scalar (vector Mat)>
        }
        <addto_vector (scalar Mat)>
      }

compiled by gcc with ``-O2 -ffp-contract=off`` (plain IEEE double) and
called through ctypes on the numpy arrays.  Globals are passed in place and
accumulate in element order, Dats are passed in place (INC adds straight
into the target), Mats are CSR arrays filled by the restated
``addto_scalar``/``addto_vector`` of ``csrc/oracle_mat.h``.

The module is also a backend: registering it in
``pyop2_amd.backends.backends['sequential']`` lets ``op2.init(backend=
'sequential')`` run the op2 API on it, which is how the CPU unit tests run.
Its classes can also be used directly (``sequential.Dat(...)``,
``sequential.par_loop(...)``) next to an initialised hip backend, which is
how the GPU parity tests compute expected values from the same inputs.
"""

import ctypes
import hashlib
import os
import subprocess
import tempfile

import numpy as np

from pyop2_amd import base
from pyop2_amd.base import (READ, WRITE, RW, INC, MIN, MAX, IdentityMap, IterationIndex,
                            IterationSpace, Halo, Arg, i)
from pyop2_amd.exceptions import MatTypeError, DatTypeError
from pyop2_amd.mpi import MPI
from pyop2_amd.utils import as_tuple, maybe_setflags, validate_type

HERE = os.path.dirname(os.path.abspath(__file__))
RT_LIB = os.path.join(HERE, 'lib', 'liboracle_rt.so')
CACHE = os.environ.get('ORACLE_CACHE_DIR',
                       os.path.join(tempfile.gettempdir(), 'pyop2_amd_oracle_%d' % os.getuid()))
CFLAGS = ['-O2', '-ffp-contract=off', '-fPIC', '-shared', '-std=gnu99', '-w']


def build_runtime():
    """Build liboracle_rt.so with the oracle's Makefile."""
    subprocess.check_call(['make', '-s', '-C', HERE])


def _rt():
    if not os.path.exists(RT_LIB):
        build_runtime()
    lib = ctypes.CDLL(RT_LIB)
    return lib


_rtlib = None


def rt():
    global _rtlib
    if _rtlib is None:
        _rtlib = _rt()
    return _rtlib


class _OracleMap(ctypes.Structure):
    _fields_ = [('values', ctypes.c_void_p), ('dim', ctypes.c_int), ('from_size', ctypes.c_int)]


class _OracleCSR(ctypes.Structure):
    _fields_ = [('rowptr', ctypes.c_void_p), ('colidx', ctypes.c_void_p),
                ('vals', ctypes.c_void_p), ('nrows', ctypes.c_int), ('err', ctypes.c_int)]


# -- classes -----------------------------------------------------------------

class Set(base.Set):
    pass


class Map(base.Map):
    _arg_class = Arg


class Dat(base.Dat):
    _arg_class = Arg

    def zero(self):
        k = Kernel("void zero(%(t)s *dat) { for (int n = 0; n < %(d)d; ++n) dat[n] = (%(t)s)0; }"
                   % {'t': self.ctype, 'd': self.cdim}, 'zero')
        ParLoop(k, self.dataset, Arg(self, IdentityMap, None, WRITE)).compute()


class Global(base.Global):
    _arg_class = Arg


class Const(base.Const):
    pass


class Kernel(base.Kernel):
    _cache = {}


class Sparsity(base.Sparsity):
    """CSR by the std::set restatement of build_sparsity_pattern_seq."""

    _cache = {}

    def _build(self):
        rmult, cmult = self._dims
        lsize = self._nrows * rmult
        keep = [np.ascontiguousarray(m.values, dtype=np.int32) for m in self._rmaps + self._cmaps]
        nm = self._nmaps
        rms = (_OracleMap * nm)(*[_OracleMap(keep[k].ctypes.data, m.dim, m.iterset.size)
                                  for k, m in enumerate(self._rmaps)])
        cms = (_OracleMap * nm)(*[_OracleMap(keep[nm + k].ctypes.data, m.dim, m.iterset.size)
                                  for k, m in enumerate(self._cmaps)])
        nnz = np.zeros(lsize, dtype=np.int32)
        rowptr = np.zeros(lsize + 1, dtype=np.int32)
        f = rt().oracle_sparsity_seq
        f.restype = ctypes.c_longlong
        f.argtypes = [ctypes.c_int] * 4 + [ctypes.c_void_p] * 5
        total = f(rmult, cmult, self._nrows, nm, ctypes.addressof(rms), ctypes.addressof(cms),
                  nnz.ctypes.data, rowptr.ctypes.data, None)
        colidx = np.zeros(max(total, 1), dtype=np.int32)
        f(rmult, cmult, self._nrows, nm, ctypes.addressof(rms), ctypes.addressof(cms),
          nnz.ctypes.data, rowptr.ctypes.data, colidx.ctypes.data)
        self._rowptr = rowptr
        self._colidx = colidx[:total]
        self._d_nnz = nnz
        self._d_nz = int(total)
        self._o_nnz = np.zeros(lsize, dtype=np.int32)
        self._o_nz = 0


class Mat(base.Mat):
    _arg_class = Arg

    def __init__(self, *args, **kwargs):
        base.Mat.__init__(self, *args, **kwargs)
        self._array = np.zeros(self._sparsity.nz, dtype=np.float64)
        self._csr = _OracleCSR(self._sparsity._rowptr.ctypes.data,
                               self._sparsity._colidx.ctypes.data, self._array.ctypes.data,
                               self._sparsity.nrows * self.dims[0], 0)

    def zero(self):
        self._array[:] = 0.0

    def zero_rows(self, rows, diag_val):
        r = np.ascontiguousarray(rows, dtype=np.int32)
        f = rt().oracle_zero_rows
        f.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_double]
        f(ctypes.addressof(self._csr), r.ctypes.data, int(r.size), float(diag_val))

    def _assemble(self):
        if self._csr.err:
            self._csr.err = 0
            raise RuntimeError("Mat %s: an element added an entry outside its sparsity pattern"
                               % self.name)

    @property
    def array(self):
        return self._array

    @property
    def values(self):
        sp = self._sparsity
        n, m = sp.nrows * self.dims[0], sp.ncols * self.dims[1]
        out = np.zeros((n, m))
        for r in range(n):
            out[r, sp._colidx[sp._rowptr[r]:sp._rowptr[r + 1]]] = \
                self._array[sp._rowptr[r]:sp._rowptr[r + 1]]
        return out

    @property
    def handle(self):
        return self


KRYLOV_NAMES = {'cg': 'pcg', 'gmres': 'gmres', 'bcgs': 'bcgs', 'bicgstab': 'bcgs'}


class Solver(base.Solver):
    """PETSc KSPCG, KSPGMRES and KSPBCGS with PCJACOBI restated (oracle_pcg,
    oracle_gmres, oracle_bcgs); linear_solver 'cg', 'gmres', or 'bcgs' /
    'bicgstab' (the reference's cusp name, pyop2/cuda.py:504-508)."""

    def solve(self, A, x, b):
        p = self.parameters
        method = KRYLOV_NAMES.get(p['linear_solver'])
        assert method is not None and p['preconditioner'] in ('jacobi', 'none')
        sp = A.sparsity
        n = sp.nrows * A.dims[0]
        bb = np.ascontiguousarray(b.data_ro, dtype=np.float64).reshape(-1)
        xx = np.zeros(n)
        maxit = int(p['maximum_iterations'])
        hist = np.zeros(maxit + 2)
        its = ctypes.c_int()
        rn = ctypes.c_double()
        tol = [float(p['relative_tolerance']), float(p['absolute_tolerance']),
               float(p['divergence_tolerance'])]
        head = [n, sp._rowptr.ctypes.data, sp._colidx.ctypes.data, A._array.ctypes.data,
                bb.ctypes.data, xx.ctypes.data]
        tail = [maxit, ctypes.byref(its), ctypes.byref(rn), hist.ctypes.data,
                int(p['preconditioner'] == 'jacobi')]
        f = getattr(rt(), 'oracle_' + method)
        f.restype = ctypes.c_int
        extra = [int(p.get('gmres_restart', 30))] if method == 'gmres' else []
        f.argtypes = [ctypes.c_int] + [ctypes.c_void_p] * 5 + [ctypes.c_int] * len(extra) + \
            [ctypes.c_double] * 3 + [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                     ctypes.c_void_p, ctypes.c_int]
        reason = f(*(head + extra + tol + tail))
        x.data[:] = xx.reshape(x.data.shape)
        self._its, self._rnorm, self._reason = its.value, rn.value, reason
        self.reshist = list(hist[:its.value + 1])
        if reason < 0:
            msg = "KSP Solver failed to converge in %d iterations: reason %d (Residual norm: %e)" \
                % (its.value, reason, rn.value)
            if p['error_on_nonconvergence']:
                raise RuntimeError(msg)

    def getIterationNumber(self):
        return self._its

    def getResidualNorm(self):
        return self._rnorm

    def getConvergedReason(self):
        return self._reason


# -- code generation (pyop2/host.py conventions) --------------------------------

def _ind(argname, mapname, mdim, idx, cdim):
    return "%s + %s[i * %s + %s] * %s" % (argname, mapname, mdim, idx, cdim)


def generate(kernel, extents, args, consts):
    """C source of the sequential wrapper for ``kernel`` over ``args``."""
    wargs, decs, kargs, vec_inits, zero_tmps, add_vec, add_scal, tens = [], [], [], [], [], [], [], []
    for k, a in enumerate(args):
        name, mname = 'arg%d' % k, 'arg%d_map' % k
        d = a.data
        if a._is_mat:
            wargs.append('void* %s, int* %s, int* %s2' % (name, mname, mname))
            decs.append('oracle_csr* %s_m = (oracle_csr*)%s' % (name, name))
        else:
            wargs.append('void* %s' % name + (', int* %s' % mname if a._is_indirect else ''))
            decs.append('%s* %s_d = (%s*)%s' % (d.ctype, name, d.ctype, name))
        dn = name + '_d'
        if a._is_vec_map:
            decs.append('%s* %s_vec[%d]' % (d.ctype, name, a.map.dim))
            vec_inits += ['%s_vec[%d] = %s' % (name, j, _ind(dn, mname, a.map.dim, j, d.cdim))
                          for j in range(a.map.dim)]
        # kernel argument (host.Arg.c_kernel_arg)
        if a._uses_itspace:
            if a._is_mat:
                maps = as_tuple(a.map)
                nrows, ncols = maps[0].dim, maps[1].dim
                if d._is_vector_field:
                    tens.append('%s p_%s[%d][%d]' % (d.ctype, name, d.dims[0], d.dims[1]))
                    kargs.append('p_%s' % name)
                    zero_tmps.append('memset(p_%s, 0, sizeof(%s) * %d)'
                                     % (name, d.ctype, d.dims[0] * d.dims[1]))
                    rm, cm = d.dims
                    for r in range(rm):
                        for c in range(cm):
                            add_vec.append(
                                'addto_scalar(%s_m, &p_%s[%d][%d], %d * %s[i * %d + i_0] + %d, '
                                '%d * %s2[i * %d + i_1] + %d, %d)'
                                % (name, name, r, c, rm, mname, nrows, r, cm, mname, ncols, c,
                                   a.access is WRITE))
                else:
                    tens.append('%s p_%s%s' % (d.ctype, name, ''.join('[%d]' % e for e in extents)))
                    kargs.append('(%s (*)[1])&p_%s[i_0][i_1]' % (d.ctype, name))
                    zero_tmps.append('p_%s[i_0][i_1] = (%s)0' % (name, d.ctype))
                    add_scal.append('addto_vector(%s_m, p_%s, %d, %s + i * %d, %d, %s2 + i * %d, %d)'
                                    % (name, name, nrows, mname, nrows, ncols, mname, ncols,
                                       a.access is WRITE))
            else:
                kargs.append(_ind(dn, mname, a.map.dim, 'i_%d' % a.idx.index, d.cdim))
        elif a._is_indirect:
            if a._is_vec_map:
                kargs.append('%s_vec' % name)
            else:
                kargs.append(_ind(dn, mname, a.map.dim, a.idx, d.cdim))
        elif a._is_global:
            kargs.append(dn)
        else:
            kargs.append('%s + i * %d' % (dn, d.cdim))
    kargs += ['i_%d' % n for n in range(len(extents))]
    cargs, cinits, cdecls = [], [], []
    for c in consts:
        cargs.append('void* c_%s' % c.name)
        cdecls.append(c._format_declaration())
        if c.cdim == 1:
            cinits.append('%s = ((%s*)c_%s)[0]' % (c.name, c.ctype, c.name))
        else:
            cinits += ['%s[%d] = ((%s*)c_%s)[%d]' % (c.name, j, c.ctype, c.name, j)
                       for j in range(c.cdim)]
    loops = ''.join('for (int i_%d = 0; i_%d < %d; ++i_%d) {\n' % (n, n, e, n)
                    for n, e in enumerate(extents))
    close = '}\n' * len(extents)
    soa = any(a._is_soa for a in args)
    src = ['#include <math.h>', '#include <string.h>', '#include <stdlib.h>',
           '#include "oracle_mat.h"']
    src += cdecls
    if soa:
        src.append('#define OP2_STRIDE(a, idx) a[idx]')
    src.append('static inline ' + kernel.code)
    if soa:
        src.append('#undef OP2_STRIDE')
    src.append('void wrap(int start, int end%s%s) {'
               % (''.join(', ' + w for w in wargs), ''.join(', ' + c for c in cargs)))
    src += ['  %s;' % x for x in decs + tens + cinits]
    src.append('  for (int i = start; i < end; i++) {')
    src += ['    %s;' % x for x in vec_inits]
    src.append(loops)
    src += ['    %s;' % x for x in zero_tmps]
    src.append('    %s(%s);' % (kernel.name, ', '.join(kargs)))
    src += ['    %s;' % x for x in add_vec]
    src.append(close)
    src += ['    %s;' % x for x in add_scal]
    src.append('  }')
    src.append('}')
    return '\n'.join(src)


_compiled = {}


def compile_wrapper(src):
    """Compile a wrapper with gcc into a cached shared object."""
    h = hashlib.md5((src + ' '.join(CFLAGS)).encode()).hexdigest()
    f = _compiled.get(h)
    if f is not None:
        return f
    os.makedirs(CACHE, exist_ok=True)
    so = os.path.join(CACHE, 'w_%s.so' % h)
    if not os.path.exists(so):
        cpath = os.path.join(CACHE, 'w_%s.c' % h)
        with open(cpath, 'w') as fh:
            fh.write(src)
        tmp = so + '.%d.tmp' % os.getpid()
        subprocess.check_call(['gcc'] + CFLAGS + ['-I', os.path.join(HERE, 'csrc'), cpath,
                                                  '-o', tmp, '-lm'])
        os.replace(tmp, so)
    lib = ctypes.CDLL(so)
    f = lib.wrap
    f.restype = None
    _compiled[h] = (f, lib)
    return _compiled[h]


class ParLoop(base.ParLoop):

    def compute(self):
        consts = base.Const._definitions()
        src = generate(self.kernel, self._it_space.extents, self.args, consts)
        fun, _lib = compile_wrapper(src)
        cargs = []
        keep = []
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
        n = self._it_space.exec_size if self.needs_exec_halo else self._it_space.size
        fun(ctypes.c_int(0), ctypes.c_int(n), *[ctypes.c_void_p(x) for x in cargs])
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


@validate_type(('M', base.Mat, MatTypeError), ('x', base.Dat, DatTypeError),
               ('b', base.Dat, DatTypeError))
def solve(M, x, b):
    Solver().solve(M, x, b)


def _setup():
    rt()


def register():
    """Make ``op2.init(backend='sequential')`` select this oracle."""
    import sys
    from pyop2_amd import backends
    backends.backends['sequential'] = sys.modules[__name__]
