"""The ``hip`` backend: op2 on one MI355X (gfx950) GPU.

Selected with ``op2.init(backend='hip')`` through the plugin boundary of
:mod:`pyop2_amd.backends` (reference: pyop2/backends.py:128-146).  It exports
the classes the boundary looks up by name (Set, Map, Dat, Global, Const,
Kernel, Sparsity, Mat, Solver, IterationSpace, Halo, Arg, ParLoop), plus
``par_loop``, ``solve``, ``_setup`` and ``MPI``.

Data live in device memory owned by torch (used only as an allocator and for
the current stream); every computation is a kernel of ``libop2hip.so`` or a
hiprtc-compiled loop:

* Dat/Global values have a host copy (numpy) and a device copy with the
  coherence states of the reference's device backends (pyop2/device.py:103-183):
  HOST (host newer), DEVICE (device newer) or BOTH.  ``.data`` brings values
  to the host, a ``par_loop`` brings its arguments to the device.
* Maps are uploaded once as int32.
* Sparsities are built on the device (bit-exact CSR), Mats hold FP64 CSR
  values on the device, ``solve`` runs the device Jacobi-PCG.
* Loops run the stream or coloured templates of :mod:`pyop2_amd.codegen`;
  coloured loops use a cached :class:`Plan`.
"""

import ctypes
import hashlib
import weakref

import numpy as np

from . import base
from . import configuration as cfg
from . import _native as nat
from . import jit
from . import cg
from . import layout
from . import autotune
from .base import (READ, WRITE, RW, INC, MIN, MAX, IdentityMap, IterationIndex, i,
                   IterationSpace, Halo, Access)
from .codegen import LoopSource, _mode as _access_mode
from .exceptions import MatTypeError, DatTypeError
from .logger import debug, warning
from .mpi import MPI
from .utils import as_tuple, maybe_setflags, validate_type

try:
    import torch
except ImportError:  # pragma: no cover - torch is part of the runtime
    torch = None

HOST, DEVICE, BOTH = 'HOST', 'DEVICE', 'BOTH'

_state = {'initialised': False, 'device': None}


def _setup():
    """Select GPU 0 and initialise the native runtime (called by op2.init)."""
    if torch is None or not torch.cuda.is_available():
        raise RuntimeError("the hip backend needs a ROCm GPU (torch.cuda is unavailable)")
    nat.check(nat.lib.op2hip_init(0), 'op2hip_init')
    lb, ncu, lcu = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    nat.check(nat.lib.op2hip_device_limits(0, ctypes.byref(lb), ctypes.byref(ncu),
                                           ctypes.byref(lcu)), 'device limits')
    _state['lds_per_block'], _state['num_cu'], _state['lds_per_cu'] = \
        lb.value, ncu.value, lcu.value
    _state['device'] = torch.device('cuda', 0)
    torch.cuda.set_device(0)
    _state['initialised'] = True


def device():
    if not _state['initialised']:
        _setup()
    return _state['device']


_raw_stream = getattr(torch._C, '_cuda_getCurrentRawStream', None)


def stream():
    """The HIP stream every loop is launched on: torch's current stream."""
    if _raw_stream is not None and _state['initialised']:
        return _raw_stream(0)
    return torch.cuda.current_stream(device()).cuda_stream


def _dev_bytes(nbytes):
    return torch.empty(max(int(nbytes), 1), dtype=torch.uint8, device=device())


def _ptr(t):
    return t.data_ptr() if t is not None else 0


# uploads up to this size go through pinned staging and do not wait
H2D_ASYNC_MAX = 64 << 20


def _h2d(dev, host):
    """Copy a host numpy array's bytes into a device byte tensor, in stream
    order.  Up to H2D_ASYNC_MAX bytes the bytes are first copied into a
    pinned staging buffer (torch's caching host allocator keeps it from
    reuse until the copy has run), so the call returns without waiting for
    the stream; larger arrays are copied synchronously."""
    n = host.nbytes
    if n == 0:
        return
    if torch.cuda.is_current_stream_capturing():
        raise RuntimeError("host data reached the device during graph capture: run the "
                           "captured function eagerly once before capturing it")
    flat = np.ascontiguousarray(host).reshape(-1).view(np.uint8)
    if n <= H2D_ASYNC_MAX:
        stage = torch.empty(n, dtype=torch.uint8, pin_memory=True)
        stage.numpy()[:] = flat
        dev[:n].copy_(stage, non_blocking=True)
    else:
        dev[:n].copy_(torch.from_numpy(flat))


def _d2h(host, dev):
    """Copy a device byte tensor into a host numpy array (blocking)."""
    if host.nbytes == 0:
        return
    flags = host.flags.writeable
    if not flags and host.flags['OWNDATA']:
        host.setflags(write=True)
    dst = torch.from_numpy(host.reshape(-1).view(np.uint8))
    dst.copy_(dev[:host.nbytes])
    if not flags and host.flags['OWNDATA']:
        host.setflags(write=False)


class _DeviceData(object):
    """Host/device coherence of one array (Dat, Global).  A Dat's device
    copy is stored in its Set's layout order (see :mod:`layout`)."""

    def _dev_init(self):
        self._dev = None
        self._state = HOST

    def _layout(self):
        ds = getattr(self, '_dataset', None)
        return layout.of(ds) if ds is not None else layout.IDENTITY

    def _soa_device(self):
        """Stored component-major on the device (structure of arrays):
        component k of row r at k * rows + r (reference: device.py:153-174).
        One-component Dats are the same either way."""
        return bool(getattr(self, '_soa', False)) and self.cdim > 1

    def _upload(self):
        a = self._layout().to_device_order(self._data)
        if self._soa_device():
            a = np.ascontiguousarray(a.reshape(len(a), -1).T)
        _h2d(self._dev, a)

    def _to_device(self):
        if self._dev is None:
            self._dev = _dev_bytes(self._data.nbytes)
            self._upload()
            self._state = BOTH
        elif self._state == HOST:
            self._upload()
            self._state = BOTH
        return self._dev

    def _sync_to_host(self):
        if getattr(self, '_state', HOST) == DEVICE:
            lay = self._layout()
            soa = self._soa_device()
            if lay.identity and not soa:
                _d2h(self._data, self._dev)
            else:
                if soa:
                    n = len(self._data)
                    t = np.empty((self.cdim, n), dtype=self._data.dtype)
                    _d2h(t, self._dev)
                    tmp = np.ascontiguousarray(t.T).reshape(self._data.shape)
                else:
                    tmp = np.empty_like(self._data)
                    _d2h(tmp, self._dev)
                w = self._data.flags.writeable
                if not w and self._data.flags['OWNDATA']:
                    self._data.setflags(write=True)
                lay.to_user_order(tmp, self._data)
                if not w and self._data.flags['OWNDATA']:
                    self._data.setflags(write=False)
            self._state = BOTH

    def _host_written(self):
        self._state = HOST

    def _device_written(self):
        self._state = DEVICE


class Arg(base.Arg):
    pass


class Set(base.Set):
    pass


_map_buffers = {}


class Map(base.Map):
    # Args of this backend built directly (no backend lookup per argument)
    _arg_class = Arg

    def __init__(self, *args, **kwargs):
        base.Map.__init__(self, *args, **kwargs)
        if self is not IdentityMap:
            layout.register_map(self)

    def _layout_values(self):
        """Map values in device numbering (host int32)."""
        return layout.map_values(self)

    def _device_values(self):
        d = getattr(self, '_dev', None)
        if d is None:
            v = self._layout_values()
            # maps with identical device values share one buffer (airfoil's
            # pecell and pevcell), so a loop reading both loads one array
            key = (v.shape, hashlib.blake2b(v.tobytes(), digest_size=16).digest())
            ref = _map_buffers.get(key)
            d = ref() if ref is not None else None
            if d is None:
                d = _dev_bytes(v.nbytes)
                _h2d(d, v)
                _map_buffers[key] = weakref.ref(d)
            self._dev = d
        return d


class Dat(_DeviceData, base.Dat):
    _arg_class = Arg

    def __init__(self, *args, **kwargs):
        base.Dat.__init__(self, *args, **kwargs)
        self._dev_init()


class Global(_DeviceData, base.Global):
    _arg_class = Arg

    def __init__(self, *args, **kwargs):
        base.Global.__init__(self, *args, **kwargs)
        self._dev_init()

    def _sync_to_host(self):
        p = _pending[0]
        if p is not None and p.glob is self:
            flush()
        _DeviceData._sync_to_host(self)

    def _upload(self):
        # a value of up to 256 bytes reaches the device in stream order, as
        # a kernel argument: creating a Global in a loop costs no host wait
        a = np.ascontiguousarray(self._data)
        if a.nbytes > 256:
            return _DeviceData._upload(self)
        nat.check(nat.lib.op2hip_store_small(ctypes.c_void_p(_ptr(self._dev)),
                                             a.ctypes.data_as(ctypes.c_void_p), a.nbytes,
                                             ctypes.c_void_p(stream())), 'Global upload')


class Const(base.Const):
    pass


class Kernel(base.Kernel):
    _cache = {}


class Sparsity(base.Sparsity):
    """CSR structure built on the device (bit-exact to the reference's
    sparsity_utils.cxx:5-51 std::set construction).

    The device CSR is in device numbering (rows and columns in the row and
    column Sets' layouts); ``rowptr``/``colidx``/``nnz`` are the same
    pattern in user numbering, derived on first access."""

    _cache = {}

    def _build(self):
        for m in self._rmaps + self._cmaps:
            v = m.values
            if v.size and (v.min() < 0 or v.max() >= m.dataset.total_size):
                raise base.MapValueError("Map %s has values outside its dataset" % m.name)
        (self._dev_rowptr, self._dev_colidx, self._dev_nnz_rows,
         self._d_nz) = self._build_device()
        self._o_nnz = np.zeros(self._nrows * self._dims[0], dtype=np.int32)
        self._o_nz = 0
        self._user = None

    def _build_device(self):
        """The device CSR of the maps: (rowptr, colidx, nnz per row, nnz),
        enqueued on the current stream."""
        rmult, cmult = self._dims
        nrows = self._nrows
        lsize = nrows * rmult
        s = stream()
        rmaps = [m._device_values() for m in self._rmaps]
        cmaps = [m._device_values() for m in self._cmaps]
        rowptr = torch.empty(lsize + 1, dtype=torch.int32, device=device())
        handle = ctypes.c_void_p()
        nnz = ctypes.c_longlong()
        nel = (ctypes.c_longlong * self._nmaps)(*[m.iterset.size for m in self._rmaps])
        rdims = nat.int_array([m.dim for m in self._rmaps])
        cdims = nat.int_array([m.dim for m in self._cmaps])
        wbytes = nat.lib.op2hip_sparsity_workspace(nrows, self._nmaps, nel, rdims, cdims)
        if wbytes < 0:
            raise RuntimeError("bad sparsity arguments")
        # the workspace comes from torch's caching allocator: repeated builds
        # reuse its blocks instead of allocating device memory
        work = _dev_bytes(wbytes)
        nat.check(nat.lib.op2hip_sparsity_build(
            rmult, cmult, nrows, self._nmaps, nat.ptr_array([_ptr(t) for t in rmaps]), rdims,
            nat.ptr_array([_ptr(t) for t in cmaps]), cdims, nel, ctypes.c_void_p(_ptr(work)),
            int(work.numel()), ctypes.c_void_p(_ptr(rowptr)),
            ctypes.byref(handle), ctypes.byref(nnz), ctypes.c_void_p(s)), 'sparsity build')
        colidx = torch.empty(max(nnz.value, 1), dtype=torch.int32, device=device())
        nnz_rows = torch.empty(max(lsize, 1), dtype=torch.int32, device=device())
        nat.check(nat.lib.op2hip_sparsity_fill(handle, ctypes.c_void_p(_ptr(rowptr)),
                                               ctypes.c_void_p(_ptr(colidx)),
                                               ctypes.c_void_p(_ptr(nnz_rows)),
                                               ctypes.c_void_p(s)), 'sparsity fill')
        del work
        return rowptr, colidx, nnz_rows, int(nnz.value)

    def _row_layout(self):
        return layout.of(self._rmaps[0].dataset)

    def _col_layout(self):
        return layout.of(self._cmaps[0].dataset)

    def _user_csr(self):
        """(rowptr, colidx, nnz, slotmap) in user numbering; slotmap[k] is
        the device CSR slot of user slot k (None when both layouts are the
        identity)."""
        if self._user is not None:
            return self._user
        rmult, cmult = self._dims
        lsize = self._nrows * rmult
        nz = self._d_nz
        rp = self._dev_rowptr.cpu().numpy()
        ci = self._dev_colidx[:nz].cpu().numpy()
        self._user = layout.csr_to_user(rp, ci, rmult, cmult, self._row_layout(),
                                        self._col_layout())
        return self._user

    @property
    def _rowptr(self):
        return self._user_csr()[0]

    @property
    def _colidx(self):
        return self._user_csr()[1]

    @property
    def _d_nnz(self):
        return self._user_csr()[2]

    def _device_rows(self, rows):
        """Device row numbers of user rows (expanded numbering)."""
        rmult = self._dims[0]
        r = np.asarray(rows, dtype=np.int64).reshape(-1)
        rl = self._row_layout()
        if rl.identity:
            return r
        return rmult * rl.iperm[r // rmult] + r % rmult


class Mat(base.Mat):
    """FP64 CSR values on the device, aligned with the Sparsity's CSR
    (reference: petsc_base.py:80-120 createAIJWithArrays)."""
    _arg_class = Arg

    def __init__(self, *args, **kwargs):
        base.Mat.__init__(self, *args, **kwargs)
        if self._datatype != np.float64:
            raise RuntimeError("Can only create a matrix of type float64, %s is not supported"
                               % self._datatype)
        nnz = self._sparsity.nz
        self._dev_vals = torch.zeros(max(nnz, 1), dtype=torch.float64, device=device())
        self._err = torch.zeros(1, dtype=torch.int32, device=device())

    def _check_err(self):
        if int(self._err.item()) != 0:
            self._err.zero_()
            raise RuntimeError("Mat %s: an element added an entry outside its sparsity pattern"
                               % self.name)

    def zero(self):
        """Zero every stored value."""
        self._dev_vals.zero_()

    def zero_rows(self, rows, diag_val):
        """Zero the given rows except the diagonal, which becomes ``diag_val``
        (reference: petsc_base.py:131-135)."""
        r = np.asarray(rows, dtype=np.int32).reshape(-1)
        if r.size == 0:
            return
        nrows = self._sparsity.nrows * self.dims[0]
        if r.min() < 0 or r.max() >= nrows:
            raise IndexError("row out of range in zero_rows")
        sp = self._sparsity
        d = torch.from_numpy(sp._device_rows(r).astype(np.int32)).to(device())
        nat.check(nat.lib.op2hip_csr_zero_rows(int(r.size), ctypes.c_void_p(_ptr(d)),
                                               ctypes.c_void_p(_ptr(sp._dev_rowptr)),
                                               ctypes.c_void_p(_ptr(sp._dev_colidx)),
                                               ctypes.c_void_p(_ptr(self._dev_vals)),
                                               float(diag_val), ctypes.c_void_p(stream())),
                  'zero_rows')
        torch.cuda.current_stream().synchronize()

    def _assemble(self):
        pass

    @property
    def array(self):
        """Host copy of the CSR values."""
        self._check_err()
        vals = self._dev_vals[:self._sparsity.nz].cpu().numpy()
        slot = self._sparsity._user_csr()[3]
        return vals if slot is None else vals[slot]

    @property
    def values(self):
        """Dense host copy of the matrix (small matrices only)."""
        sp = self._sparsity
        n = sp.nrows * self.dims[0]
        m = sp.ncols * self.dims[1]
        out = np.zeros((n, m), dtype=np.float64)
        vals = self.array
        rp, ci = sp._rowptr, sp._colidx
        for r in range(n):
            out[r, ci[rp[r]:rp[r + 1]]] = vals[rp[r]:rp[r + 1]]
        return out

    @property
    def handle(self):
        return self


# linear_solver names: PETSc's ('cg', 'gmres', 'bcgs') and cusp's 'bicgstab'
# (pyop2/cuda.py:504-508)
_KRYLOV = {'cg': 0, 'gmres': 1, 'bcgs': 2, 'bicgstab': 2}


class Solver(base.Solver):
    """Jacobi-preconditioned Krylov solves on the device with PETSc 3.3
    semantics (reference: petsc_base.py:160-220): KSPCG for 'cg', KSPGMRES
    (restart ``gmres_restart``) for 'gmres', KSPBCGS for 'bcgs' or
    'bicgstab'."""

    def solve(self, A, x, b):
        p = self.parameters
        method = _KRYLOV.get(p['linear_solver'])
        if method is None:
            raise NotImplementedError("linear_solver %r: the hip backend implements %s"
                                      % (p['linear_solver'], sorted(_KRYLOV)))
        if p['preconditioner'] not in ('jacobi', 'none'):
            raise NotImplementedError("preconditioner %r: the hip backend implements 'jacobi' "
                                      "and 'none'" % p['preconditioner'])
        if p['plot_convergence']:
            p['monitor_convergence'] = True
        for d in (x, b):
            if d._soa_device():
                raise NotImplementedError("solve on a structure-of-arrays Dat: the matrix rows "
                                          "interleave its components")
        if method:
            return self._krylov(method, A, x, b)
        A._check_err()
        sp = A.sparsity
        n = sp.nrows * A.dims[0]
        bdev = b._to_device()
        xdev = x._to_device()
        work = torch.empty((5, max(n, 1)), dtype=torch.float64, device=device())
        maxit = int(p['maximum_iterations'])
        s = stream()
        sysd = cg.make_system(_ptr(sp._dev_rowptr), _ptr(sp._dev_colidx), _ptr(A._dev_vals),
                              _ptr(bdev), _ptr(xdev), [_ptr(work[k]) for k in range(5)], 0, n,
                              float(p['relative_tolerance']), float(p['absolute_tolerance']),
                              float(p['divergence_tolerance']), maxit,
                              jacobi=p['preconditioner'] == 'jacobi')
        hist_len = maxit + 1 if p['monitor_convergence'] else 0
        batch = cg.CGBatch(1, n, hist_len)
        out = batch.solve([sysd], s, chunk=int(p.get('hip_cg_chunk', 16)))[0]
        its, rnorm, reason = out.its, out.dp, out.reason
        if hist_len:
            self.reshist = list(batch.history(0, s)[:its + 1])
        self._finish(x, its, rnorm, reason)

    def _krylov(self, method, A, x, b):
        """GMRES (method 1) or BiCGStab (2) through op2hip_krylov_solve."""
        p = self.parameters
        A._check_err()
        sp = A.sparsity
        n = sp.nrows * A.dims[0]
        restart = int(p.get('gmres_restart', 30))
        if method == 1 and restart < 1:
            raise ValueError("gmres_restart must be at least 1")
        bdev = b._to_device()
        xdev = x._to_device()
        maxit = int(p['maximum_iterations'])
        # the work space (and history) is kept between solves of one size, so
        # repeated solves present the same arguments and replay the
        # library's cached round graphs
        key = (method, n, restart, maxit if p['monitor_convergence'] else None)
        if getattr(self, '_kry', None) is None or self._kry[0] != key:
            nwork = nat.lib.op2hip_krylov_work(method, n, restart)
            if nwork < 0:
                raise ValueError("bad Krylov solve arguments")
            work = torch.empty(max(int(nwork), 1), dtype=torch.float64, device=device())
            hist = None
            if p['monitor_convergence']:
                hist = torch.zeros(maxit + 1, dtype=torch.float64, device=device())
            self._kry = (key, work, hist)
        _, work, hist = self._kry
        its, rnorm, reason = ctypes.c_int(), ctypes.c_double(), ctypes.c_int()
        nat.check(nat.lib.op2hip_krylov_solve(
            method, n, ctypes.c_void_p(_ptr(sp._dev_rowptr)), ctypes.c_void_p(_ptr(sp._dev_colidx)),
            ctypes.c_void_p(_ptr(A._dev_vals)), ctypes.c_void_p(_ptr(bdev)),
            ctypes.c_void_p(_ptr(xdev)), ctypes.c_void_p(_ptr(work)), int(work.numel()), restart,
            1 if p['preconditioner'] == 'jacobi' else 0, float(p['relative_tolerance']), float(p['absolute_tolerance']),
            float(p['divergence_tolerance']), maxit,
            ctypes.c_void_p(_ptr(hist) if hist is not None else None),
            0 if hist is None else int(hist.numel()), ctypes.byref(its), ctypes.byref(rnorm),
            ctypes.byref(reason), ctypes.c_void_p(stream())), 'krylov solve')
        its, rnorm, reason = its.value, rnorm.value, reason.value
        if hist is not None:
            self.reshist = list(hist[:its + 1].cpu().numpy()) if n else [0.0]
        self._finish(x, its, rnorm, reason)

    def _finish(self, x, its, rnorm, reason):
        p = self.parameters
        x._device_written()
        x.needs_halo_update = True
        self._its, self._rnorm, self._reason = its, rnorm, reason
        debug("Converged reason: %s" % cg.REASONS.get(reason, reason))
        debug("Iterations: %s" % its)
        debug("Residual norm: %s" % rnorm)
        if reason < 0:
            msg = "KSP Solver failed to converge in %d iterations: %s (Residual norm: %e)" \
                % (its, cg.REASONS.get(reason, reason), rnorm)
            if p['error_on_nonconvergence']:
                raise RuntimeError(msg)
            warning(msg)

    def getIterationNumber(self):
        return self._its

    def getResidualNorm(self):
        return self._rnorm

    def getConvergedReason(self):
        return self._reason


# -- plans -------------------------------------------------------------------

class Plan(object):
    """Partition and colouring of an iteration set for a coloured loop,
    built natively (csrc/plan.cpp) and kept on the device.

    Properties mirror the reference's plan (pyop2/plan.pyx:378-455):
    ``nblocks``, ``ncolors``, ``ncolblk``, ``blkmap``, ``offset``,
    ``nelems``, ``nthrcol``, ``thrcol``; ``elems`` is the execution order."""

    ARRAYS = ('elems', 'offset', 'nelems', 'thrcol', 'nthrcol', 'blkcol', 'blkmap', 'ncolblk')

    def __init__(self, n, part_size, reorder, race):
        self.n = n
        self.part_size = part_size
        self.race = race  # [(map values, dim, idx, space, space_size)]
        spaces = []
        for r in race:
            if r[3] not in spaces:
                spaces.append(r[3])
        sizes = [0] * len(spaces)
        for r in race:
            sizes[spaces.index(r[3])] = r[4]
        self._maps_keep = [np.ascontiguousarray(r[0], dtype=np.int32) for r in race]
        self._args = (nat.ptr_array([m.ctypes.data for m in self._maps_keep]),
                      nat.int_array([r[1] for r in race]),
                      nat.int_array([r[2] for r in race]),
                      nat.int_array([spaces.index(r[3]) for r in race]), len(spaces),
                      nat.int_array(sizes))
        h = ctypes.c_void_p()
        mp, dims, idxs, sp, nsp, ssz = self._args
        nat.check(nat.lib.op2hip_plan_build(n, part_size, int(reorder), len(race), mp, dims,
                                            idxs, sp, nsp, ssz, ctypes.byref(h)), 'plan build')
        try:
            v = ctypes.c_int64()
            nat.check(nat.lib.op2hip_plan_scalar(h, 1, ctypes.byref(v)))
            self.nblocks = v.value
            nat.check(nat.lib.op2hip_plan_scalar(h, 2, ctypes.byref(v)))
            self.ncolors = v.value
            sizes = {'elems': n, 'offset': self.nblocks, 'nelems': self.nblocks, 'thrcol': n,
                     'nthrcol': self.nblocks, 'blkcol': self.nblocks, 'blkmap': self.nblocks,
                     'ncolblk': self.ncolors}
            for k, name in enumerate(self.ARRAYS):
                a = np.zeros(sizes[name], dtype=np.int32)
                if a.size:
                    nat.check(nat.lib.op2hip_plan_array(h, k, a.ctypes.data_as(ctypes.c_void_p)))
                setattr(self, name, a)
        finally:
            nat.lib.op2hip_plan_destroy(h)
        self.nthrcol_max = int(self.nthrcol.max()) if self.nblocks else 0
        self.colour_start = np.concatenate([[0], np.cumsum(self.ncolblk)]).astype(np.int64)
        self._dev = None

    def block_elements(self, b):
        """(elements in slot order, primary flags) of active block ``b``."""
        ex = self.exec[self.ex_off[b]:self.ex_off[b + 1]].view(np.uint32)
        rn = int(self.rng_n[b])
        els = np.concatenate([np.arange(self.rng_lo[b], self.rng_lo[b] + rn),
                              (ex & 0x7fffffff).astype(np.int64)])
        prim = np.concatenate([np.ones(rn, dtype=bool), (ex >> 31) == 0])
        return els, prim

    def device_arrays(self):
        """The plan arrays the coloured loops read, uploaded on first use."""
        if self._dev is None:
            self._dev = {}
            for name in ('elems', 'blkmap', 'offset', 'nelems', 'nthrcol', 'thrcol'):
                a = getattr(self, name)
                d = _dev_bytes(a.nbytes)
                _h2d(d, a)
                self._dev[name] = d
        return self._dev

    def staging(self, m):
        """The reference plan's staging lists for map ``m`` (a hip Map on this
        plan's iteration set; plan.pyx:112-212, ind_map / loc_map / ind_sizes
        / ind_offs), built on the device (:func:`stage_lists`) for the plan's
        blocks: a dict of host arrays ``ind_map`` (block b's distinct targets
        ascending at ``ind_offs[b]``, ``ind_sizes[b]`` of them), ``loc_map``
        (per element slot of the plan's ``elems`` and map column, the
        position of its target in its block's list) and ``ind_offs``
        (offset * arity), in device numbering."""
        off = self.offset.astype(np.int64)
        ne = self.nelems.astype(np.int64)
        dim = m.dim
        if self.nblocks and (off[0] != 0 or (off[1:] != off[:-1] + ne[:-1]).any()):
            raise RuntimeError("plan blocks are not consecutive runs of elems")
        xoff = np.concatenate([off, [off[-1] + ne[-1]] if self.nblocks else [0]]).astype(np.int32)
        if int(xoff[-1]) > len(self.elems):
            raise RuntimeError("plan blocks run past its elements")
        elems = self.device_arrays()['elems']
        st = stage_lists(torch.from_numpy(xoff).to(device()), elems,
                         m._device_values().view(torch.int32), dim, int(xoff[-1]) * dim,
                         int(ne.max()) * dim if self.nblocks else 0)
        if st is None:
            raise RuntimeError("a plan block names more targets than the staging kernel holds")
        ind_map, sizes = st[4], st[5]
        nsl = int(xoff[-1]) * dim
        return {'ind_map': ind_map[:nsl].cpu().numpy(),
                'loc_map': st[2][:nsl].cpu().numpy().view(np.uint16),
                'ind_sizes': sizes.cpu().numpy(), 'ind_offs': xoff[:-1].astype(np.int64) * dim}

    def verify(self):
        """True when the plan is conflict-free for its race maps."""
        ok = ctypes.c_int()
        mp, dims, idxs, sp, nsp, ssz = self._args
        c = lambda a: a.ctypes.data_as(ctypes.c_void_p)
        nat.check(nat.lib.op2hip_plan_verify(self.n, self.nblocks, c(self.elems), c(self.offset),
                                             c(self.nelems), c(self.thrcol), c(self.blkcol),
                                             len(self.race), mp, dims, idxs, sp, nsp, ssz,
                                             ctypes.byref(ok)), 'plan verify')
        return bool(ok.value)


class DevicePlan(Plan):
    """A coloured plan built on the device (csrc/plan_colour.hip:
    op2hip_plan_dbuild/dfill): mode 0 colours blocks of ``part_size``
    consecutive elements and threads inside them (the reference's two
    levels), mode 2 colours elements and cuts each colour into blocks.
    ``race`` as for :class:`Plan`, with the hip Maps (their device values)
    first: [(map, dim, idx, space, space_size)].  The arrays are fetched to
    the host (the attributes of :class:`Plan`), where :meth:`verify` checks
    them."""

    def __init__(self, n, part_size, mode, race):
        self.n = n
        self.part_size = part_size
        spaces = []
        for r in race:
            if r[3] not in spaces:
                spaces.append(r[3])
        sizes = [0] * len(spaces)
        for r in race:
            sizes[spaces.index(r[3])] = r[4]
        self.race = [(r[0]._layout_values()[:n], r[1], r[2], r[3], r[4]) for r in race]
        self._maps_keep = [np.ascontiguousarray(r[0], dtype=np.int32) for r in self.race]
        self._args = (nat.ptr_array([m.ctypes.data for m in self._maps_keep]),
                      nat.int_array([r[1] for r in race]),
                      nat.int_array([r[2] for r in race]),
                      nat.int_array([spaces.index(r[3]) for r in race]), len(spaces),
                      nat.int_array(sizes))
        _, dims, idxs, sp, nsp, ssz = self._args
        nac = sum(r[1] if r[2] < 0 else 1 for r in race)
        wbytes = nat.lib.op2hip_plan_workspace(n, part_size, nac, ssz, nsp)
        if wbytes < 0:
            raise RuntimeError("bad device plan arguments")
        work = _dev_bytes(wbytes)
        elems = torch.empty(max(n, 1), dtype=torch.int32, device=device())
        thrcol = torch.empty(max(n, 1), dtype=torch.int32, device=device())
        dmaps = [r[0]._device_values() for r in race]
        h = ctypes.c_void_p()
        nb, nc = ctypes.c_int(), ctypes.c_int()
        s = stream()
        nat.check(nat.lib.op2hip_plan_dbuild(
            n, part_size, mode, len(race), nat.ptr_array([_ptr(t) for t in dmaps]), dims, idxs,
            sp, nsp, ssz, ctypes.c_void_p(_ptr(work)), int(work.numel()),
            ctypes.c_void_p(_ptr(thrcol)), ctypes.byref(h), ctypes.byref(nb), ctypes.byref(nc),
            ctypes.c_void_p(s)), 'device plan build')
        self.nblocks, self.ncolors = nb.value, nc.value
        small = {k: np.zeros(self.nblocks, dtype=np.int32)
                 for k in ('offset', 'nelems', 'nthrcol', 'blkcol', 'blkmap')}
        small['ncolblk'] = np.zeros(self.ncolors, dtype=np.int32)
        P = lambda a: a.ctypes.data_as(ctypes.c_void_p)
        nat.check(nat.lib.op2hip_plan_dfill(
            h, ctypes.c_void_p(_ptr(elems)), ctypes.c_void_p(_ptr(thrcol)), P(small['offset']),
            P(small['nelems']), P(small['nthrcol']), P(small['blkcol']), P(small['blkmap']),
            P(small['ncolblk']), ctypes.c_void_p(s)), 'device plan fill')
        del work
        for k, v in small.items():
            setattr(self, k, v)
        self.elems = elems[:n].cpu().numpy().astype(np.int32)
        self.thrcol = thrcol[:n].cpu().numpy().astype(np.int32)
        self.nthrcol_max = int(self.nthrcol.max()) if self.nblocks else 0
        self.colour_start = np.concatenate([[0], np.cumsum(self.ncolblk)]).astype(np.int64)
        self._dev = None


_plan_cache = {}


def _race_maps(args):
    """Race targets of a loop: (map, idx, space key) per conflicting write.

    INC through a map and every Mat (its row map) need conflict handling,
    as in the reference plan (pyop2/plan.pyx:225-238 with matrix colouring)."""
    race = []
    for a in args:
        if a._is_mat:
            rmap = a.map[0]
            race.append((rmap, -1, ('mat', id(rmap.dataset)), rmap.dataset.total_size))
        elif a._is_indirect and a.access is INC:
            idx = a.idx if isinstance(a.idx, int) else -1
            race.append((a.map, idx, ('dat', id(a.map.dataset)), a.map.dataset.total_size))
    return race


def get_plan(iterset, n, args, part_size=None, reorder=None):
    """The cached plan of a coloured loop over ``n`` elements of ``iterset``:
    element colours or the reference's two-level block/thread colouring
    (``hip_colouring`` 'elements' or 'blocks'; 'auto', the default, colours
    elements for Mat assembly, whose many thread colours per block would
    serialise its addto, and blocks otherwise)."""
    if part_size is None:
        part_size = cfg.get('hip_partition_size') or 256
    if reorder is None:
        reorder = cfg.get('hip_reorder', True)
        how = cfg.get('hip_colouring', 'auto')
        if how == 'elements' or (how == 'auto' and any(a._is_mat for a in args)):
            reorder = 2
    race = _race_maps(args)
    # one entry per distinct (map, idx, space)
    seen, uniq = set(), []
    for r in race:
        k = (id(r[0]), r[1], r[2])
        if k not in seen:
            seen.add(k)
            uniq.append(r)
    dev = bool(cfg.get('hip_device_plan', True))
    key = (id(iterset), n, part_size, int(reorder), dev,
           tuple((id(r[0]), r[1], r[2]) for r in uniq))
    p = _plan_cache.get(key)
    if p is None and dev:
        # on the device: element colours (reorder 2), or the two levels over
        # blocks of consecutive elements of the device layout (reorder 0 and
        # 1: the layout already orders the iteration set into compact runs)
        try:
            p = DevicePlan(n, part_size, 2 if int(reorder) == 2 else 0,
                           [(r[0], r[0].dim, r[1], r[2], r[3]) for r in uniq])
        except nat.NativeError as e:
            # more colours than the device colouring keeps (a target shared
            # by hundreds of elements): the host build colours in rounds
            debug("device plan refused (%s); building on the host" % e)
            p = None
    if p is None:
        p = Plan(n, part_size, reorder,
                 [(r[0]._layout_values()[:n], r[0].dim, r[1], r[2], r[3]) for r in uniq])
    if key not in _plan_cache:
        p._keep = [r[0] for r in uniq]
        _plan_cache[key] = p
    return p


class OwnerPlan(object):
    """Owner-computes plan (csrc/plan.cpp op2hip_owner_plan_build): targets
    owned by blocks in contiguous ranges of ``k``; each block runs the
    elements naming its targets and sums each owned target's increments in
    ascending element order."""

    NATIVE = ('blk_t0', 'blk_t1', 'ex_off', 'exec', 'tp', 'contrib', 'tg_off', 'tgt',
              'rng_lo', 'rng_n')
    ARRAYS = NATIVE + ('tg_cb',)

    def __init__(self, n, ntargets, cols, k, elem_order=None):
        self.n, self.ntargets, self.k = n, ntargets, k
        keep = [np.ascontiguousarray(v, dtype=np.int32) for v, _, _ in cols]
        order = None if elem_order is None else np.ascontiguousarray(elem_order, dtype=np.int32)
        h = ctypes.c_void_p()
        nat.check(nat.lib.op2hip_owner_plan_build(
            n, ntargets, len(cols), nat.ptr_array([v.ctypes.data for v in keep]),
            nat.int_array([d for _, d, _ in cols]), nat.int_array([i for _, _, i in cols]), k,
            None if order is None else order.ctypes.data_as(ctypes.c_void_p),
            ctypes.byref(h)), 'owner plan build')
        try:
            v = ctypes.c_int64()
            sc = []
            for f in range(6):
                nat.check(nat.lib.op2hip_owner_plan_scalar(h, f, ctypes.byref(v)))
                sc.append(v.value)
            self.nactive, self.max_exec, self.nac, nexec, ncontrib, ntgt = sc
            sizes = {'blk_t0': self.nactive, 'blk_t1': self.nactive,
                     'ex_off': self.nactive + 1, 'exec': nexec, 'tp': ntargets + 1,
                     'contrib': ncontrib, 'tg_off': self.nactive + 1, 'tgt': ntgt,
                     'rng_lo': self.nactive, 'rng_n': self.nactive}
            for w, name in enumerate(self.NATIVE):
                arr = np.zeros(sizes[name], dtype=np.int32)
                if arr.size:
                    nat.check(nat.lib.op2hip_owner_plan_array(
                        h, w, arr.ctypes.data_as(ctypes.c_void_p)))
                setattr(self, name, arr)
        finally:
            nat.lib.op2hip_owner_plan_destroy(h)
        # contribution offsets by touched-target position (targets without
        # increments have empty ranges, so the next entry ends each range)
        self.tg_cb = np.append(self.tp[self.tgt], self.tp[-1]).astype(np.int32)
        # elements run per block (run of primaries plus listed ones)
        self.nexec = np.diff(self.ex_off) + self.rng_n if self.nactive else \
            np.zeros(0, dtype=np.int64)
        # contribution codes per block (its owned targets' range of contrib)
        self.max_codes = int((self.tp[self.blk_t1] - self.tp[self.blk_t0]).max()) \
            if self.nactive else 0
        self.sum_exec = int(self.nexec.sum()) if self.nactive else 0
        # device lists carry one padding entry (tgt 0, tg_cb its end), so a
        # kernel may load entry ng of a block unconditionally
        self._pad = {'tgt': 0, 'tg_cb': int(self.tp[-1])}
        self._dev = None

    def block_elements(self, b):
        """(elements in slot order, primary flags) of active block ``b``."""
        ex = self.exec[self.ex_off[b]:self.ex_off[b + 1]].view(np.uint32)
        rn = int(self.rng_n[b])
        els = np.concatenate([np.arange(self.rng_lo[b], self.rng_lo[b] + rn),
                              (ex & 0x7fffffff).astype(np.int64)])
        prim = np.concatenate([np.ones(rn, dtype=bool), (ex >> 31) == 0])
        return els, prim

    @property
    def c16(self):
        """Contribution codes fit 16 bits (the device copy is uint16)."""
        return self.max_exec * self.nac <= 0xffff

    @property
    def exec_per_block(self):
        return self.sum_exec / float(self.nactive) if self.nactive else 0.0

    def device_arrays(self):
        if self._dev is None:
            self._dev = {}
            for name in self.ARRAYS:
                a = getattr(self, name)
                if name == 'contrib' and self.c16:
                    a = a.astype(np.uint16)
                if name in self._pad:
                    a = np.append(a, np.int32(self._pad[name])).astype(np.int32)
                d = _dev_bytes(a.nbytes)
                _h2d(d, a)
                self._dev[name] = d
        return self._dev


class DeviceOwnerPlan(object):
    """The owner-computes plan built on the device
    (csrc/plan_device.hip: op2hip_owner_plan_dbuild/dfill), entry for entry
    the arrays of :class:`OwnerPlan`, which stays the host restatement the
    tests compare it with.  ``cols`` are (map, idx) pairs of hip Maps (their
    device values are read); ``rank`` (device int32, may be None) orders each
    target's increments.  Host copies of the arrays are fetched on first
    access (tests, the Mat scatter program)."""

    def __init__(self, n, ntargets, cols, k, rank=None):
        self.n, self.ntargets, self.k = n, ntargets, k
        maps = [m._device_values() for m, _ in cols]
        dims = nat.int_array([m.dim for m, _ in cols])
        idxs = nat.int_array([i for _, i in cols])
        nac = sum(m.dim if i < 0 else 1 for m, i in cols)
        wbytes = nat.lib.op2hip_owner_plan_workspace(n, ntargets, nac, k)
        if wbytes < 0:
            raise RuntimeError("bad owner plan arguments")
        work = _dev_bytes(wbytes)
        tp = torch.empty(ntargets + 1, dtype=torch.int32, device=device())
        h = ctypes.c_void_p()
        sc = (ctypes.c_longlong * 6)()
        s = stream()
        nat.check(nat.lib.op2hip_owner_plan_dbuild(
            n, ntargets, len(cols), nat.ptr_array([_ptr(t) for t in maps]), dims, idxs, k,
            ctypes.c_void_p(_ptr(rank)), ctypes.c_void_p(_ptr(work)), int(work.numel()),
            ctypes.c_void_p(_ptr(tp)), ctypes.byref(h), sc, ctypes.c_void_p(s)),
            'device owner plan build')
        self.nactive, nexec, ntgt, self.max_exec, self.max_codes, self.sum_exec = \
            [int(x) for x in sc]
        self.nac = nac
        ninc = n * nac
        i32 = lambda m: torch.empty(max(m, 1), dtype=torch.int32, device=device())
        d = {'blk_t0': i32(self.nactive), 'blk_t1': i32(self.nactive),
             'rng_lo': i32(self.nactive), 'rng_n': i32(self.nactive),
             'ex_off': i32(self.nactive + 1), 'tg_off': i32(self.nactive + 1),
             'exec': i32(nexec), 'tgt': i32(ntgt + 1), 'tg_cb': i32(ntgt + 2), 'tp': tp,
             'contrib': torch.empty(max(ninc, 1), dtype=torch.int16 if self.c16 else torch.int32,
                                    device=device())}
        P = lambda name: ctypes.c_void_p(_ptr(d[name]))
        nat.check(nat.lib.op2hip_owner_plan_dfill(
            h, P('tp'), P('blk_t0'), P('blk_t1'), P('rng_lo'), P('rng_n'), P('ex_off'), P('exec'),
            P('tg_off'), P('tgt'), P('tg_cb'), P('contrib'), int(self.c16), ctypes.c_void_p(s)),
            'device owner plan fill')
        del work
        self._sizes = {'blk_t0': self.nactive, 'blk_t1': self.nactive, 'rng_lo': self.nactive,
                       'rng_n': self.nactive, 'ex_off': self.nactive + 1,
                       'tg_off': self.nactive + 1, 'exec': nexec, 'tgt': ntgt,
                       'tg_cb': ntgt + 1, 'tp': ntargets + 1, 'contrib': ninc}
        self._dev = d
        self._host = {}

    @property
    def c16(self):
        return self.max_exec * self.nac <= 0xffff

    @property
    def exec_per_block(self):
        return self.sum_exec / float(self.nactive) if self.nactive else 0.0

    def device_arrays(self):
        return self._dev

    def __getattr__(self, name):
        sizes = self.__dict__.get('_sizes')
        if sizes is None or name not in sizes:
            raise AttributeError(name)
        a = self._host.get(name)
        if a is None:
            t = self._dev[name][:sizes[name]]
            a = t.cpu().numpy()
            if name == 'contrib' and self.c16:
                a = a.view(np.uint16)
            a = a.astype(np.int32)
            self._host[name] = a
        return a

    def block_elements(self, b):
        return OwnerPlan.block_elements(self, b)


class RowsMatPlan(object):
    """Row-block plan of the owner Mat "rows" template, built on the device
    (csrc/plan_device.hip op2hip_mat_rows_build).  Block b owns the row
    nodes [b*k, min(nnodes, (b+1)*k)) of the device layout; its elements
    (``exec[xoff[b]:xoff[b+1]]``, device numbering) are those naming one of
    its nodes as a row, in ascending user order (``rank``); per row node,
    ``tp`` delimits its records: ``code`` = slot * e0 + i0 (uint16 when
    ``c16``) in ascending (user element, row position) order and ``pos``,
    ceil(e1 / 4) words of column positions (byte j: the position of the
    element's column node j among the row's node columns).  ``fits`` is
    False when a row has more than 256 node columns (the positions would not
    fit a byte); such a plan is not used."""

    def __init__(self, n, arg, e0, e1, k, rank=None):
        rm, cm = arg.map[0], arg.map[1]
        sp = arg.data.sparsity
        rd, cd = arg.data.dims
        nnodes = rm.dataset.total_size
        self.n, self.k, self.nnodes, self.e0, self.e1 = n, k, nnodes, e0, e1
        self.nblk = (nnodes + k - 1) // k
        ninc = n * e0
        W = (e1 + 3) // 4
        wbytes = nat.lib.op2hip_mat_rows_workspace(n, nnodes, e0, k)
        if wbytes < 0:
            raise RuntimeError("bad row-block plan arguments")
        work = _dev_bytes(wbytes)
        i32 = lambda m: torch.empty(max(m, 1), dtype=torch.int32, device=device())
        # padding the kernel may load past the ends (never used): exec
        # ROWS_EXEC_PAD zeros (element 0), records ROWS_RECORDS_PAD entries
        d = {'tp': i32(nnodes + 1), 'xoff': i32(self.nblk + 1),
             'exec': torch.zeros(ninc + ROWS_EXEC_PAD, dtype=torch.int32, device=device()),
             'pos': torch.zeros((ninc + ROWS_RECORDS_PAD) * W, dtype=torch.int32, device=device())}
        c32 = i32(ninc)
        sc = (ctypes.c_longlong * 5)()
        P = lambda t: ctypes.c_void_p(_ptr(t))
        nat.check(nat.lib.op2hip_mat_rows_build(
            n, nnodes, P(rm._device_values()), rm.dim, e0, P(cm._device_values()), cm.dim, e1,
            P(rank), P(sp._dev_rowptr), P(sp._dev_colidx), rd, cd, k, P(work),
            int(work.numel()), P(d['tp']), P(d['xoff']), P(d['exec']), P(c32), P(d['pos']), sc,
            ctypes.c_void_p(stream())), 'row-block Mat plan')
        del work
        self.max_code, self.max_exec, self.max_slice, self.sum_exec, flags = [int(x) for x in sc]
        if flags & 1:
            raise RuntimeError("an element names a matrix entry outside the sparsity pattern")
        if flags & 4:
            raise RuntimeError("inconsistent row-block Mat plan")
        # the template reads 16-bit codes: a block of more than 65536 / e0
        # elements does not fit
        self.fits = not flags & 2 and self.max_code <= 0xffff
        self.c16 = True
        d['code'] = torch.zeros(ninc + ROWS_RECORDS_PAD, dtype=torch.int16, device=device())
        if self.fits and ninc:
            nat.check(nat.lib.op2hip_mat_rows_pack16(P(c32), ninc, P(d['code']),
                                                      ctypes.c_void_p(stream())), 'pack codes')
        self._dev = d
        self._xmaps = {}
        self._stages = {}
        self._sizes = {'tp': nnodes + 1, 'xoff': self.nblk + 1, 'exec': ninc, 'code': ninc,
                       'pos': ninc * W}

    @property
    def redundancy(self):
        """Element executions per element."""
        return self.sum_exec / float(self.n) if self.n else 0.0

    def device_arrays(self):
        return self._dev

    def stage(self, m):
        """Per-block staging lists through map ``m`` (the reference's
        per-block ind_map and loc_map, plan.pyx:112-212), cached per map
        array: (sn, so, xl, smax).  Block b's distinct targets, ascending,
        are sn[so[b]:so[b+1]]; xl holds per exec slot the positions of its
        element's m.dim targets in its block's list (uint16, padded like
        exec with position 0); smax is the longest list.  None when a list
        has more than op2hip_stage_max_keys() keys.  Built on the device by
        :func:`stage_lists`."""
        dv = m._device_values()
        key = dv.data_ptr()
        if key not in self._stages:
            st = stage_lists(self._dev['xoff'][:self.nblk + 1], self._dev['exec'],
                             dv.view(torch.int32), m.dim, self.sum_exec * m.dim,
                             self.max_exec * m.dim, ROWS_EXEC_PAD)
            self._stages[key] = None if st is None else st[:4]
        return self._stages[key]

    def xmap(self, m):
        """The rows of map ``m`` (device values) in exec-list order, padded
        like exec (element 0's row): what the template's element pipeline
        reads instead of element numbers and the map (cached per map
        array)."""
        dv = m._device_values()
        hit = self._xmaps.get(dv.data_ptr())
        if hit is None:
            idx = self._dev['exec'][:self.sum_exec + ROWS_EXEC_PAD].long()
            hit = dv.view(torch.int32).view(-1, m.dim).index_select(0, idx).contiguous()
            self._xmaps[dv.data_ptr()] = hit
        return hit

    def host(self, name):
        """A host copy of one plan array (tests)."""
        t = self._dev[name][:self._sizes[name]]
        a = t.cpu().numpy()
        if name == 'code' and self.c16:
            a = a.view(np.uint16)
        if name == 'pos':
            a = a.view(np.uint32)
        return a.astype(np.int64)


def stage_lists(xoff, elems, mvals, dim, nslots, maxkeys, pad=0):
    """Per-block staging lists on the device (csrc/staging.hip; the
    reference's ind_map / loc_map, plan.pyx:112-212): block b runs the
    elements ``elems[xoff[b]:xoff[b+1]]`` (device int32 tensors; ``xoff``
    has nblocks + 1 ascending entries, ``elems`` at least ``nslots / dim``)
    and names targets through ``mvals`` (device int32 map values, arity
    ``dim``).  Returns (sn, so, xl, smax, ind_map, ind_sizes): block b's
    distinct targets ascending are sn[so[b]:so[b+1]] (also ind_map from
    xoff[b] * dim, ind_sizes[b] of them); xl holds per slot (xoff[b] * dim
    + e * dim + j) the position of its target in its block's list (int16
    holding uint16, ``pad`` zero slots appended); smax is the longest list
    (at least 2).  None when a block has more than ``maxkeys`` keys or
    ``maxkeys`` exceeds what the device's LDS holds."""
    nb = xoff.numel() - 1
    if maxkeys > nat.lib.op2hip_stage_max_keys():
        return None
    dev = xoff.device
    i32 = lambda m: torch.zeros(max(m, 1), dtype=torch.int32, device=dev)
    ind_map, sizes, err = i32(nslots), i32(nb), i32(1)
    xl = torch.zeros(nslots + pad * dim, dtype=torch.int16, device=dev)
    so, sn = i32(nb + 1), i32(nslots)
    P = lambda t: ctypes.c_void_p(_ptr(t))
    s = ctypes.c_void_p(stream())
    nat.check(nat.lib.op2hip_stage_lists(nb, P(xoff), P(elems), P(mvals), dim, maxkeys, nslots,
                                         P(ind_map), P(xl), P(sizes), P(err), s),
              'staging lists')
    nat.check(nat.lib.op2hip_stage_compact(nb, P(xoff), dim, P(sizes), P(ind_map), P(so), P(sn),
                                           s), 'staging compaction')
    if int(err[0]):
        return None
    smax = int(sizes[:nb].max()) if nb else 0
    return sn, so, xl, max(smax, 2), ind_map, sizes[:nb]


_rows_cache = {}
# largest local matrix (values per element) the rows template takes
ROWS_MAX_ENTRIES = 16
# padding entries of the row-block plan arrays (loads past the ends need no
# guard): exec for three chunks of up to 1024 threads ahead, records for one
# register batch
ROWS_EXEC_PAD = 4 * 1024
ROWS_RECORDS_PAD = 16
# LDS a row block may use by default (a 64-thread block of 32 Laplace rows
# takes ~12 KB)
ROWS_LDS_LIMIT = 52 * 1024


def get_rows_plan(iterset, n, args, bs, lds_of, lds_limit, stage_maps=()):
    """Cached row-block plan of an owner Mat assembly, or None: ``k`` row
    nodes per block (``hip_mat_rows_k``, default the block size), halved
    until the block's LDS ``lds_of(k, largest slice, longest staging list
    through stage_maps)`` fits ``lds_limit``; None when no k from 8 up fits,
    a row's column positions do not fit a byte or a staging list is too
    long."""
    a = [x for x in args if x._is_mat][0]
    sp = a.data.sparsity
    e0, e1 = a.map[0].dim, a.map[1].dim
    k = min(int(cfg.get('hip_mat_rows_k') or bs), bs)
    key = (id(iterset), n, id(a.map[0]), id(a.map[1]), id(sp), k, bs, lds_limit,
           tuple(id(m) for m in stage_maps))
    hit = _rows_cache.get(key)
    if hit is not None and hit[0] is sp:
        return hit[1]
    lay = layout.of(iterset)
    rank = None if lay.identity else _device_perm(lay)
    plan = None
    # the chunk's local matrices alone may not fit
    while k >= 8 and lds_of(k, 0, 0) <= lds_limit:
        p = RowsMatPlan(n, a, e0, e1, k, rank)
        if not p.fits:
            break
        st = [p.stage(m) for m in stage_maps]
        if any(x is None for x in st):
            break
        if lds_of(k, p.max_slice, max([x[3] for x in st] or [0])) <= lds_limit:
            plan = p
            break
        k //= 2
    if len(_rows_cache) >= _CALL_CACHE_MAX:
        _rows_cache.pop(next(iter(_rows_cache)))
    _rows_cache[key] = (sp, plan)
    return plan


_owner_cache = {}
OWNER_LDS_LIMIT = 64 * 1024
# LDS kept free of dynamic allocations for the kernels' static arrays
STATIC_LDS_MARGIN = 1024
# MI355X (gfx950) compute unit: 160 KB of LDS, 4 SIMDs of at most 8 waves
LDS_PER_CU = 160 * 1024
WAVES_PER_CU = 32


def _device_perm(lay):
    """A layout's permutation as a device int32 array (cached)."""
    d = getattr(lay, '_dev_perm', None)
    if d is None:
        d = torch.from_numpy(np.ascontiguousarray(lay.perm, dtype=np.int32)).to(device())
        lay._dev_perm = d
    return d


def mat_program_host(plan, cmap, e1, nnodes, rowptr, colidx, rd, cd):
    """Host arrays (poff, prog, wide) of the scatter program of an owner Mat
    assembly (csrc/plan.cpp op2hip_mat_program): per node pair of the
    pattern (rowptr/colidx: the scalar CSR in device numbering), the codes
    c * e1 + j of the block-local contributions and column positions that
    add into it, in sequential order."""
    rp = np.ascontiguousarray(rowptr, dtype=np.int32)
    ci = np.ascontiguousarray(colidx, dtype=np.int32)
    cmap = np.ascontiguousarray(cmap, dtype=np.int32)
    npairs = int(rp[nnodes * rd]) // (rd * cd)
    wide = int(max(plan.max_codes, 1) * e1 > 0xffff)
    poff = np.empty(npairs + 1, dtype=np.int32)
    prog = np.empty(len(plan.contrib) * e1, dtype=np.int32 if wide else np.uint16)
    arrs = [np.ascontiguousarray(x, dtype=np.int32) for x in
            (plan.blk_t0, plan.blk_t1, plan.ex_off, plan.exec, plan.rng_lo, plan.rng_n, plan.tp,
             plan.contrib)]
    P = lambda x: x.ctypes.data_as(ctypes.c_void_p)
    nat.check(nat.lib.op2hip_mat_program(
        plan.nactive, *[P(x) for x in arrs], plan.nac, P(cmap), cmap.shape[1], e1, nnodes,
        P(rp), P(ci), rd, cd, wide, P(poff), P(prog)), 'Mat scatter program')
    return poff, prog, wide


def _mat_program(plan, arg, e1):
    """Device scatter program of an owner Mat assembly, cached on the plan
    per sparsity: (poff, prog, wide)."""
    sp = arg.data.sparsity
    progs = plan.__dict__.setdefault('_progs', {})
    hit = progs.get(id(sp))
    if hit is not None and hit[0] is sp:
        return hit[1]
    rd, cd = arg.data.dims
    if isinstance(plan, DeviceOwnerPlan):
        res = _mat_program_device(plan, arg, e1, sp, rd, cd)
    else:
        poff, prog, wide = mat_program_host(
            plan, arg.map[1]._layout_values(), e1, arg.map[0].dataset.total_size,
            sp._dev_rowptr.cpu().numpy().view(np.int32),
            sp._dev_colidx.cpu().numpy().view(np.int32), rd, cd)
        d_poff, d_prog = _dev_bytes(poff.nbytes), _dev_bytes(prog.nbytes)
        _h2d(d_poff, poff)
        _h2d(d_prog, prog)
        res = (d_poff, d_prog, wide)
    progs[id(sp)] = (sp, res)
    return res


def _mat_program_device(plan, arg, e1, sp, rd, cd):
    """The scatter program built on the device from a device owner plan
    (csrc/plan_device.hip op2hip_mat_program_device): (poff, prog, wide)."""
    nnodes = arg.map[0].dataset.total_size
    npairs = sp._d_nz // (rd * cd)
    wide = int(max(plan.max_codes, 1) * e1 > 0xffff)
    ncodes = plan.n * plan.nac * e1
    d_poff = torch.empty(npairs + 1, dtype=torch.int32, device=device())
    d_prog = _dev_bytes(max(ncodes, 1) * (4 if wide else 2))
    work = _dev_bytes(nat.lib.op2hip_mat_program_workspace(npairs))
    d = plan.device_arrays()
    P = lambda t: ctypes.c_void_p(_ptr(t))
    cmap = arg.map[1]
    nat.check(nat.lib.op2hip_mat_program_device(
        plan.nactive, P(d['blk_t0']), P(d['blk_t1']), P(d['ex_off']), P(d['exec']),
        P(d['rng_lo']), P(d['rng_n']), P(d['tp']), P(d['contrib']), int(plan.c16), plan.nac,
        P(cmap._device_values()), cmap.dim, e1, nnodes, P(sp._dev_rowptr), P(sp._dev_colidx),
        rd, cd, npairs, wide, P(work), int(work.numel()), P(d_poff), P(d_prog),
        ctypes.c_void_p(stream())), 'device Mat program')
    return (d_poff, d_prog, wide)


def _owner_columns(args):
    """(map, idx or -1) of every race argument, in argument order (a Mat's
    race columns are its row map's)."""
    cols = []
    for a in args:
        if a._is_mat:
            cols.append((a.map[0], -1))
        elif a._is_indirect and a.access is INC:
            idx = a.idx if isinstance(a.idx, int) else -1
            cols.append((a.map, idx))
    return cols


def get_owner_plan(iterset, n, args, slot_bytes, k=None, block=256, lds_of=None,
                   lds_limit=None):
    """Cached owner plan.  The owned range ``k`` is chosen so a block runs
    about one element per thread (a first build measures elements run per
    owned target), and shrunk until a block's LDS (``lds_of(plan)``; by
    default ``slot_bytes`` per element) fits ``lds_limit`` (default
    ``OWNER_LDS_LIMIT``) bytes."""
    if lds_of is None:
        lds_of = lambda p: p.max_exec * slot_bytes
    cols = _owner_columns(args)
    target = cols[0][0].dataset
    k_cfg = k or cfg.get('hip_owner_targets')
    key = (id(iterset), n, k_cfg, block, cfg.get('hip_owner_fill'), lds_limit,
           tuple((id(m), i) for m, i in cols))
    if key in _owner_cache:
        return _owner_cache[key]
    lay = layout.of(iterset)
    if cfg.get('hip_device_plan', True):
        # increments are summed in user element order (the sequential
        # order): device element e is user element perm[e]
        rank = None if lay.identity else _device_perm(lay)

        def build(kk):
            return DeviceOwnerPlan(n, target.total_size, cols, kk, rank)
    else:
        # the host build: the device position of user element i is iperm[i]
        order = None if lay.identity else lay.iperm[:n]
        vals = [(m._layout_values()[:n], m.dim, i) for m, i in cols]

        def build(kk):
            return OwnerPlan(n, target.total_size, vals, kk, order)

    # the kernels' static LDS (the deferred combine pass's 128 bytes, Global
    # reduction scratch) comes on top of the dynamic bytes sized here
    limit = cfg.get('hip_owner_lds') or (min(OWNER_LDS_LIMIT, _state.get('lds_per_block') or 65536)
                                         - STATIC_LDS_MARGIN)
    if lds_limit:
        limit = min(limit, lds_limit)
    if k_cfg:
        kk = int(k_cfg)
        p = build(kk)
    else:
        kk = 128
        p = build(kk)
        if p.nactive:
            # elements run and LDS bytes per owned target, from the trial
            per_target = p.exec_per_block / kk
            lds_per_target = lds_of(p) / float(kk)
            fill = cfg.get('hip_owner_fill') or 1.0
            kk = min(1 << 16, max(1, int(fill * block / max(per_target, 1e-9))))
            kk = max(1, min(kk, int(limit / max(lds_per_target, 1e-9))))
            p = build(kk)
    while lds_of(p) > limit and kk > 1:
        kk = max(1, int(kk * 0.8))
        p = build(kk)
    if lds_of(p) > limit:
        _owner_cache[key] = None
        return None
    if p.nactive and p.exec_per_block < 32 and not k_cfg:
        # blocks would run too few elements to keep a workgroup busy
        _owner_cache[key] = None
        return None
    p._keep = [m for m, _ in cols]
    _owner_cache[key] = p
    return p


# -- loops -------------------------------------------------------------------

class _CompiledLoop(object):

    def __init__(self, source):
        self.source = source
        self.module = jit.Module(jit.compile_source(source.source()))
        self.fn = self.module.function(source.entry)
        self.reduce_fn = self.module.function(source.reduce_entry) if source.reduce_entry else None
        self.const_versions = {}


_loop_cache = {}
# generated sources not compiled yet, by the same keys
_source_cache = {}


class JITModule(base.JITModule):
    """Cache key holder of compiled loops (key as the reference's JITModule)."""
    _cache = {}

    def __init__(self, kernel, itspace_extents, *args, **kwargs):
        pass


def _block_size():
    return cfg.get('hip_block_size') or 256


class _NoOwnerPlan(Exception):
    """The loop's owner plan would not fit LDS with enough work per block."""


class _LaunchEntry(object):
    """Everything one recurring ``par_loop`` call needs, resolved once: the
    compiled loop, device pointers packed into ctypes argument arrays per
    launch, reduction scratch, and the coherence updates to apply.  Holds
    strong references to the loop's data objects, so the pointers stay
    valid for as long as the entry is cached."""

    __slots__ = ('loop', 'launches', 'to_device', 'written', 'dats', 'globals_red', 'mats',
                 'consts', 'keep', 'has_reduce', 'red_index', 'red_info', 'red_glob')

    def rebindable(self, args):
        """A copy whose Global arguments can be exchanged for others of
        the same shape (``rebind``): its own argument arrays, and per Global
        argument position the (launch, slot) entries holding its device
        pointer.  Reference-style drivers build a new Global per call
        (demo/airfoil.py:124-130: ``rms = op2.Global(...)`` in the loop)."""
        c = object.__new__(_LaunchEntry)
        for k in ('loop', 'consts', 'dats', 'mats', 'globals_red', 'has_reduce', 'red_index',
                  'red_info', 'red_glob'):
            setattr(c, k, getattr(self, k, None))
        c.launches = [(fn, grid, bs, ((ctypes.c_uint64 * n)(*arr), n), lds)
                      for fn, grid, bs, (arr, n), lds in self.launches]
        c.to_device = list(self.to_device)
        c.written = list(self.written)
        c.keep = list(self.keep)
        slots = {}
        for pos, a in enumerate(args):
            if a._is_global:
                ptr = _ptr(a.data._dev)
                slots[pos] = (a.data, [(li, ai) for li, (_, _, _, (arr, n), _) in
                                       enumerate(c.launches) for ai in range(n)
                                       if arr[ai] == ptr])
        c.globals_red = slots
        return c

    def rebind(self, args):
        """Point the Global arguments at ``args``' Globals."""
        for pos, (cur, where) in list(self.globals_red.items()):
            new = args[pos].data
            if new is cur:
                continue
            new._to_device()
            ptr = _ptr(new._dev)
            for li, ai in where:
                self.launches[li][3][0][ai] = ptr
            self.to_device = [new if d is cur else d for d in self.to_device]
            self.written = [new if d is cur else d for d in self.written]
            if self.red_glob is cur:
                self.red_glob = new
            # the bound Globals stay alive through globals_red / to_device
            self.globals_red[pos] = (new, where)

    def __init__(self, pl, mode, bs):
        consts = base.Const._definitions()
        self.consts = consts
        args = pl.args
        n = pl._it_space.exec_size if pl.needs_exec_halo else pl._it_space.size
        iterset = pl._it_space.iterset
        self.keep = [pl.kernel, args, iterset]
        self.globals_red = None
        # direct Dats on another Set of the same size share the iteration
        # set's layout (element e is one position in both)
        for a in args:
            if a._is_direct and a.data.dataset is not iterset:
                layout.link(a.data.dataset, iterset)
        if mode in ('owner', 'rows'):
            m0, i0 = _owner_columns(args)[0]
            layout.prefer(iterset, m0, max(i0, 0))
        layout.of(iterset)
        self.to_device = []
        for a in args:
            if (a._is_dat or a._is_global) and a.data not in self.to_device:
                self.to_device.append(a.data)
        for d in self.to_device:
            d._to_device()
        # the generated source (layouts decided: it names which arguments
        # share a map array) decides the owner plan; it is compiled once the
        # plan exists (a loop falling back to colouring compiles only its
        # coloured form)
        _, src = pl._source(mode, bs, consts)
        plan = None
        lds = 0
        if mode == 'rows':
            limit = min(int(cfg.get('hip_mat_rows_lds') or ROWS_LDS_LIMIT),
                        (_state.get('lds_per_block') or 65536) - STATIC_LDS_MARGIN)
            smaps = [args[m].map for m in src.rows_maps()] if src.rows_stage else []
            plan = get_rows_plan(iterset, n, args, bs, src.rows_lds_bytes, limit, smaps)
            if plan is None:
                raise _NoOwnerPlan()
            nparts = plan.nblk
            stage_max = max([plan.stage(m)[3] for m in smaps] or [0])
            lds = src.rows_lds_bytes(plan.k, plan.max_slice, stage_max)
            grid_stream = 0
        elif mode == 'owner':
            slot_bytes = src.owner_slot_bytes()
            lds_limit = None
            if any(a._is_mat for a in args):
                # Mat blocks keep a local-matrix row per contribution: their
                # LDS is kept to twice a CU's share per block at full
                # occupancy (4 of 8 waves per SIMD), so wide element matrices
                # get smaller owned ranges rather than fewer blocks per CU
                # (128^3: 256 threads with 40 KB measured 0.995 / 1.96 ms for
                # the Laplace / vector mass against 1.011 / 2.05 ms at 20 KB)
                lds_limit = int(cfg.get('hip_owner_mat_lds') or
                                2 * LDS_PER_CU * bs // (64 * WAVES_PER_CU))
            plan = get_owner_plan(iterset, n, args, slot_bytes, block=bs,
                                  lds_of=lambda p: src.owner_lds_bytes(p.k, p.max_exec, p.max_codes),
                                  lds_limit=lds_limit)
            if plan is None:
                raise _NoOwnerPlan()
            nparts = max(plan.nactive, 1)
            lds = src.owner_lds_bytes(plan.k, plan.max_exec, plan.max_codes)
            grid_stream = 0
        elif mode == 'coloured':
            plan = get_plan(pl._it_space.iterset, n, args,
                            part_size=cfg.get('hip_partition_size') or bs)
            nparts = plan.nblocks
            grid_stream = 0
        else:
            chunk = bs * src.stream_unroll()
            grid_stream = max(1, min((n + chunk - 1) // chunk,
                                     int(cfg.get('hip_stream_blocks') or 4096)))
            nparts = grid_stream
        loop = pl._compiled(mode, bs, consts)
        self.loop = loop
        partials = {}
        vals = []
        for role, pos in loop.source.param_roles():
            a = args[pos] if pos is not None else None
            if role in ('start', 'count') or role.startswith('plan_') or role.startswith('own_'):
                vals.append(role)
            elif role == 'gbl':
                vals.append(_ptr(a.data._dev))
            elif role == 'partial':
                t = _dev_bytes(max(nparts, 1) * a.data.cdim * a.data.dtype.itemsize)
                partials[pos] = t
                vals.append(_ptr(t))
            elif role == 'data':
                vals.append(_ptr(a.data._dev))
            elif role == 'rows_xmap':
                vals.append(_ptr(plan.xmap(a.map)))
            elif role in ('rows_sn', 'rows_so', 'rows_xl'):
                vals.append(_ptr(plan.stage(a.map)[('rows_sn', 'rows_so', 'rows_xl').index(role)]))
            elif role == 'map':
                vals.append(_ptr(a.map._device_values()))
            elif role == 'mat_vals':
                vals.append(_ptr(a.data._dev_vals))
            elif role == 'mat_rowptr':
                vals.append(_ptr(a.data.sparsity._dev_rowptr))
            elif role == 'mat_colidx':
                vals.append(_ptr(a.data.sparsity._dev_colidx))
            elif role == 'mat_rmap':
                vals.append(_ptr(a.map[0]._device_values()))
            elif role == 'mat_cmap':
                vals.append(_ptr(a.map[1]._device_values()))
            elif role == 'err_flag':
                vals.append(_ptr([x for x in args if x._is_mat][0].data._err))
            elif role.startswith('pend_'):
                vals.append(0)
            else:
                raise RuntimeError("unknown parameter role %s" % role)
        self.keep.append(partials)
        launches = []
        if n > 0:
            if mode == 'stream':
                argv = [0 if r == 'start' else n if r == 'count' else r for r in vals]
                launches.append((loop.fn, grid_stream, bs, _argv(argv), 0))
            elif mode == 'rows':
                d = plan.device_arrays()
                self.keep.append(plan)
                names = {'own_tp': 'tp', 'own_xoff': 'xoff', 'own_exec': 'exec',
                         'own_contrib': 'code', 'own_pos': 'pos'}
                scal = {'start': 0, 'count': plan.nblk, 'own_c16': int(plan.c16),
                        'own_k': plan.k, 'own_nnodes': plan.nnodes, 'own_slice': plan.max_slice,
                        'own_stage': stage_max}
                argv = [scal[r] if r in scal else _ptr(d[names[r]]) if r in names else r
                        for r in vals]
                if plan.nblk:
                    launches.append((loop.fn, plan.nblk, bs, _argv(argv), lds))
            elif mode == 'owner':
                pdev = plan.device_arrays()
                self.keep.append(plan)
                prog = (None, None, 0)
                if any(x._is_mat for x in args):
                    ma = [x for x in args if x._is_mat][0]
                    prog = _mat_program(plan, ma, pl._it_space.extents[1])
                names = {'own_t0': 'blk_t0', 'own_t1': 'blk_t1', 'own_tgcb': 'tg_cb',
                         'own_rlo': 'rng_lo', 'own_rn': 'rng_n',
                         'own_tgoff': 'tg_off', 'own_tgt': 'tgt', 'own_exoff': 'ex_off',
                         'own_exec': 'exec', 'own_tp': 'tp', 'own_contrib': 'contrib'}
                argv = [0 if r == 'start' else plan.nactive if r == 'count' else
                        plan.max_exec if r == 'own_pitch' else plan.k if r == 'own_k' else
                        max(plan.max_codes, 1) if r == 'own_cpitch' else
                        _ptr(prog[0]) if r == 'own_poff' else _ptr(prog[1]) if r == 'own_prog' else
                        prog[2] if r == 'own_wide' else int(plan.c16) if r == 'own_c16' else
                        _ptr(pdev[names[r]]) if isinstance(r, str) else r for r in vals]
                if plan.nactive:
                    launches.append((loop.fn, plan.nactive, bs, _argv(argv), lds))
            else:
                pdev = plan.device_arrays()
                self.keep.append(plan)
                names = {'plan_elems': 'elems', 'plan_blkmap': 'blkmap',
                         'plan_offset': 'offset', 'plan_nelem': 'nelems',
                         'plan_nthrcol': 'nthrcol', 'plan_thrcol': 'thrcol'}
                for col in range(plan.ncolors):
                    cnt = int(plan.ncolblk[col])
                    if cnt == 0:
                        continue
                    argv = []
                    for r in vals:
                        if r == 'start':
                            argv.append(int(plan.colour_start[col]))
                        elif r == 'count':
                            argv.append(cnt)
                        elif isinstance(r, str):
                            argv.append(_ptr(pdev[names[r]]))
                        else:
                            argv.append(r)
                    launches.append((loop.fn, cnt, bs, _argv(argv), 0))
            if loop.reduce_fn is not None:
                rargv = []
                for role, pos in loop.source.reduce_roles():
                    if role == 'nparts':
                        rargv.append(nparts)
                    elif role == 'gbl_out':
                        rargv.append(_ptr(args[pos].data._dev))
                    else:
                        rargv.append(_ptr(partials[pos]))
                launches.append((loop.reduce_fn, 1, 1024, _argv(rargv), 0))
        self.launches = launches
        # one FP64 Global reduction: its combine pass is deferred to the
        # next launch (see _Pending)
        self.red_index = self.red_info = self.red_glob = None
        reds = [a for a in args if a._is_global_reduction]
        if loop.reduce_fn is not None and launches and launches[-1][0] is loop.reduce_fn:
            self.has_reduce = True
            op = _red_mode(reds[0].access) if len(reds) == 1 else None
            if (op is not None and reds[0].data.dtype == np.float64 and
                    reds[0].data.cdim < 0x10000 and nparts * reds[0].data.cdim < 2 ** 31 and
                    cfg.get('hip_defer_reductions', True)):
                self.red_index = len(launches) - 1
                self.red_info = reds[0].data.cdim | (op << 16)
                self.red_glob = reds[0].data
        else:
            self.has_reduce = loop.reduce_fn is not None
        self.written = [a.data for a in args if (a._is_dat and a.access is not READ) or
                        a._is_global_reduction]
        self.dats = [a.data for a in args if a._is_dat]
        self.mats = [a.data for a in args if a._is_mat]

    def prepare(self):
        """Upload changed Consts and host-side data (synchronising work that
        cannot run inside a graph capture)."""
        s = stream()
        loop = self.loop
        for c in self.consts:
            if loop.const_versions.get(c.name) != (id(c), c._version):
                ptr, nbytes = loop.module.global_symbol(c.name)
                data = np.ascontiguousarray(c.data)
                nat.check(nat.lib.op2hip_set_global(ptr, data.ctypes.data_as(ctypes.c_void_p),
                                                    min(nbytes, data.nbytes), ctypes.c_void_p(s)),
                          'const upload')
                loop.const_versions[c.name] = (id(c), c._version)
        for d in self.to_device:
            if d._state == HOST:
                d._to_device()

    def run(self):
        self.prepare()
        cur = stream()
        sp = ctypes.c_void_p(cur)
        launch = nat.lib.op2hip_launch
        pend = _pending[0]
        if pend is not None and (self.has_reduce or pend.stream != cur or
                                 any(d is pend.glob for d in self.to_device)):
            # this loop reduces (its partials, its Global) or uses the
            # deferred Global: the combine pass runs first, on its own
            flush()
            pend = None
        red, fn0 = self.red_index, self.loop.fn
        for li, (fn, grid, bs, (arr, n), lds) in enumerate(self.launches):
            if li == red:
                _pending[0] = _Pending(fn, arr, self.red_info, self.red_glob, self, cur)
                continue
            if pend is not None and fn is fn0 and bs % 64 == 0:
                # the deferred combine pass rides along as one more workgroup
                arr[n - 4], arr[n - 3], arr[n - 2], arr[n - 1] = (
                    pend.argv[0], pend.argv[2], pend.argv[1], pend.info)
                st = launch(fn, grid + 1, bs, lds, sp, arr, n)
                arr[n - 4] = arr[n - 3] = arr[n - 2] = arr[n - 1] = 0
                _pending[0] = pend = None
            else:
                st = launch(fn, grid, bs, lds, sp, arr, n)
            if st != 0:
                nat.check(st, 'kernel launch')
        for d in self.written:
            d._state = DEVICE
        for d in self.dats:
            # host views go read-only once (Dat.data makes them writable again)
            a = d._data
            if a.flags.writeable:
                maybe_setflags(a, write=False)


class _Pending(object):
    """The combine pass of an FP64 Global reduction whose loop has run: the
    next loop launch that neither reduces nor names the Global runs it in
    one extra workgroup (codegen's op2_pending_reduce, in the combine
    kernel's order), else :func:`flush` launches the combine kernel."""

    __slots__ = ('fn', 'argv', 'info', 'glob', 'keep', 'stream')

    def __init__(self, fn, arr, info, glob, keep, strm):
        # a copy of the combine launch's arguments (nparts, Global,
        # partials): a later call of the same loop may rebind its Global
        self.fn, self.argv, self.info = fn, (ctypes.c_uint64 * 3)(*arr[:3]), info
        self.glob, self.keep, self.stream = glob, keep, strm


# at most one deferred reduction, in stream order behind the launches made
_pending = [None]


def _red_mode(access):
    """Combine code of a Global reduction's access: INC 0, MIN 1, MAX 2."""
    return {'INC': 0, 'MIN': 1, 'MAX': 2}.get(_access_mode(access))


def flush():
    """Launch the deferred Global reduction, if any, on the stream its loop
    ran on.  Reading a Global's value, capturing a graph, a launch on
    another stream and loops that use the Global or reduce themselves do
    this first."""
    p = _pending[0]
    if p is None:
        return
    _pending[0] = None
    st = nat.lib.op2hip_launch(p.fn, 1, 1024, 0, ctypes.c_void_p(p.stream), p.argv, 3)
    if st != 0:
        nat.check(st, 'kernel launch')


def _argv(values):
    n = len(values)
    return ((ctypes.c_uint64 * n)(*[int(v) & 0xFFFFFFFFFFFFFFFF for v in values]), n)


_call_cache = {}
_CALL_CACHE_MAX = 512


def _arg_key(a):
    k = a._fkey
    if k is None:
        # an Arg never changes once made: its key is computed once
        k = a._fkey = _arg_key_of(a)
    return k


def _arg_key_of(a):
    m = a.map
    if isinstance(m, (list, tuple)):
        mk = tuple(id(x) for x in m)
    else:
        mk = id(m)
    idx = a.idx
    if isinstance(idx, (list, tuple)):
        ik = tuple(getattr(x, 'index', x) for x in idx)
    elif isinstance(idx, IterationIndex):
        ik = ('i', idx.index)
    else:
        ik = idx
    return (id(a.data), mk, ik, id(a.access))


class ParLoop(base.ParLoop):

    def _mode(self):
        args = self.args
        mats = [a for a in args if a._is_mat]
        race = [a for a in args if a._is_indirect and a.access is INC]
        if not mats and not race:
            return 'stream'
        if not cfg.get('hip_owner', True):
            return 'coloured'
        ext = self._it_space.extents
        if mats:
            # owner-computes Mat assembly: one Mat, no Dat increments, the
            # iteration space spanning both maps' arities
            if len(mats) > 1 or race or len(ext) != 2:
                return 'coloured'
            a = mats[0]
            if ext[0] != a.map[0].dim or ext[1] != a.map[1].dim:
                return 'coloured'
            if a.idx[0].index != 0 or a.idx[1].index != 1:
                return 'coloured'
            if any(x._is_dat and x.access is not READ for x in args):
                return 'coloured'
            return 'owner'
        # owner-computes Dat increments: into one Dat, every other access a
        # read (an element may run in two blocks; only its increments to
        # owned targets and, in its primary block, its Global reductions count)
        dat = race[0].data
        if any(a.data is not dat for a in race):
            return 'coloured'
        for a in race:
            if isinstance(a.idx, IterationIndex) and ext[a.idx.index] != a.map.dim:
                return 'coloured'
        for a in args:
            if a._is_dat and a not in race and a.access is not READ:
                return 'coloured'
            if a._is_dat and a not in race and a.data is dat:
                return 'coloured'
        return 'owner'

    def _owner_stage(self):
        """Arguments an owner loop may stage in LDS: read-only FP64 Dats
        gathered through a single map column whose Set shares the layout of
        the increment target (so the block's owned positions are theirs)."""
        args = self.args
        race = [a for a in args if a._is_indirect and a.access is INC]
        if not race or any(a._is_mat for a in args):
            return ()
        tgroup = layout._group(race[0].data.dataset)
        written = set(id(a.data) for a in args if not a._is_global and a.access is not READ)
        out = []
        for k, a in enumerate(args):
            if (a._is_indirect and a.access is READ and not a._is_vec_map and not a._is_soa and
                    not isinstance(a.idx, IterationIndex) and a.data.dtype == np.float64 and
                    id(a.data) not in written and a.data.cdim <= 16 and
                    layout._group(a.data.dataset) is tgroup):
                out.append(k)
        return tuple(out)

    def _compiled(self, mode, bs, consts):
        """The compiled loop of this loop's signature for ``mode`` (cached)."""
        key, src = self._source(mode, bs, consts)
        c = _loop_cache.get(key)
        if c is None:
            c = _CompiledLoop(src)
            _loop_cache[key] = c
        return c

    def _source(self, mode, bs, consts):
        """(cache key, generated source) of this loop's signature for
        ``mode`` under the current configuration."""
        key = base.JITModule._cache_key(self.kernel, self._it_space.extents, *self.args)
        # structure-of-arrays Dats are indexed with their row count compiled in
        key += (tuple(a.data.dataset.total_size if a._is_dat and a._is_soa else 0
                      for a in self.args),)
        pf = int(cfg.get('hip_owner_prefetch', 1))
        sr = bool(cfg.get('hip_stage_reads', False))
        sw = bool(cfg.get('hip_stage_writes', True))
        un = int(cfg.get('hip_stream_unroll', 1))
        st = self._owner_stage() if mode == 'owner' and cfg.get('hip_owner_stage') else ()
        rs = mode == 'rows' and bool(cfg.get('hip_mat_rows_stage'))
        mat = any(a._is_mat for a in self.args)
        mw = int(cfg.get('hip_owner_waves', 0)) if mode == 'owner' and not mat else 0
        alias = self._map_alias()
        key = key + (mode, bs, pf, sr, sw, un, st, mw, alias, rs)
        c = _loop_cache.get(key)
        if c is not None:
            return key, c.source
        src = _source_cache.get(key)
        if src is None:
            src = LoopSource(self.kernel, self._it_space.extents, self.args, consts, mode, bs,
                             prefetch=pf, stage_reads=sr, stage_writes=sw, unroll=un,
                             owner_stage=st, rows_stage=rs)
            src.min_waves = mw
            src.map_of = {i: r for i, r in enumerate(alias) if r != i}
            if len(_source_cache) >= _CALL_CACHE_MAX:
                _source_cache.pop(next(iter(_source_cache)))
            _source_cache[key] = src
        return key, src

    def _map_alias(self):
        """Per argument, the first argument whose map has the same device
        array (and arity); Dat arguments through maps that share one array
        (the same Map, or identical maps whose target Sets share a layout)
        then read one map parameter, so the kernel loads each map entry once.
        Needs the layouts decided: it asks for the maps' device arrays."""
        seen, out = {}, []
        for i, a in enumerate(self.args):
            r = i
            if a._is_indirect:
                r = seen.setdefault((a.map._device_values().data_ptr(), a.map.dim), i)
            out.append(r)
        return tuple(out)

    def _entry(self, mode):
        """A launch entry for ``mode`` under the current configuration
        (owner loops whose plan does not fit fall back to colouring)."""
        bs = _block_size()
        if mode == 'owner' and any(a._is_mat for a in self.args):
            # Mat assembly keeps a local-matrix row per contribution in LDS:
            # smaller blocks and owned ranges hide its latency better
            obs = cfg.get('hip_owner_mat_block_size') or bs
            fill = cfg.get('hip_owner_mat_fill') or cfg.get('hip_owner_fill')
        else:
            obs = cfg.get('hip_owner_block_size') or bs
            fill = cfg.get('hip_owner_fill')
        if (mode == 'owner' and any(a._is_mat for a in self.args) and
                cfg.get('hip_mat_rows', True) and
                not any(a._is_global_reduction for a in self.args)):
            # row blocks first for small local matrices: each element run by
            # fewer blocks, the CSR slice summed in LDS.  (Larger ones, e.g.
            # the 2x2-block vector mass with 64 values, keep the owner
            # template: their chunk of local matrices crowds the LDS.)
            m = [a for a in self.args if a._is_mat][0]
            ext = self._it_space.extents
            if ext[0] * ext[1] * m.data.dims[0] * m.data.dims[1] <= ROWS_MAX_ENTRIES:
                try:
                    return _LaunchEntry(self, 'rows', cfg.get('hip_mat_rows_block_size') or bs)
                except _NoOwnerPlan:
                    pass
        try:
            with cfg.override(hip_owner_fill=fill):
                return _LaunchEntry(self, mode, obs if mode == 'owner' else bs)
        except _NoOwnerPlan:
            return _LaunchEntry(self, 'coloured', bs)

    def _call_key(self):
        return (id(self.kernel), self._it_space.extents, id(self._it_space.iterset),
                _variant_key(), tuple(_arg_key(a) for a in self.args), _const_key())

    def _entry_cached(self):
        """The cached launch entry of this loop under the current (and, with
        the autotuner on, tuned) configuration."""
        if cfg.get('hip_autotune'):
            mode = self._mode()
            params = autotune.params_for(self, mode, lambda cand: _trial(self, mode, cand))
            if params:
                with cfg.override(**params):
                    return self._entry_for_key()
        return self._entry_for_key()

    def _entry_for_key(self):
        key = self._call_key()
        ent = _call_cache.get(key)
        if ent is None:
            if len(_call_cache) >= _CALL_CACHE_MAX:
                _call_cache.pop(next(iter(_call_cache)))
            ent = self._entry(self._mode())
            _call_cache[key] = ent
        return ent

    def compute(self):
        self._entry_cached().run()
        self.maybe_set_halo_update_needed()


_key_cache = {'cfg': (None, None), 'const': (None, None)}


def _variant_key():
    """The configuration entries of _VARIANT_KEYS, recomputed only when the
    configuration changed."""
    v = cfg.version()
    hit = _key_cache['cfg']
    if hit[0] != v:
        hit = (v, tuple(cfg.get(k) for k in _VARIANT_KEYS))
        _key_cache['cfg'] = hit
    return hit[1]


def _const_key():
    """(name, identity) of the declared Consts, recomputed only when one is
    declared or removed."""
    v = base.Const._defs_version
    hit = _key_cache['const']
    if hit[0] != v:
        hit = (v, tuple((c.name, id(c)) for c in base.Const._definitions()))
        _key_cache['const'] = hit
    return hit[1]


# configuration entries that shape a loop's launches (part of the call key)
_VARIANT_KEYS = ('hip_block_size', 'hip_owner_block_size', 'hip_owner_fill', 'hip_owner_stage',
                 'hip_owner_mat_block_size', 'hip_owner_mat_fill',
                 'hip_owner_targets', 'hip_owner', 'hip_partition_size', 'hip_reorder',
                 'hip_stage_reads', 'hip_stage_writes', 'hip_stream_unroll',
                 'hip_stream_blocks', 'hip_owner_prefetch', 'hip_owner_waves',
                 'hip_owner_mat_lds', 'hip_colouring', 'hip_mat_rows', 'hip_mat_rows_block_size',
                 'hip_mat_rows_k', 'hip_mat_rows_lds', 'hip_mat_rows_stage',
                 'hip_defer_reductions')


def _written_buffers(pl):
    """Device tensors a loop writes (Dats, reduced Globals, Mat values)."""
    out = []
    for a in pl.args:
        if a._is_mat:
            out.append((a.data, a.data._dev_vals))
        elif (a._is_dat and a.access is not READ) or a._is_global_reduction:
            a.data._to_device()
            out.append((a.data, a.data._dev))
    return out


def _trial(pl, mode, params, budget_s=2e-3):
    """Seconds per call of ``pl`` launched under ``params``; the data it
    writes is restored afterwards."""
    with cfg.override(**params):
        ent = pl._entry(mode)
    bufs = _written_buffers(pl)
    saved = [(d, t.clone(), getattr(d, '_state', None)) for d, t in bufs]
    s = ctypes.c_void_p(stream())
    ev = [ctypes.c_void_p(), ctypes.c_void_p()]
    for e in ev:
        nat.check(nat.lib.op2hip_event_create(ctypes.byref(e)), 'event')
    try:
        ent.run()
        torch.cuda.synchronize()

        def timed(reps):
            nat.lib.op2hip_event_record(ev[0], s)
            for _ in range(reps):
                ent.run()
            nat.lib.op2hip_event_record(ev[1], s)
            torch.cuda.synchronize()
            ms = ctypes.c_float()
            nat.check(nat.lib.op2hip_event_elapsed_ms(ev[0], ev[1], ctypes.byref(ms)), 'timing')
            return ms.value / 1e3 / reps

        one = timed(1)
        reps = int(min(50, max(3, budget_s / max(one, 1e-7))))
        return min(timed(reps) for _ in range(3))
    finally:
        flush()
        torch.cuda.synchronize()
        for d, t, st in saved:
            buf = d._dev_vals if isinstance(d, Mat) else d._dev
            buf.copy_(t)
            if st is not None:
                d._state = st
        torch.cuda.synchronize()
        for e in ev:
            nat.lib.op2hip_event_destroy(e)


# Eager calls already seen: (kernel, iteration set, extents, per argument
# its data, map(s), index and access) -> (state key, launch entry, Dats whose
# halos the loop dirties).  A call with the same objects was validated and
# planned when it was first seen (the entry keeps those objects alive, so
# their ids stay theirs), so a repeat under the same configuration and
# Const declarations goes straight to the launch.
_fast_cache = {}


def _fast_key(kernel, it_space, args):
    if isinstance(it_space, IterationSpace):
        head = (id(kernel), id(it_space.iterset), it_space.extents)
    else:
        head = (id(kernel), id(it_space), ())
    ks = []
    for a in args:
        k = a._fkey
        if k is None:
            if a._is_global:
                # Globals by shape and by which positions share one Global:
                # a cached call rebinds them
                k = ('gbl', a.data.cdim, a.data.dtype.char, id(a.access),
                     next(j for j, b in enumerate(args) if b.data is a.data))
            else:
                k = a._fkey = _arg_key_of(a)
        ks.append(k)
    return head + tuple(ks)


def par_loop(kernel, it_space, *args):
    """Run ``kernel`` over ``it_space`` (reference: op2.py:125-158)."""
    try:
        key = _fast_key(kernel, it_space, args)
    except (AttributeError, TypeError):
        key = None
    state = (cfg.version(), base.Const._defs_version)
    hit = _fast_cache.get(key) if key is not None else None
    if hit is not None and hit[0] == state:
        ent = hit[1]
        if ent.globals_red:
            ent.rebind(args)
        ent.run()
        for d in hit[2]:
            d.needs_halo_update = True
        return
    pl = ParLoop(kernel, it_space, *args)
    pl.compute()
    if key is not None and not cfg.get('hip_autotune'):
        if len(_fast_cache) >= _CALL_CACHE_MAX:
            _fast_cache.pop(next(iter(_fast_cache)))
        ent = pl._entry_cached()
        if any(a._is_global for a in pl.args):
            ent = ent.rebindable(pl.args)
        _fast_cache[key] = (state, ent,
                            [a.data for a in pl.args
                             if a._is_dat and a.access in (INC, WRITE, RW)])


@validate_type(('M', base.Mat, MatTypeError), ('x', base.Dat, DatTypeError),
               ('b', base.Dat, DatTypeError))
def solve(M, x, b):
    Solver().solve(M, x, b)


class Graph(object):
    """A captured sequence of op2 calls, replayed as one hipGraph.

    ``fn`` must have run eagerly before (so loops are compiled, plans built
    and scratch allocated) and must not touch host data (``.data``) or change
    Consts.  Replaying re-executes every captured launch on the current
    stream; host-side coherence states are those left by the capture."""

    def __init__(self, fn):
        self._g = torch.cuda.CUDAGraph()
        flush()
        with torch.cuda.graph(self._g):
            fn()
            # a reduction deferred by the last loop is part of the graph
            flush()

    def replay(self):
        self._g.replay()


def capture(fn):
    """Capture ``fn`` (a sequence of par_loops) into a :class:`Graph`."""
    return Graph(fn)
