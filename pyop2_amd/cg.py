"""Host side of the device batched Jacobi-PCG (csrc/cg_batch.hip).

A :class:`CGBatch` holds up to ``max_sys`` systems; each is a row range of
an FP64 CSR matrix with its vectors on the device.  ``solve`` runs systems
to completion, ``load``/``step``/``read`` let a caller (the serving loop)
add and retire systems between iteration launches.  PETSc KSPCG semantics
per system (reference: pyop2/petsc_base.py:172-220)."""

import ctypes

import numpy as np

from . import _native as nat

SYS_EMPTY, SYS_NEW, SYS_RUNNING, SYS_DONE = 0, 1, 2, 3

REASONS = {2: 'CONVERGED_RTOL', 3: 'CONVERGED_ATOL', -2: 'DIVERGED_NULL', -3: 'DIVERGED_ITS',
           -4: 'DIVERGED_DTOL', -5: 'DIVERGED_BREAKDOWN', -8: 'DIVERGED_INDEFINITE_MAT',
           -9: 'DIVERGED_NANORINF'}

_p = ctypes.c_void_p
_d = ctypes.c_double
_i = ctypes.c_int


class CGSystem(ctypes.Structure):
    """Mirror of ``op2hip_cg_system`` (include/op2hip.h)."""
    _fields_ = [('rowptr', _p), ('colidx', _p), ('vals', _p), ('b', _p), ('x', _p),
                ('inv_diag', _p), ('r', _p), ('zp', _p), ('w', _p),
                ('row0', _i), ('n', _i), ('chunk0', _i), ('nchunks', _i),
                ('rtol', _d), ('atol', _d), ('dtol', _d), ('maxit', _i), ('state', _i),
                ('no_precond', _i), ('beta', _d), ('betaold', _d), ('dpi', _d), ('dpiold', _d), ('dp', _d),
                ('rnorm0', _d), ('ttol', _d), ('alpha', _d), ('its', _i), ('reason', _i)]


CHUNK = None


def chunk_rows():
    global CHUNK
    if CHUNK is None:
        CHUNK = nat.lib.op2hip_cgb_chunk_rows()
    return CHUNK


def make_system(rowptr, colidx, vals, b, x, work, row0, n, rtol, atol, dtol, maxit,
                jacobi=True):
    """A NEW system over rows [row0, row0+n).  ``work`` holds the base
    addresses of the 5 rows of one (5, length) device array of doubles laid
    out like x (indexed by global row): inv_diag, r, then z and p
    interleaved over rows 2 and 3 (z_i at 2 i, p_i at 2 i + 1), and w.
    ``jacobi`` False runs CG without a preconditioner (PETSc's PCNONE)."""
    pitch = work[1] - work[0]
    if any(work[j + 1] - work[j] != pitch for j in range(4)) or pitch <= 0:
        raise ValueError("work must be the rows of one (5, length) array")
    if work[2] % 16:
        raise ValueError("work rows must be 16-byte aligned")
    s = CGSystem()
    s.rowptr, s.colidx, s.vals, s.b, s.x = rowptr, colidx, vals, b, x
    s.inv_diag, s.r, s.zp, s.w = work[0], work[1], work[2], work[4]
    s.row0, s.n = row0, n
    s.rtol, s.atol, s.dtol, s.maxit = rtol, atol, dtol, maxit
    s.state = SYS_NEW
    s.no_precond = 0 if jacobi else 1
    return s


class CGBatch(object):

    def __init__(self, max_sys, max_rows, hist_len=0):
        self.max_sys = max_sys
        ch = chunk_rows()
        self.max_chunks = max(1, (max_rows + ch - 1) // ch + max_sys)
        h = ctypes.c_void_p()
        nat.check(nat.lib.op2hip_cgb_create(max_sys, self.max_chunks, hist_len, ctypes.byref(h)),
                  'cg batch create')
        self._h = h
        self.hist_len = hist_len
        self.systems = (CGSystem * max_sys)()
        self.nsys = 0

    def load(self, systems, stream):
        n = len(systems)
        for k, s in enumerate(systems):
            self.systems[k] = s
        self.nsys = n
        nat.check(nat.lib.op2hip_cgb_load(self._h, n, ctypes.byref(self.systems),
                                          ctypes.c_void_p(stream)), 'cg batch load')

    def step(self, iters, stream, graph=False):
        """Enqueue ``iters`` iterations; ``graph`` replays them from a
        captured hipGraph (kept while the chunk list and count stay)."""
        f = nat.lib.op2hip_cgb_step_graph if graph else nat.lib.op2hip_cgb_step
        nat.check(f(self._h, iters, ctypes.c_void_p(stream)), 'cg step')

    def read(self, stream):
        nat.check(nat.lib.op2hip_cgb_read(self._h, ctypes.byref(self.systems),
                                          ctypes.c_void_p(stream)), 'cg batch read')
        return [self.systems[k] for k in range(self.nsys)]

    def history(self, k, stream):
        out = np.zeros(self.hist_len)
        nat.check(nat.lib.op2hip_cgb_history(self._h, k, out.ctypes.data_as(ctypes.c_void_p),
                                             ctypes.c_void_p(stream)), 'cg history')
        return out

    def solve(self, systems, stream, chunk=16, graph=False):
        """Run ``systems`` (all NEW) to completion; returns their final
        descriptors.  With ``graph`` the steps are replayed hipGraphs of
        ``chunk`` iterations, and finished systems leave the launches only
        when they hold half of the rows launched (each reload re-captures)."""
        self.load(systems, stream)
        maxit = max(s.maxit for s in systems) if systems else 0
        ch = chunk_rows()
        live = lambda ss: sum((s.n + ch - 1) // ch for s in ss if s.state != SYS_DONE)
        loaded = live(systems)
        done_iters = 0
        ndone = 0
        while True:
            k = chunk if (done_iters or graph) else min(chunk, 4)
            self.step(k, stream, graph)
            done_iters += k
            out = self.read(stream)
            d = sum(s.state == SYS_DONE for s in out)
            if d == len(out) or done_iters > maxit + chunk:
                return out
            if d > ndone and (not graph or 2 * live(out) <= loaded):
                # drop finished systems' rows from the launches
                self.load(out, stream)
                ndone = d
                loaded = live(out)

    def __del__(self):
        try:
            nat.lib.op2hip_cgb_destroy(self._h)
        except Exception:
            pass
