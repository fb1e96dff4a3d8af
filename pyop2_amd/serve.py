"""Serving loop: assembly + CG requests on a resident pool of meshes.

Every pool mesh keeps its sets, maps, coordinates, Sparsity, device layout
and loop plans resident on the device.  A request names a mesh and brings a
diffusion coefficient per vertex; it is served by assembling A = M + K(kappa)
and b = M f on that mesh and solving A x = b with CG to ``rtol``.

The engine batches continuously at CG-iteration granularity: each engine
step admits the requests that have arrived, as long as a slot of their mesh
is free or the HBM budget allows a new one (a slot holds one request's
matrix values, vectors and CG work space), assembles the admitted requests,
then advances every admitted system by ``iters`` iterations of one batched
CG launch sequence.  Systems that converged leave the batch and free their
slot; new ones join at the next step while the others keep iterating.  A
request's result does not depend on what it shares the device with: its
assembly sums each entry's element contributions in element order, and its
CG reductions are per chunk of its own rows (csrc/cg_batch.hip).

``serve_sequential`` is the plain path for comparison: one request at a
time, in arrival order, through ``op2`` par_loops and ``Solver.solve``.
"""

import time

import numpy as np

from . import cg
from . import layout
from .workloads import fem


class ResidentMesh(object):
    """One pool mesh with its device-resident structure (2-D triangles or
    3-D tetrahedra, P1)."""

    def __init__(self, W, mesh, name):
        self.W = W
        self.mesh = mesh
        d = mesh.coords.shape[1]
        self.dim = d
        nv = d + 1
        self.n = mesh.nvert
        self.nodes = W.Set(mesh.nvert, 1, name + '_nodes')
        self.xnodes = W.Set(mesh.nvert, d, name + '_xnodes')
        self.elems = W.Set(mesh.nelem, 1, name + '_elems')
        self.en = W.Map(self.elems, self.nodes, nv, mesh.elem_node, name + '_en')
        self.ex = W.Map(self.elems, self.xnodes, nv, mesh.elem_node, name + '_ex')
        self.coords = W.Dat(self.xnodes, mesh.coords.copy(), np.float64, name + '_x')
        self.f = W.Dat(self.nodes, np.ones(mesh.nvert), np.float64, name + '_f')
        self.sparsity = W.Sparsity((self.en, self.en), name + '_sp')
        if d == 2:
            self.k_mass = W.Kernel(fem.MASS_TRI, 'mass_tri')
            self.k_lap = W.Kernel(fem.LAPLACE_TRI, 'laplace_tri')
            self.k_rhs = W.Kernel(fem.RHS_TRI, 'rhs_tri')
        else:
            self.k_mass = W.Kernel(fem.MASS_TET, 'mass_tet')
            self.k_lap = W.Kernel(fem.LAPLACE_TET, 'laplace_tet')
            self.k_rhs = W.Kernel(fem.RHS_TET, 'rhs_tet')
        self.free = []
        self.nslots = 0
        self.name = name

    @property
    def slot_bytes(self):
        return self.sparsity.nz * 8 + self.n * 8 * 9

    def new_slot(self):
        self.nslots += 1
        return Slot(self, '%s_s%d' % (self.name, self.nslots))


class Slot(object):
    """One request's device state on a resident mesh."""

    def __init__(self, rm, name):
        from . import hip
        W = rm.W
        self.rm = rm
        self.kappa = W.Dat(rm.nodes, np.ones(rm.n), np.float64, name + '_k')
        self.b = W.Dat(rm.nodes, np.zeros(rm.n), np.float64, name + '_b')
        self.x = W.Dat(rm.nodes, np.zeros(rm.n), np.float64, name + '_u')
        self.A = W.Mat(rm.sparsity, np.float64, name + '_A')
        self.work = hip.torch.empty((5, max(rm.n, 1)), dtype=hip.torch.float64,
                                    device=hip.device())

    def assemble(self, kappa):
        W, rm = self.rm.W, self.rm
        nv = rm.dim + 1
        self.kappa.data[:] = kappa.reshape(self.kappa.data.shape)
        self.A.zero()
        self.b.zero()
        W.par_loop(rm.k_mass, rm.elems(nv, nv), self.A((rm.en[W.i[0]], rm.en[W.i[1]]), W.INC),
                   rm.coords(rm.ex, W.READ))
        W.par_loop(rm.k_lap, rm.elems(nv, nv), self.A((rm.en[W.i[0]], rm.en[W.i[1]]), W.INC),
                   rm.coords(rm.ex, W.READ), self.kappa(rm.en, W.READ))
        W.par_loop(rm.k_rhs, rm.elems(nv), self.b(rm.en[W.i[0]], W.INC),
                   rm.coords(rm.ex, W.READ), rm.f(rm.en, W.READ))

    def system(self, rtol, maxit):
        from . import hip
        sp = self.rm.sparsity
        return cg.make_system(hip._ptr(sp._dev_rowptr), hip._ptr(sp._dev_colidx),
                              hip._ptr(self.A._dev_vals), hip._ptr(self.b._to_device()),
                              hip._ptr(self.x._to_device()),
                              [hip._ptr(self.work[j]) for j in range(5)], 0, self.rm.n,
                              rtol, 1e-50, 1e4, maxit)

    def solution(self):
        self.x._device_written()
        return np.array(self.x.data_ro).reshape(-1)


class Request(object):

    def __init__(self, rid, arrival, mesh_id, seed):
        self.rid, self.arrival, self.mesh_id, self.seed = rid, arrival, mesh_id, seed
        self.admitted = self.done = None
        self.its = self.reason = None
        self.solution = None


def kappa_of(req, n):
    """The request's seeded lognormal(0, 1) coefficient field."""
    return np.random.default_rng(req.seed).lognormal(0.0, 1.0, n)


def poisson_requests(nreq, rate, pool_size, seed=0):
    """``nreq`` requests with exponential inter-arrival times at ``rate``
    per second, uniformly random meshes and coefficient seeds."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, size=nreq))
    t -= t[0]
    ids = rng.integers(0, pool_size, size=nreq)
    seeds = rng.integers(0, 1 << 31, size=nreq)
    return [Request(k, float(t[k]), int(ids[k]), int(seeds[k])) for k in range(nreq)]


class Server(object):
    """Continuous-batching engine over a pool of :class:`ResidentMesh`."""

    def __init__(self, pool, budget_bytes, max_active=64, iters=16, rtol=1e-10, maxit=20000,
                 keep_solutions=False):
        self.pool = pool
        self.budget = budget_bytes
        self.used = 0
        self.max_active = max_active
        self.iters = iters
        self.rtol, self.maxit = rtol, maxit
        self.keep = keep_solutions
        self.cg = cg.CGBatch(max_active, max_active * max(m.n for m in pool))
        self.rejected = 0

    def _take_slot(self, rm):
        if rm.free:
            return rm.free.pop()
        # release idle slots of other meshes until the new one fits
        for other in self.pool:
            while other.free and self.used + rm.slot_bytes > self.budget:
                other.free.pop()
                self.used -= other.slot_bytes
        if self.used + rm.slot_bytes > self.budget:
            return None
        self.used += rm.slot_bytes
        return rm.new_slot()

    def run(self, requests):
        """Serve ``requests`` (arrival times relative to the start); returns
        them with admitted/done times, iterations and reasons filled in.

        Each engine step enqueues the batch's iterations first and then, while
        the device runs them, admits the requests that have arrived and
        enqueues their assembly (behind the iterations on the same stream);
        they join the batch at the next step."""
        from . import hip
        s = hip.stream()
        queue = sorted(requests, key=lambda r: r.arrival)
        qi = 0
        active = []          # (request, slot, system)
        pending = []         # admitted and assembled, joining at the next step
        t0 = time.perf_counter()
        while qi < len(queue) or active or pending:
            active += pending
            pending = []
            if active:
                self.cg.load([a[2] for a in active], s)
                self.cg.step(self.iters, s)
            now = time.perf_counter() - t0
            while qi < len(queue) and queue[qi].arrival <= now and \
                    len(active) + len(pending) < self.max_active:
                req = queue[qi]
                rm = self.pool[req.mesh_id]
                slot = self._take_slot(rm)
                if slot is None:
                    break        # HBM budget full: wait for a slot to free up
                req.admitted = now
                slot.assemble(kappa_of(req, rm.n))
                pending.append((req, slot, slot.system(self.rtol, self.maxit)))
                qi += 1
            if not active:
                if not pending:
                    if qi < len(queue) and queue[qi].arrival > now:
                        time.sleep(min(queue[qi].arrival - now, 0.002))
                        continue
                    if qi < len(queue):
                        raise RuntimeError("request %d does not fit the HBM budget"
                                           % queue[qi].rid)
                continue
            out = self.cg.read(s)
            keep = []
            for (req, slot, _), st in zip(active, out):
                if st.state == cg.SYS_DONE:
                    req.done = time.perf_counter() - t0
                    req.its, req.reason = st.its, st.reason
                    if self.keep:
                        req.solution = slot.solution()
                    slot.rm.free.append(slot)
                else:
                    keep.append((req, slot, st))
            active = keep
        return requests


def serve_sequential(pool, requests, rtol=1e-10, maxit=20000, keep_solutions=False):
    """The plain path: one request at a time, in arrival order, each
    assembled with par_loops and solved with ``Solver.solve``."""
    from . import hip
    slots = {}
    t0 = time.perf_counter()
    for req in sorted(requests, key=lambda r: r.arrival):
        now = time.perf_counter() - t0
        if req.arrival > now:
            time.sleep(req.arrival - now)
        rm = pool[req.mesh_id]
        slot = slots.get(req.mesh_id)
        if slot is None:
            slot = slots[req.mesh_id] = rm.new_slot()
        req.admitted = time.perf_counter() - t0
        slot.assemble(kappa_of(req, rm.n))
        solver = hip.Solver(relative_tolerance=rtol, maximum_iterations=maxit,
                            error_on_nonconvergence=False)
        solver.solve(slot.A, slot.x, slot.b)
        req.its, req.reason = solver.getIterationNumber(), solver.getConvergedReason()
        if keep_solutions:
            req.solution = slot.solution()
        req.done = time.perf_counter() - t0
    return requests


def latency_stats(requests):
    lat = np.array([r.done - r.arrival for r in requests])
    span = max(r.done for r in requests) - min(r.arrival for r in requests)
    return {'p50_ms': float(np.percentile(lat, 50) * 1e3),
            'p99_ms': float(np.percentile(lat, 99) * 1e3),
            'solves_per_s': len(requests) / span if span > 0 else float('inf'),
            'requests': len(requests)}
