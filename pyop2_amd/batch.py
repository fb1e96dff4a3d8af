"""Batched assembly and solve of many independent meshes.

``MeshBatch`` concatenates the meshes into one op2 problem: their vertex
and element sets are laid end to end and every map entry is offset by its
mesh's first vertex, so one ``par_loop`` (one launch of the owner-computes
template) assembles every mesh, and one ``Sparsity`` holds the
block-diagonal pattern.  The solve is one batched CG (csrc/cg_batch.hip)
with a system per mesh: a row range of the shared CSR with its own
scalars, iteration count and convergence test.

Each mesh's values do not depend on its batch: a target's increments are
summed in ascending element order within the mesh, and a system's
reductions are per 256-row chunk of that system, so a mesh solved alone
and in a batch gives bit-identical results.

The per-mesh problem is the reaction-diffusion form used by the batching
and serving workloads: A = M + K(kappa) (P1 mass plus Laplace with a
per-vertex coefficient) and b = M f, solved to a relative preconditioned
residual ``rtol``.
"""

import numpy as np

from . import cg
from . import layout
from .workloads import fem


class MeshBatch(object):
    """Meshes (objects with ``elem_node``, ``coords``, ``nvert``, ``nelem``)
    assembled and solved together in world ``W`` (the op2 module)."""

    def __init__(self, W, meshes, f=None, kappa=None, name='batch'):
        self.W = W
        self.meshes = list(meshes)
        nv = np.array([m.nvert for m in self.meshes], dtype=np.int64)
        ne = np.array([m.nelem for m in self.meshes], dtype=np.int64)
        self.voff = np.concatenate([[0], np.cumsum(nv)])
        self.eoff = np.concatenate([[0], np.cumsum(ne)])
        tv, te = int(self.voff[-1]), int(self.eoff[-1])
        en = np.concatenate([m.elem_node.astype(np.int64) + self.voff[k]
                             for k, m in enumerate(self.meshes)]).astype(np.int32)
        xy = np.concatenate([m.coords for m in self.meshes])
        if f is None:
            f = [np.ones(m.nvert) for m in self.meshes]
        if kappa is None:
            kappa = [np.ones(m.nvert) for m in self.meshes]
        self.nodes = W.Set(tv, 1, name + '_nodes')
        self.vnodes = W.Set(tv, 2, name + '_vnodes')
        self.elems = W.Set(te, 1, name + '_elems')
        self.elem_node = W.Map(self.elems, self.nodes, 3, en, name + '_elem_node')
        self.elem_vnode = W.Map(self.elems, self.vnodes, 3, en, name + '_elem_vnode')
        self.coords = W.Dat(self.vnodes, xy, np.float64, name + '_coords')
        self.f = W.Dat(self.nodes, np.concatenate(f), np.float64, name + '_f')
        self.kappa = W.Dat(self.nodes, np.concatenate(kappa), np.float64, name + '_kappa')
        self.b = W.Dat(self.nodes, np.zeros(tv), np.float64, name + '_b')
        self.x = W.Dat(self.nodes, np.zeros(tv), np.float64, name + '_x')
        self.sparsity = W.Sparsity((self.elem_node, self.elem_node), name + '_sp')
        self.A = W.Mat(self.sparsity, np.float64, name + '_A')
        self.k_mass = W.Kernel(fem.MASS_TRI, 'mass_tri')
        self.k_lap = W.Kernel(fem.LAPLACE_TRI, 'laplace_tri')
        self.k_rhs = W.Kernel(fem.RHS_TRI, 'rhs_tri')
        self._cg = None

    def assemble(self):
        """A = M + K(kappa), b = M f for every mesh: three par_loops."""
        W, en, ev = self.W, self.elem_node, self.elem_vnode
        self.A.zero()
        self.b.zero()
        W.par_loop(self.k_mass, self.elems(3, 3), self.A((en[W.i[0]], en[W.i[1]]), W.INC),
                   self.coords(ev, W.READ))
        W.par_loop(self.k_lap, self.elems(3, 3), self.A((en[W.i[0]], en[W.i[1]]), W.INC),
                   self.coords(ev, W.READ), self.kappa(en, W.READ))
        W.par_loop(self.k_rhs, self.elems(3), self.b(en[W.i[0]], W.INC),
                   self.coords(ev, W.READ), self.f(en, W.READ))

    def systems(self, rtol=1e-10, atol=1e-50, dtol=1e4, maxit=10000):
        """One CG system per mesh: the device row range of its vertices."""
        from . import hip
        lay = layout.of(self.nodes)
        sp = self.sparsity
        nt = int(self.voff[-1])
        if self._cg is None:
            self._work = hip.torch.empty((5, max(nt, 1)), dtype=hip.torch.float64,
                                         device=hip.device())
        bdev = self.b._to_device()
        xdev = self.x._to_device()
        out = []
        for k in range(len(self.meshes)):
            lo, hi = int(self.voff[k]), int(self.voff[k + 1])
            if lay.identity:
                r0 = lo
            else:
                pos = lay.iperm[lo:hi]
                r0 = int(pos.min())
                if int(pos.max()) - r0 + 1 != hi - lo:
                    raise RuntimeError("mesh %d is not contiguous in the device layout" % k)
            out.append(cg.make_system(hip._ptr(sp._dev_rowptr), hip._ptr(sp._dev_colidx),
                                      hip._ptr(self.A._dev_vals), hip._ptr(bdev),
                                      hip._ptr(xdev),
                                      [hip._ptr(self._work[j]) for j in range(5)], r0, hi - lo,
                                      rtol, atol, dtol, maxit))
        return out

    def solve(self, rtol=1e-10, maxit=10000, chunk=32, graph=True):
        """Batched CG of every mesh to ``rtol``; returns per-mesh
        (iterations, residual norm, PETSc reason)."""
        from . import hip
        systems = self.systems(rtol=rtol, maxit=maxit)
        if self._cg is None or self._cg.max_sys < len(systems):
            self._cg = cg.CGBatch(len(systems), int(self.voff[-1]))
        out = self._cg.solve(systems, hip.stream(), chunk=chunk, graph=graph)
        self.x._device_written()
        return [(s.its, s.dp, s.reason) for s in out]

    def solution(self, k):
        """Host copy of mesh k's solution."""
        return np.array(self.x.data_ro[self.voff[k]:self.voff[k + 1]]).reshape(-1)
