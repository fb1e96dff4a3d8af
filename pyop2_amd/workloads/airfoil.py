"""Airfoil-style finite-volume workload on a synthetic quad mesh.

The five user kernels below are the workload of the benchmark: the
non-linear 2-D inviscid airfoil code of OP2 (Copyright (c) 2011, Mike Giles
and others, BSD licence), whose kernels the reference ships as C strings
(demo/airfoil_kernels.py:64-187).  They are input data here: the element
math of the measured loops has to be the reference's, so the strings keep
its arithmetic exactly.  The driver follows demo/airfoil.py:84-135:
per iteration save_soln, then twice adt_calc, res_calc, bres_calc, update.

``build(W, mesh)`` declares the sets, maps, dats and consts with the op2-like
module ``W`` (``pyop2_amd.op2`` after ``op2.init``, or the CPU oracle's
backend module) and returns an :class:`Airfoil` whose ``iteration()`` runs
one time step.
"""

import re
from math import atan, sqrt

import numpy as np

SAVE_SOLN = """
void save_soln(double *q, double *qold){
  for (int n=0; n<4; n++) qold[n] = q[n];
}
"""

ADT_CALC = """
void adt_calc(double *x1,double *x2,double *x3,double *x4,double *q,double *adt){
  double dx,dy, ri,u,v,c;

  ri =  1.0f/q[0];
  u  =   ri*q[1];
  v  =   ri*q[2];
  c  = sqrt(gam*gm1*(ri*q[3]-0.5f*(u*u+v*v)));

  dx = x2[0] - x1[0];
  dy = x2[1] - x1[1];
  *adt  = fabs(u*dy-v*dx) + c*sqrt(dx*dx+dy*dy);

  dx = x3[0] - x2[0];
  dy = x3[1] - x2[1];
  *adt += fabs(u*dy-v*dx) + c*sqrt(dx*dx+dy*dy);

  dx = x4[0] - x3[0];
  dy = x4[1] - x3[1];
  *adt += fabs(u*dy-v*dx) + c*sqrt(dx*dx+dy*dy);

  dx = x1[0] - x4[0];
  dy = x1[1] - x4[1];
  *adt += fabs(u*dy-v*dx) + c*sqrt(dx*dx+dy*dy);

  *adt = (*adt) / cfl;
}
"""

RES_CALC = """
void res_calc(double *x1,  double *x2,  double *q1,  double *q2,
              double *adt1,double *adt2,double *res1,double *res2) {
  double dx,dy,mu, ri, p1,vol1, p2,vol2, f;

  dx = x1[0] - x2[0];
  dy = x1[1] - x2[1];

  ri   = 1.0f/q1[0];
  p1   = gm1*(q1[3]-0.5f*ri*(q1[1]*q1[1]+q1[2]*q1[2]));
  vol1 =  ri*(q1[1]*dy - q1[2]*dx);

  ri   = 1.0f/q2[0];
  p2   = gm1*(q2[3]-0.5f*ri*(q2[1]*q2[1]+q2[2]*q2[2]));
  vol2 =  ri*(q2[1]*dy - q2[2]*dx);

  mu = 0.5f*((*adt1)+(*adt2))*eps;

  f = 0.5f*(vol1* q1[0]         + vol2* q2[0]        ) + mu*(q1[0]-q2[0]);
  res1[0] += f;
  res2[0] -= f;
  f = 0.5f*(vol1* q1[1] + p1*dy + vol2* q2[1] + p2*dy) + mu*(q1[1]-q2[1]);
  res1[1] += f;
  res2[1] -= f;
  f = 0.5f*(vol1* q1[2] - p1*dx + vol2* q2[2] - p2*dx) + mu*(q1[2]-q2[2]);
  res1[2] += f;
  res2[2] -= f;
  f = 0.5f*(vol1*(q1[3]+p1)     + vol2*(q2[3]+p2)    ) + mu*(q1[3]-q2[3]);
  res1[3] += f;
  res2[3] -= f;
}
"""

BRES_CALC = """
void bres_calc(double *x1,  double *x2,  double *q1,
               double *adt1,double *res1,int *bound) {
  double dx,dy,mu, ri, p1,vol1, p2,vol2, f;

  dx = x1[0] - x2[0];
  dy = x1[1] - x2[1];

  ri = 1.0f/q1[0];
  p1 = gm1*(q1[3]-0.5f*ri*(q1[1]*q1[1]+q1[2]*q1[2]));

  if (*bound==1) {
    res1[1] += + p1*dy;
    res1[2] += - p1*dx;
  }
  else {
    vol1 =  ri*(q1[1]*dy - q1[2]*dx);

    ri   = 1.0f/qinf[0];
    p2   = gm1*(qinf[3]-0.5f*ri*(qinf[1]*qinf[1]+qinf[2]*qinf[2]));
    vol2 =  ri*(qinf[1]*dy - qinf[2]*dx);

    mu = (*adt1)*eps;

    f = 0.5f*(vol1* q1[0]         + vol2* qinf[0]        ) + mu*(q1[0]-qinf[0]);
    res1[0] += f;
    f = 0.5f*(vol1* q1[1] + p1*dy + vol2* qinf[1] + p2*dy) + mu*(q1[1]-qinf[1]);
    res1[1] += f;
    f = 0.5f*(vol1* q1[2] - p1*dx + vol2* qinf[2] - p2*dx) + mu*(q1[2]-qinf[2]);
    res1[2] += f;
    f = 0.5f*(vol1*(q1[3]+p1)     + vol2*(qinf[3]+p2)    ) + mu*(q1[3]-qinf[3]);
    res1[3] += f;
  }
}
"""

UPDATE = """
void update(double *qold, double *q, double *res, double *adt, double *rms){
  double del, adti;

  adti = 1.0f/(*adt);

  for (int n=0; n<4; n++) {
    del    = adti*res[n];
    q[n]   = qold[n] - del;
    res[n] = 0.0f;
    *rms  += del*del;
  }
}
"""


def free_stream(gam=1.4, mach=0.4):
    """The free-stream state of the OP2 airfoil setup: density 1, pressure 1,
    speed mach * sound speed along x."""
    gm1 = gam - 1.0
    r, p = 1.0, 1.0
    u = sqrt(gam * p / r) * mach
    e = p / (r * gm1) + 0.5 * u * u
    return np.array([r, r * u, 0.0, r * e])


CONSTS = {'gam': 1.4, 'gm1': 0.4, 'cfl': 0.9, 'eps': 0.05, 'mach': 0.4,
          'alpha': 3.0 * atan(1.0) / 45.0}


def initial_q(ncell, seed=1, noise=0.01):
    """Free stream with seeded relative noise on every component."""
    rng = np.random.default_rng(seed)
    q = np.tile(free_stream(), (ncell, 1))
    return q * (1.0 + noise * rng.standard_normal(q.shape))


# Algorithmic bytes per loop: every distinct (array, access) once, direct
# READ/WRITE n*cdim*8, RW/INC twice; indirect READ over the unique targets,
# INC twice; each distinct map n*mdim*4.  For a mesh where every cell, node
# and boundary item is touched, unique targets are the whole target set.
def unique_bytes(mesh):
    nc, nn, ne, nb = mesh.ncell, mesh.nnode, mesh.nedge, mesh.nbedge
    d = 8
    return {
        'save_soln': nc * 4 * d * 2,
        'adt_calc': nc * 4 * 4 + nn * 2 * d + nc * 4 * d + nc * d,
        'res_calc': ne * 2 * 4 * 2 + nn * 2 * d + nc * 4 * d + nc * d + 2 * nc * 4 * d,
        'bres_calc': (nb * 2 * 4 + nb * 4 + nb * 4 +
                      len(np.unique(mesh.pbedge)) * 2 * d +
                      len(np.unique(mesh.pbecell)) * (4 * d + d + 2 * 4 * d)),
        'update': nc * 4 * d + nc * 4 * d + 2 * nc * 4 * d + nc * d,
    }


def _vector_form(code, signature, names):
    """A kernel string with its pointer arguments ``names`` (old -> new
    element expression prefix) rewritten for vector maps: ``x1[k]`` becomes
    ``x[0][k]`` and so on; the arithmetic is untouched."""
    head, body = code.split('{', 1)
    for old, new in names:
        body = re.sub(r'\(\*%s\)' % old, '(%s[0])' % new, body)
        body = re.sub(r'\b%s\[' % old, '%s[' % new, body)
    return signature + ' {' + body


# The vector-map form of adt_calc, res_calc and bres_calc (the reference's
# demo/airfoil_vector.py drives them, demo/airfoil_vector_kernels.py holds
# them): the same arithmetic as the kernels above, reached through pointer
# arrays of vector maps.
ADT_CALC_VEC = _vector_form(
    ADT_CALC, '\nvoid adt_calc(double *x[2], double *q, double *adt)',
    [('x1', 'x[0]'), ('x2', 'x[1]'), ('x3', 'x[2]'), ('x4', 'x[3]')])
RES_CALC_VEC = _vector_form(
    RES_CALC, '\nvoid res_calc(double *x[2], double *q[4], double *adt[1], double *res[4])',
    [('x1', 'x[0]'), ('x2', 'x[1]'), ('q1', 'q[0]'), ('q2', 'q[1]'), ('adt1', 'adt[0]'),
     ('adt2', 'adt[1]'), ('res1', 'res[0]'), ('res2', 'res[1]')])
BRES_CALC_VEC = _vector_form(
    BRES_CALC, '\nvoid bres_calc(double *x[2], double *q1, double *adt1, double *res1, int *bound)',
    [('x1', 'x[0]'), ('x2', 'x[1]')])


class Airfoil(object):
    """Sets, maps, dats and kernels of one airfoil problem in world ``W``."""

    def __init__(self, W, mesh, q0=None, consts=True, soa=False):
        """``soa`` stores q, q_old and res (four components per cell) as
        structure-of-arrays Dats."""
        self.W = W
        self.mesh = mesh
        nodes = W.Set(mesh.nnode, 2, 'nodes')
        edges = W.Set(mesh.nedge, 1, 'edges')
        bedges = W.Set(mesh.nbedge, 1, 'bedges')
        cells = W.Set(mesh.ncell, 1, 'cells')
        vcells = W.Set(mesh.ncell, 4, 'vcells')
        self.nodes, self.edges, self.bedges, self.cells, self.vcells = \
            nodes, edges, bedges, cells, vcells
        self.pedge = W.Map(edges, nodes, 2, mesh.pedge, 'pedge')
        self.pecell = W.Map(edges, cells, 2, mesh.pecell, 'pecell')
        self.pevcell = W.Map(edges, vcells, 2, mesh.pecell, 'pevcell')
        self.pbedge = W.Map(bedges, nodes, 2, mesh.pbedge, 'pbedge')
        self.pbecell = W.Map(bedges, cells, 1, mesh.pbecell, 'pbecell')
        self.pbevcell = W.Map(bedges, vcells, 1, mesh.pbecell, 'pbevcell')
        self.pcell = W.Map(cells, nodes, 4, mesh.pcell, 'pcell')
        q0 = initial_q(mesh.ncell) if q0 is None else q0
        self.p_bound = W.Dat(bedges, mesh.bound.copy(), np.int32, 'p_bound')
        self.p_x = W.Dat(nodes, mesh.x.copy(), np.float64, 'p_x')
        self.p_q = W.Dat(vcells, q0.copy(), np.float64, 'p_q', soa=soa)
        self.p_qold = W.Dat(vcells, np.zeros((mesh.ncell, 4)), np.float64, 'p_qold', soa=soa)
        self.p_adt = W.Dat(cells, np.zeros(mesh.ncell), np.float64, 'p_adt')
        self.p_res = W.Dat(vcells, np.zeros((mesh.ncell, 4)), np.float64, 'p_res', soa=soa)
        self.rms = W.Global(1, 0.0, np.float64, 'rms')
        self.consts = []
        if consts:
            for k, v in CONSTS.items():
                self.consts.append(W.Const(1, v, k, np.float64))
            self.consts.append(W.Const(4, free_stream(), 'qinf', np.float64))
        self._kernels()

    CONST_NAMES = ('gam', 'gm1', 'cfl', 'eps', 'mach', 'alpha', 'qinf')

    @classmethod
    def fromhdf5(cls, W, f, consts=True):
        """The problem read from HDF5 data ``f`` (an h5py-like mapping such
        as ``pyop2_amd.hdf5.File``) the way the reference's demo reads
        ``new_grid.h5`` (demo/airfoil.py:45-76): Sets, Maps, Dats and Consts
        through their ``fromhdf5`` constructors."""
        self = cls.__new__(cls)
        self.W = W
        self.mesh = None
        self.nodes = W.Set.fromhdf5(f, 'nodes', dim=2)
        self.edges = W.Set.fromhdf5(f, 'edges')
        self.bedges = W.Set.fromhdf5(f, 'bedges')
        self.cells = W.Set.fromhdf5(f, 'cells')
        self.vcells = W.Set.fromhdf5(f, 'cells', dim=4)
        self.pedge = W.Map.fromhdf5(self.edges, self.nodes, f, 'pedge')
        self.pecell = W.Map.fromhdf5(self.edges, self.cells, f, 'pecell')
        self.pevcell = W.Map.fromhdf5(self.edges, self.vcells, f, 'pecell')
        self.pbedge = W.Map.fromhdf5(self.bedges, self.nodes, f, 'pbedge')
        self.pbecell = W.Map.fromhdf5(self.bedges, self.cells, f, 'pbecell')
        self.pbevcell = W.Map.fromhdf5(self.bedges, self.vcells, f, 'pbecell')
        self.pcell = W.Map.fromhdf5(self.cells, self.nodes, f, 'pcell')
        self.p_bound = W.Dat.fromhdf5(self.bedges, f, 'p_bound')
        self.p_x = W.Dat.fromhdf5(self.nodes, f, 'p_x')
        self.p_q = W.Dat.fromhdf5(self.vcells, f, 'p_q')
        self.p_qold = W.Dat.fromhdf5(self.vcells, f, 'p_qold')
        self.p_adt = W.Dat.fromhdf5(self.cells, f, 'p_adt')
        self.p_res = W.Dat.fromhdf5(self.vcells, f, 'p_res')
        self.rms = W.Global(1, 0.0, np.float64, 'rms')
        self.consts = [W.Const.fromhdf5(f, k) for k in cls.CONST_NAMES] if consts else []
        self._kernels()
        return self

    def _kernels(self):
        W = self.W
        self.k_save_soln = W.Kernel(SAVE_SOLN, 'save_soln')
        self.k_adt_calc = W.Kernel(ADT_CALC, 'adt_calc')
        self.k_res_calc = W.Kernel(RES_CALC, 'res_calc')
        self.k_bres_calc = W.Kernel(BRES_CALC, 'bres_calc')
        self.k_update = W.Kernel(UPDATE, 'update')

    def remove_consts(self):
        for c in self.consts:
            c.remove_from_namespace()

    def save_soln(self):
        W = self.W
        W.par_loop(self.k_save_soln, self.cells,
                   self.p_q(W.IdentityMap, W.READ), self.p_qold(W.IdentityMap, W.WRITE))

    def adt_calc(self):
        W, x, pc = self.W, self.p_x, self.pcell
        W.par_loop(self.k_adt_calc, self.cells,
                   x(pc[0], W.READ), x(pc[1], W.READ), x(pc[2], W.READ), x(pc[3], W.READ),
                   self.p_q(W.IdentityMap, W.READ), self.p_adt(W.IdentityMap, W.WRITE))

    def res_calc(self):
        W = self.W
        W.par_loop(self.k_res_calc, self.edges,
                   self.p_x(self.pedge[0], W.READ), self.p_x(self.pedge[1], W.READ),
                   self.p_q(self.pevcell[0], W.READ), self.p_q(self.pevcell[1], W.READ),
                   self.p_adt(self.pecell[0], W.READ), self.p_adt(self.pecell[1], W.READ),
                   self.p_res(self.pevcell[0], W.INC), self.p_res(self.pevcell[1], W.INC))

    def bres_calc(self):
        W = self.W
        W.par_loop(self.k_bres_calc, self.bedges,
                   self.p_x(self.pbedge[0], W.READ), self.p_x(self.pbedge[1], W.READ),
                   self.p_q(self.pbevcell[0], W.READ), self.p_adt(self.pbecell[0], W.READ),
                   self.p_res(self.pbevcell[0], W.INC), self.p_bound(W.IdentityMap, W.READ))

    def update(self, rms=None):
        W = self.W
        W.par_loop(self.k_update, self.cells,
                   self.p_qold(W.IdentityMap, W.READ), self.p_q(W.IdentityMap, W.WRITE),
                   self.p_res(W.IdentityMap, W.RW), self.p_adt(W.IdentityMap, W.READ),
                   (self.rms if rms is None else rms)(W.INC))

    def iteration(self, inner=2):
        """One time step: save_soln, then ``inner`` x (adt, res, bres, update)."""
        self.save_soln()
        for _ in range(inner):
            self.adt_calc()
            self.res_calc()
            self.bres_calc()
            self.update()

    def demo_iteration(self, inner=2):
        """One time step as the reference's demo drives it
        (demo/airfoil.py:89-133): a new ``rms`` Global for each update, and
        the last one read back on the host at the end of the step.  Returns
        sqrt(rms / cells)."""
        W = self.W
        self.save_soln()
        for _ in range(inner):
            self.adt_calc()
            self.res_calc()
            self.bres_calc()
            rms = W.Global(1, 0.0, np.float64, 'rms')
            self.update(rms)
        return float(np.sqrt(rms.data[0] / self.cells.size))


class AirfoilVector(Airfoil):
    """The same problem driven as the reference's vector-map demo
    (demo/airfoil_vector.py:84-135): adt_calc reads the four nodes of a
    cell through one vector map, res_calc reads and increments through the
    edge's vector maps, bres_calc reads its nodes through one.  Results
    equal :class:`Airfoil`'s bit for bit."""

    def _kernels(self):
        super(AirfoilVector, self)._kernels()
        W = self.W
        self.k_adt_calc = W.Kernel(ADT_CALC_VEC, 'adt_calc')
        self.k_res_calc = W.Kernel(RES_CALC_VEC, 'res_calc')
        self.k_bres_calc = W.Kernel(BRES_CALC_VEC, 'bres_calc')

    def adt_calc(self):
        W = self.W
        W.par_loop(self.k_adt_calc, self.cells, self.p_x(self.pcell, W.READ),
                   self.p_q(W.IdentityMap, W.READ), self.p_adt(W.IdentityMap, W.WRITE))

    def res_calc(self):
        W = self.W
        W.par_loop(self.k_res_calc, self.edges, self.p_x(self.pedge, W.READ),
                   self.p_q(self.pevcell, W.READ), self.p_adt(self.pecell, W.READ),
                   self.p_res(self.pevcell, W.INC))

    def bres_calc(self):
        W = self.W
        W.par_loop(self.k_bres_calc, self.bedges, self.p_x(self.pbedge, W.READ),
                   self.p_q(self.pbevcell[0], W.READ), self.p_adt(self.pbecell[0], W.READ),
                   self.p_res(self.pbevcell[0], W.INC), self.p_bound(W.IdentityMap, W.READ))
