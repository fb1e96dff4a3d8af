"""Write tests/golden/airfoil_24x12.h5: a small airfoil mesh in the layout
the reference's airfoil demo reads (demo/airfoil.py:45-76; the OP2
new_grid.h5 layout): Set sizes as one-element int32 datasets, maps as
(n, dim) int32, Dats as (n, dim) with a 'type' attribute, Consts as 1-D
float64.  The mesh is pyop2_amd.workloads.meshes.QuadMesh(24, 12, seed=3)
with the seeded initial state of pyop2_amd.workloads.airfoil.initial_q.

    python3 tests/golden/make_airfoil_h5.py
"""

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import h5write  # noqa: E402
from pyop2_amd.workloads import airfoil, meshes  # noqa: E402

PATH = os.path.join(HERE, 'airfoil_24x12.h5')


def datasets():
    m = meshes.QuadMesh(24, 12, seed=3)
    q = airfoil.initial_q(m.ncell)
    i32 = lambda a: np.asarray(a, dtype=np.int32)
    d = {
        'nodes': (i32([m.nnode]), {}),
        'edges': (i32([m.nedge]), {}),
        'bedges': (i32([m.nbedge]), {}),
        'cells': (i32([m.ncell]), {}),
        'pedge': (i32(m.pedge), {}),
        'pecell': (i32(m.pecell), {}),
        'pbedge': (i32(m.pbedge), {}),
        'pbecell': (i32(m.pbecell).reshape(-1, 1), {}),
        'pcell': (i32(m.pcell), {}),
        'p_bound': (i32(m.bound).reshape(-1, 1), {'type': 'int'}),
        'p_x': (np.asarray(m.x, dtype=np.float64), {'type': 'double'}),
        'p_q': (q, {'type': 'double'}),
        'p_qold': (np.zeros_like(q), {'type': 'double'}),
        'p_adt': (np.zeros((m.ncell, 1)), {'type': 'double'}),
        'p_res': (np.zeros_like(q), {'type': 'double'}),
        'qinf': (airfoil.free_stream(), {}),
    }
    for k, v in airfoil.CONSTS.items():
        d[k] = (np.array([v]), {})
    return m, q, d


if __name__ == '__main__':
    h5write.write(PATH, datasets()[2])
    print(PATH, os.path.getsize(PATH))
