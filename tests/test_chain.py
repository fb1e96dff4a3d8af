"""Owner loops chained inside a fusing capture (gpu): res_calc and
bres_calc, which both increment res and touch it no other way, run as one
kernel in the captured airfoil step, and the replayed steps stay bit-exact
to the oracle (each cell's increments keep the sequential order)."""

import numpy as np
import pytest

from oracle import sequential as seq
from pyop2_amd.workloads import airfoil, meshes

pytestmark = pytest.mark.gpu


def test_fused_airfoil_step_bit_exact(hip):
    from pyop2_amd import hip as H
    mesh = meshes.QuadMesh(64, 32, seed=9)
    A = airfoil.Airfoil(hip, mesh)
    try:
        B = airfoil.Airfoil(seq, mesh, consts=False)
        A.iteration()
        B.iteration()
        g = H.capture(A.iteration, fuse=True)
        chains = [e for e in H._chain_cache.values() if e is not None]
        assert any(e.loop.source.entry == 'op2_res_calc_bres_calc_chain' for e in chains)
        for _ in range(3):
            g.replay()
            B.iteration()
        assert np.array_equal(A.p_q.data_ro, B.p_q.data_ro)
        assert np.array_equal(A.p_res.data_ro, B.p_res.data_ro)
        assert abs(A.rms.data[0] - B.rms.data[0]) <= 1e-12 * abs(B.rms.data[0])
    finally:
        A.remove_consts()


def test_unfused_capture_matches_fused(hip):
    from pyop2_amd import hip as H
    mesh = meshes.QuadMesh(40, 20, seed=4)
    outs = []
    for fuse in (True, False):
        A = airfoil.Airfoil(hip, mesh)
        try:
            A.iteration()
            g = H.capture(A.iteration, fuse=fuse)
            g.replay()
            g.replay()
            outs.append(A.p_q.data_ro.copy())
        finally:
            A.remove_consts()
    assert np.array_equal(outs[0], outs[1])
