"""The airfoil step in the reference's vector-map form
(demo/airfoil_vector.py:84-135; kernels through pointer arrays, res
incremented through the edge's vector map) against the scalar form: the
kernels do the same arithmetic, so every Dat and the rms Global are equal
bit for bit, on the CPU oracle and on hip; and hip's vector form against
the oracle's, as the scalar form's parity tests check it."""

import numpy as np
import pytest

from oracle import sequential as seq
from pyop2_amd.workloads import airfoil, meshes


def _run(A, steps=2):
    rms = []
    for _ in range(steps):
        A.rms.data = 0.0
        A.iteration()
        rms.append(float(A.rms.data[0]))
    return rms


def _state(A):
    return [np.array(d.data_ro) for d in (A.p_q, A.p_qold, A.p_adt, A.p_res)]


def test_vector_equals_scalar_oracle():
    """On the oracle: the vector-map kernels reproduce the scalar ones."""
    mesh = meshes.QuadMesh(24, 12, seed=3)
    q0 = airfoil.initial_q(mesh.ncell, seed=2)
    S = airfoil.Airfoil(seq, mesh, q0=q0)
    try:
        V = airfoil.AirfoilVector(seq, mesh, q0=q0, consts=False)
        rs, rv = _run(S), _run(V)
        assert rs == rv
        for a, b in zip(_state(S), _state(V)):
            assert np.array_equal(a, b)
    finally:
        S.remove_consts()


@pytest.mark.gpu
def test_vector_hip_equals_scalar_hip_and_oracle(hip):
    """On hip: the vector form gives the scalar form's results bit for bit
    (owner-computes INC through a vector map sums in the same element
    order), and both match the oracle: q, adt and res bit-exact, rms within
    1e-12."""
    mesh = meshes.QuadMesh(96, 48, seed=11)
    q0 = airfoil.initial_q(mesh.ncell, seed=5)
    S = airfoil.Airfoil(hip, mesh, q0=q0)
    try:
        V = airfoil.AirfoilVector(hip, mesh, q0=q0, consts=False)
        O = airfoil.AirfoilVector(seq, mesh, q0=q0, consts=False)
        rs, rv, ro = _run(S), _run(V), _run(O)
        assert rs == rv
        for a, b in zip(_state(S), _state(V)):
            assert np.array_equal(a, b)
        for a, b in zip(_state(V), _state(O)):
            assert np.array_equal(a, b)
        for a, b in zip(rv, ro):
            assert abs(a - b) <= 1e-12 * abs(b)
    finally:
        S.remove_consts()
