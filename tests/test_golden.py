"""The committed golden fixtures (tests/golden/, written by
tests/golden/make_golden.py with the CPU oracle): small seeded airfoil,
mass2d and 3-D assembly cases.

CPU: the oracle reproduces the fixture exactly (any change to the oracle's
arithmetic shows here).  GPU: hip reproduces it to the parity bars: maps and
CSR bit-exact; airfoil q, qold, adt and res bit-exact (owner-computes sums
in element order; one increment per kernel call); the rms Global, Mat
values and the right-hand side (three increments per call, summed
privately first) within 1e-12 relative to the array's largest magnitude;
the CG solution within 1e-10 of the oracle's."""

import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, 'golden'))

import make_golden  # noqa: E402

GOLD = os.path.join(HERE, 'golden', 'oracle_small.npz')

EXACT = ('airfoil_q', 'airfoil_qold', 'airfoil_adt', 'airfoil_res',
         'mass2d_rowptr', 'mass2d_colidx', 'laplace_rowptr', 'laplace_colidx',
         'vmass_rowptr', 'vmass_colidx')
CLOSE = ('airfoil_rms', 'mass2d_values', 'mass2d_b', 'laplace_values', 'vmass_values')


def _gold():
    with np.load(GOLD, allow_pickle=False) as d:
        return {k: d[k] for k in d.files}


def _rel(a, b):
    scale = max(np.abs(b).max(), 1e-300)
    return np.abs(np.asarray(a, dtype=np.float64) - b).max() / scale


def test_oracle_reproduces_golden():
    from oracle import sequential
    gold = _gold()
    got = make_golden.cases(sequential)
    assert sorted(got) == sorted(gold)
    for k in gold:
        assert np.array_equal(got[k], gold[k]), k


@pytest.mark.gpu
def test_hip_matches_golden(hip):
    gold = _gold()
    got = make_golden.cases(hip)
    for k in EXACT:
        assert np.array_equal(got[k], gold[k]), k
    for k in CLOSE:
        assert got[k].shape == gold[k].shape, k
        assert _rel(got[k], gold[k]) <= 1e-12, k
    assert _rel(got['mass2d_x'], gold['mass2d_x']) <= 1e-10
