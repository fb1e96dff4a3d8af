"""Launch-parameter autotuner: the table logic on the CPU; on the GPU,
tuned loops give the oracle's results bit for bit and the trials leave the
data as it was."""

import json

import numpy as np
import pytest

from pyop2_amd import autotune


class _FakeLoop(object):
    """Just what autotune.signature reads."""

    class _K(object):
        name = 'k'
        code = 'void k(double* a) {}'

    class _A(object):
        _is_mat = False
        _is_global = False
        _is_direct = True

        class access(object):
            _mode = 'RW'

        class data(object):
            cdim = 2

    class _S(object):
        size = 1000

    kernel = _K()
    args = [_A()]
    _it_space = _S()


def test_table_picks_fastest_and_persists(tmp_path, monkeypatch):
    path = tmp_path / 'autotune.json'
    monkeypatch.setattr(autotune, '_path', lambda: str(path))
    autotune.reset()
    cost = {128: 3e-6, 256: 1e-6, 512: 2e-6}
    calls = []

    def trial(p):
        calls.append(p)
        return cost[p['hip_block_size']]
    best = autotune.params_for(_FakeLoop(), 'stream', trial)
    assert best == {'hip_block_size': 256}
    assert len(calls) == 3
    # cached: no more trials, and on disk
    assert autotune.params_for(_FakeLoop(), 'stream', trial) == best
    assert len(calls) == 3
    saved = json.load(open(path))
    (key, entry), = saved.items()
    assert key.endswith('/stream/dir:RW:2/2^9')
    assert entry['params'] == best and len(entry['tried']) == 3
    autotune.reset()
    assert autotune.params_for(_FakeLoop(), 'stream', trial) == best   # reloaded from disk
    assert len(calls) == 3


def test_failing_candidates_are_skipped(tmp_path, monkeypatch):
    monkeypatch.setattr(autotune, '_path', lambda: str(tmp_path / 'a.json'))
    autotune.reset()

    def trial(p):
        if p['hip_owner_block_size'] == 512:
            raise RuntimeError('does not fit')
        return p['hip_owner_fill']
    best = autotune.params_for(_FakeLoop(), 'owner', trial)
    assert best == {'hip_owner_block_size': 256, 'hip_owner_fill': 0.6}
    entry = list(autotune.table().values())[0]
    assert sum('error' in t for t in entry['tried']) == 3


@pytest.mark.gpu
def test_tuned_airfoil_matches_oracle(hip, tmp_path, monkeypatch):
    from oracle import sequential as seq
    from pyop2_amd import configuration as cfg
    from pyop2_amd.workloads import airfoil, meshes
    monkeypatch.setattr(autotune, '_path', lambda: str(tmp_path / 'tune.json'))
    autotune.reset()
    mesh = meshes.QuadMesh(48, 24, seed=3)
    with cfg.override(hip_autotune=True):
        A = airfoil.Airfoil(hip, mesh)
        try:
            A.iteration()
            q, rms = A.p_q.data_ro.copy(), np.array(A.rms.data).copy()
            B = airfoil.Airfoil(seq, mesh, consts=False)
            B.iteration()
        finally:
            A.remove_consts()
    assert np.array_equal(q, B.p_q.data_ro)
    assert abs(rms[0] - B.rms.data[0]) <= 1e-12 * abs(B.rms.data[0])
    t = autotune.table()
    names = set(k.split('/')[0] for k in t)
    assert {'save_soln', 'adt_calc', 'res_calc', 'bres_calc', 'update'} <= names
    assert all(e['tried'] for e in t.values())
