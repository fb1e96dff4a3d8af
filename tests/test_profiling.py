"""Timers (the reference's test/unit/test_profiling.py, plus the decorator,
the summary and device-synchronised timers)."""

import pytest

from pyop2_amd.profiling import Timer, get_timers, profile, reset, summary, tic, toc


def test_create():
    tic('create')
    toc('create')
    assert 'create' in get_timers().keys()


def test_elapsed_nonstarted_fails():
    t = Timer('test_elapsed_nonstarted_fails')
    with pytest.raises(AssertionError):
        t.elapsed


def test_stop_nonstarted_fails():
    t = Timer('test_stop_nonstarted_fails')
    with pytest.raises(AssertionError):
        t.stop()


def test_ncalls():
    t = Timer('test_ncalls')
    for i in range(10):
        t.start()
        t.stop()
    assert t.ncalls == 10
    assert Timer('test_ncalls') is t


def test_reset():
    tic('test_reset')
    toc('test_reset')
    reset()
    assert list(get_timers().keys()) == []


def test_profile_decorator_and_summary(tmp_path, capsys):
    @profile('twice')
    def f(x):
        return 2 * x

    assert f(3) == 6 and f(4) == 8
    assert get_timers()['twice'].ncalls == 2
    tic('synced', sync=True)
    toc('synced')
    summary()
    out = capsys.readouterr().out
    assert 'twice' in out and 'Calls' in out
    p = str(tmp_path / 't.csv')
    summary(p)
    lines = open(p).read().splitlines()
    assert lines[0] == 'Timer,Total time,Calls,Average time'
    assert any(l.startswith('twice,') for l in lines)
