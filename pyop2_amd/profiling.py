"""Named wall-clock timers (the reference's pyop2/profiling.py interface:
``Timer``, ``tic``/``toc``, ``profile``, ``summary``, ``get_timers``,
``reset``).

Work on the hip backend is enqueued asynchronously, so a host clock read
right after a par_loop does not include the loop's GPU time.  Timers made
with ``sync=True`` (``tic(name, sync=True)``) wait for the device before
each clock read.
"""

import functools
import time

import numpy as np


def _device_sync():
    try:
        import torch
        if torch.cuda.is_available() and torch.cuda.is_initialized():
            torch.cuda.synchronize()
    except ImportError:
        pass


class Timer(object):
    """A named timer; one instance per name (constructing the name again
    returns it).  ``timer`` is the clock (seconds), ``sync`` waits for the
    device before each reading."""

    _timers = {}

    def __new__(cls, name=None, timer=time.perf_counter, sync=False):
        n = name or 'timer%d' % len(cls._timers)
        if n in cls._timers:
            return cls._timers[n]
        t = super(Timer, cls).__new__(cls)
        t._name = n
        t._timer = timer
        t._sync = sync
        t._start = None
        t._timings = []
        cls._timers[n] = t
        return t

    def __init__(self, name=None, timer=time.perf_counter, sync=False):
        pass

    def _now(self):
        if self._sync:
            _device_sync()
        return self._timer()

    def start(self):
        """Start the timer."""
        self._start = self._now()

    def stop(self):
        """Stop the timer and record the interval."""
        assert self._start is not None, "Timer %s has not been started yet." % self._name
        self._timings.append(self._now() - self._start)
        self._start = None

    @property
    def name(self):
        return self._name

    @property
    def elapsed(self):
        """Time since the running timer started."""
        assert self._start is not None, "Timer %s has not been started yet." % self._name
        return self._now() - self._start

    @property
    def ncalls(self):
        """Number of recorded intervals."""
        return len(self._timings)

    @property
    def total(self):
        return float(sum(self._timings))

    @property
    def average(self):
        return float(np.mean(self._timings)) if self._timings else 0.0

    @classmethod
    def summary(cls, filename=None):
        """A table of every timer on stdout, or CSV into ``filename``."""
        if not cls._timers:
            return
        heads = ("Timer", "Total time", "Calls", "Average time")
        rows = [(t.name, t.total, t.ncalls, t.average) for t in cls._timers.values()]
        if isinstance(filename, str):
            import csv
            with open(filename, 'w', newline='') as f:
                w = csv.writer(f, lineterminator='\n')
                w.writerow(heads)
                w.writerows(rows)
            return
        cells = [heads] + [(n, '%g' % tot, '%d' % nc, '%g' % av) for n, tot, nc, av in rows]
        widths = [max(len(r[k]) for r in cells) for k in range(4)]
        for r in cells:
            print(' | '.join(c.rjust(w) for c, w in zip(r, widths)))

    @classmethod
    def get_timers(cls):
        """The timers by name."""
        return cls._timers

    @classmethod
    def reset(cls):
        """Forget every timer."""
        cls._timers = {}


class profile(Timer):
    """Decorator timing each call of a function (named after it unless a
    name is given)."""

    def __call__(self, f):
        @functools.wraps(f)
        def wrapper(*args, **kwargs):
            self.start()
            try:
                return f(*args, **kwargs)
            finally:
                self.stop()
        return wrapper


def tic(name, sync=False):
    """Start the timer ``name``."""
    Timer(name, sync=sync).start()


def toc(name):
    """Stop the timer ``name``."""
    Timer(name).stop()


def summary(filename=None):
    Timer.summary(filename)


def get_timers():
    return Timer.get_timers()


def reset():
    Timer.reset()
