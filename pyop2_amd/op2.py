"""The op2 API (reference: pyop2/op2.py:49-164).

``init`` selects a backend once; user classes dispatch by name to the
selected backend through :class:`backends._BackendSelector`; ``par_loop`` and
``solve`` forward to the backend.  ``op2.init(backend='hip')`` selects the
MI355X backend of this package."""

import atexit

from . import backends
from . import base
from . import configuration as cfg
from .base import READ, WRITE, RW, INC, MIN, MAX, IdentityMap, i
from .exceptions import MatTypeError, DatTypeError
from .logger import debug, info, warning, error, critical, set_log_level
from .mpi import MPI
from .utils import validate_type


def init(**kwargs):
    """Select the backend and configure op2.

    Calling ``init`` again with another backend, or after :func:`exit`,
    raises ``RuntimeError``; calling it again with the same backend (or none)
    updates the configuration."""
    backend = backends.get_backend()
    if backend == 'finalised':
        raise RuntimeError("Calling init() after exit() is illegal.")
    if 'backend' in kwargs and backend not in ('void', kwargs['backend']):
        raise RuntimeError("Changing the backend is not possible once set.")
    cfg.configure(**kwargs)
    set_log_level(cfg['log_level'])
    if backend == 'void':
        backends.set_backend(cfg.backend)
        try:
            backends._BackendSelector._backend._setup()
        except Exception:
            backends._BackendSelector._backend = backends.void
            backends._BackendSelector._name = 'void'
            raise
        if 'comm' in kwargs:
            backends._BackendSelector._backend.MPI.comm = kwargs['comm']
        global MPI
        MPI = backends._BackendSelector._backend.MPI


@atexit.register
def exit():
    """Exit op2."""
    cfg.reset()
    if backends.get_backend() != 'void':
        backends.unset_backend()


class IterationSpace(base.IterationSpace, metaclass=backends._BackendSelector):
    pass


class Kernel(base.Kernel, metaclass=backends._BackendSelector):
    pass


class Set(base.Set, metaclass=backends._BackendSelector):
    pass


class Halo(base.Halo, metaclass=backends._BackendSelector):
    pass


class Dat(base.Dat, metaclass=backends._BackendSelector):
    pass


class Mat(base.Mat, metaclass=backends._BackendSelector):
    pass


class Const(base.Const, metaclass=backends._BackendSelector):
    pass


class Global(base.Global, metaclass=backends._BackendSelector):
    pass


class Map(base.Map, metaclass=backends._BackendSelector):
    pass


class Sparsity(base.Sparsity, metaclass=backends._BackendSelector):
    pass


class Solver(base.Solver, metaclass=backends._BackendSelector):
    pass


def par_loop(kernel, it_space, *args):
    """Run ``kernel`` over every element of ``it_space`` with ``args``.

    ``it_space`` is a Set, or ``set(e0)`` / ``set(e0, e1)`` for a local
    iteration space; args are built by calling Dats, Globals and Mats with a
    map (and index) and an access descriptor."""
    return backends._BackendSelector._backend.par_loop(kernel, it_space, *args)


@validate_type(('M', base.Mat, MatTypeError), ('x', base.Dat, DatTypeError),
               ('b', base.Dat, DatTypeError))
def solve(M, x, b):
    """Solve ``M x = b`` with the default Solver."""
    Solver().solve(M, x, b)
