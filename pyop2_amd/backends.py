"""The backend plugin boundary (reference: pyop2/backends.py:44-150).

User-level classes in :mod:`op2` carry the :class:`_BackendSelector`
metaclass: constructing one looks up the class of the same name in the
selected backend module and instantiates that.  ``backends`` maps backend
names to modules; :func:`set_backend` imports ``pyop2_amd.<name>`` unless the
name is already registered, so a backend defined outside the package (the
CPU oracle used by the tests) is plugged in by adding it to this dict."""

import importlib

from . import void
from . import finalised
from .logger import warning

backends = {'void': void, 'finalised': finalised}


_selectors = {}


def _make_object(obj, *args, **kwargs):
    """Instantiate the backend class named ``obj`` (the selector resolves
    the backend at each call, so one selector class per name serves every
    backend)."""
    sel = _selectors.get(obj)
    if sel is None:
        sel = _selectors[obj] = _BackendSelector(obj, (object,), {})
    return sel(*args, **kwargs)


class _BackendSelector(type):
    """Metaclass creating the backend class corresponding to the requested
    class."""

    _backend = void
    _name = 'void'

    def __new__(cls, name, bases, dct):
        if not dct.get('__doc__'):
            for mro_cls in (c for b in bases for c in b.mro()):
                if mro_cls.__doc__:
                    dct['__doc__'] = mro_cls.__doc__
                    break
        return type.__new__(cls, name, bases, dct)

    def __call__(cls, *args, **kwargs):
        try:
            t = cls._backend.__dict__[cls.__name__]
        except KeyError:
            warning('Backend %s does not appear to implement class %s'
                    % (cls._backend.__name__, cls.__name__))
            raise
        return t(*args, **kwargs)

    def __instancecheck__(cls, instance):
        # isinstance(x, op2.Dat) is true for the selected backend's Dat
        t = cls._backend.__dict__.get(cls.__name__)
        if t is not None and isinstance(instance, t):
            return True
        return type.__instancecheck__(cls, instance)

    def fromhdf5(cls, *args, **kwargs):
        try:
            return cls._backend.__dict__[cls.__name__].fromhdf5(*args, **kwargs)
        except AttributeError:
            warning("op2 object %s does not implement fromhdf5 method" % cls.__name__)
            raise


def get_backend():
    """Name of the selected backend ('void' before init, 'finalised' after
    exit)."""
    return _BackendSelector._name


def set_backend(backend):
    """Select the backend (once)."""
    if _BackendSelector._backend is not void:
        raise RuntimeError("The backend can only be set once!")
    mod = backends.get(backend)
    if mod is None:
        try:
            mod = importlib.import_module('pyop2_amd.%s' % backend)
        except ImportError:
            warning('Unable to import backend %s' % backend)
            raise
    backends[backend] = mod
    _BackendSelector._backend = mod
    _BackendSelector._name = backend


def unset_backend():
    """Mark op2 as finalised."""
    _BackendSelector._backend = finalised
    _BackendSelector._name = 'finalised'
