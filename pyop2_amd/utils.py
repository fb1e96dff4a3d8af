"""Common utilities: argument validation, reshaping, the command-line parser
and kernel preprocessing (reference: pyop2/utils.py)."""

import argparse
import functools
import inspect
import os
import subprocess

import numpy as np

from .exceptions import DataTypeError, DataValueError


def as_tuple(item, type=None, length=None):
    """Turn ``item`` into a tuple (reference: utils.py:48-63).

    ``None`` becomes the empty tuple, an iterable becomes a tuple of its items
    and anything else is repeated ``length`` (default 1) times."""
    if item is None:
        t = ()
    else:
        try:
            t = tuple(item)
        except TypeError:
            t = (item,) * (length or 1)
    if length and not len(t) == length:
        raise ValueError("Tuple needs to be of length %d" % length)
    if type and not all(isinstance(i, type) for i in t):
        raise TypeError("Items need to be of type %s" % type)
    return t


def as_type(obj, typ):
    """Return ``obj`` if it has dtype ``typ``, else a cast copy
    (reference: utils.py:65-77)."""
    try:
        return obj if obj.dtype == typ else obj.astype(typ)
    except AttributeError:
        if isinstance(obj, (int, np.integer)):
            return np.int64(obj).astype(typ)
        elif isinstance(obj, (float, np.floating)):
            return np.float64(obj).astype(typ)
        raise TypeError("Invalid type %s" % type(obj))


class _validate_base(object):
    """Decorator validating named arguments of the decorated function.

    Formal parameters the function does not have, and actual arguments that
    are not given, are ignored; an argument equal to its declared default is
    accepted (reference: utils.py:79-121)."""

    def __init__(self, *checks):
        self._checks = checks

    def __call__(self, f):
        sig = inspect.signature(f)
        names = list(sig.parameters)
        defaults = {k: p.default for k, p in sig.parameters.items()
                    if p.default is not inspect.Parameter.empty}
        code = f.__code__
        where = (code.co_filename, code.co_firstlineno + 1)

        # (name, position, has default, default, condition, exception) of
        # each check on a parameter the function has, resolved once
        checks = [(a, names.index(a), a in defaults, defaults.get(a), cond, exc,
                   self.quick(cond))
                  for a, cond, exc in self._checks if a in names]
        check_arg = self.check_arg

        @functools.wraps(f)
        def wrapper(*args, **kwargs):
            for argname, i, has_default, default, cond, exc, quick in checks:
                if argname in kwargs:
                    arg = kwargs[argname]
                elif i < len(args):
                    arg = args[i]
                else:
                    continue
                if quick(arg):
                    continue
                if has_default:
                    try:
                        if arg is default or arg == default:
                            continue
                    except Exception:
                        pass
                check_arg(arg, cond, exc, where)
            return f(*args, **kwargs)
        return wrapper


    def quick(self, cond):
        """A predicate accepting (some) valid arguments cheaply; arguments
        it does not accept go through the full check."""
        return lambda arg: False


class validate_type(_validate_base):
    """Check that named arguments have the given types."""

    def quick(self, argtype):
        return lambda arg: isinstance(arg, argtype)

    def check_arg(self, arg, argtype, exception, where):
        if not isinstance(arg, argtype):
            raise exception("%s:%d Parameter %s must be of type %r"
                            % (where[0], where[1], arg, argtype))


class validate_in(_validate_base):
    """Check that named arguments are in a list of legal values."""

    def quick(self, values):
        ids = frozenset(id(v) for v in values)
        return lambda arg: id(arg) in ids

    def check_arg(self, arg, values, exception, where):
        if not any(arg is v or arg == v for v in values):
            raise exception("%s:%d %s must be one of %s"
                            % (where[0], where[1], arg, values))


class validate_range(_validate_base):
    """Check that named arguments lie in a closed numeric range."""

    def check_arg(self, arg, rng, exception, where):
        if not rng[0] <= arg <= rng[1]:
            raise exception("%s:%d %s must be within range %s"
                            % (where[0], where[1], arg, rng))


def verify_reshape(data, dtype, shape, allow_none=False):
    """Convert ``data`` to an array of ``dtype`` and reshape it in place to
    ``shape`` (reference: utils.py:164-187).

    As in the reference, a caller's contiguous numpy array of the right dtype
    is reshaped destructively, not copied."""
    try:
        t = np.dtype(dtype) if dtype is not None else None
    except TypeError:
        raise DataTypeError("Invalid data type: %s" % dtype)
    if data is None and allow_none:
        return np.asarray([], dtype=t)
    elif data is None:
        raise DataValueError("Invalid data: None is not allowed!")
    try:
        a = np.asarray(data, dtype=t)
    except (ValueError, TypeError):
        raise DataValueError("Invalid data: cannot convert to %s!" % dtype)
    try:
        a.shape = shape
        return a
    except (ValueError, AttributeError):
        raise DataValueError("Invalid data: expected %d values, got %d!"
                             % (np.prod(shape), np.asarray(data).size))


def align(nbytes, alignment=16):
    """Round ``nbytes`` up to a multiple of ``alignment``."""
    return ((nbytes + alignment - 1) // alignment) * alignment


def uniquify(iterable):
    """Remove duplicates from ``iterable`` keeping the first occurrence."""
    seen = set()
    return [x for x in iterable if x not in seen and (seen.add(x) or True)]


def parser(description=None, group=False):
    """The default argument parser of op2 programs (reference:
    utils.py:198-224); the backend choices are the registered backends,
    ``hip`` by default."""
    p = argparse.ArgumentParser(description=description, add_help=True,
                                prefix_chars="-",
                                formatter_class=argparse.RawDescriptionHelpFormatter)
    g = p.add_argument_group('pyop2', 'backend configuration options') if group else p
    from . import backends
    names = sorted(set(['hip']) | set(k for k in backends.backends
                                      if k not in ('void', 'finalised')))
    g.add_argument('-b', '--backend', default='hip', choices=names,
                   help='select backend' if group else 'select pyop2 backend')
    g.add_argument('-d', '--debug', default=argparse.SUPPRESS, type=int,
                   choices=range(8),
                   help='set debug level' if group else 'set pyop2 debug level')
    g.add_argument('-l', '--log-level', default='WARN',
                   choices=['CRITICAL', 'ERROR', 'WARN', 'INFO', 'DEBUG'],
                   help='set logging level (default=WARN)')
    g.add_argument('-c', '--config', default=argparse.SUPPRESS,
                   type=argparse.FileType('r'),
                   help='specify alternate configuration')
    g.add_argument('--legacy-plan', dest='python_plan', action='store_false',
                   default=argparse.SUPPRESS, help='accepted and ignored')
    return p


def parse_args(*args, **kwargs):
    """Parse the command line with :func:`parser`."""
    return vars(parser(*args, **kwargs).parse_args())


def maybe_setflags(array, write=None, align=None, uic=None):
    """Set flags on a numpy array, but only touch the write flag of arrays
    that own their data (reference: utils.py:226-232)."""
    write = write if array.flags['OWNDATA'] else None
    array.setflags(write=write, align=align, uic=uic)


def preprocess(text):
    """Run the C preprocessor over a kernel string and drop line markers
    (reference: utils.py:241-245).

    Kernels without preprocessor directives are returned unchanged, which
    skips the subprocess for the common case."""
    if '#' not in text:
        return text
    p = subprocess.Popen(['cpp', '-E', '-P', '-I' + os.path.dirname(__file__)],
                         stdin=subprocess.PIPE, stdout=subprocess.PIPE,
                         universal_newlines=True)
    out = p.communicate(text)[0]
    return '\n'.join(l for l in out.split('\n') if not l.startswith('#'))
