"""Backend-independent op2 data model and ``par_loop`` semantics.

This is the object model every backend shares (reference: pyop2/base.py):
access descriptors, sets, maps, dats, globals, consts, sparsities, matrices,
kernels and parallel loops.  Backends subclass these classes to add device
storage and execution; the classes here hold host data as numpy arrays and
validate arguments with the exceptions of :mod:`exceptions`.
"""

import itertools
import operator
from hashlib import md5

import numpy as np

from .backends import _make_object
from .caching import Cached
from .exceptions import (DataTypeError, DataValueError, DimTypeError, IndexTypeError,
                         IndexValueError, MapTypeError, MapValueError, ModeValueError,
                         NameTypeError, SetTypeError, SizeTypeError, SparsityTypeError)
from .mpi import MPI
from .utils import (as_tuple, as_type, maybe_setflags, preprocess, validate_in,
                    validate_type, verify_reshape)
from . import configuration as cfg


def _new_arg(owner, **kwargs):
    """Build an :class:`Arg` for ``owner``: its class's own ``_arg_class``
    when it names one (backends usable without op2.init), else the selected
    backend's Arg."""
    cls = getattr(owner, '_arg_class', None)
    if cls is not None:
        return cls(**kwargs)
    return _make_object('Arg', **kwargs)


# Access descriptors (reference: base.py:56-90)

class Access(object):
    """How a :class:`Dat`, :class:`Global` or :class:`Mat` is accessed in a
    ``par_loop``.  Use the predefined values READ, WRITE, RW, INC, MIN, MAX."""

    _modes = ["READ", "WRITE", "RW", "INC", "MIN", "MAX"]

    @validate_in(('mode', _modes, ModeValueError))
    def __init__(self, mode):
        self._mode = mode

    def __str__(self):
        return "OP2 Access: %s" % self._mode

    def __repr__(self):
        return "Access('%s')" % self._mode


READ = Access("READ")
WRITE = Access("WRITE")
RW = Access("RW")
INC = Access("INC")
MIN = Access("MIN")
MAX = Access("MAX")


class Arg(object):
    """An argument to a ``par_loop``: a data carrier with the map (and index)
    it is reached through and its access mode (reference: base.py:94-259).

    User code builds these by calling a Dat, Global or Mat."""

    # a backend's key of this argument (it never changes once made)
    _fkey = None

    def __init__(self, data=None, map=None, idx=None, access=None):
        self._dat = data
        self._map = map
        self._idx = idx
        self._access = access
        self._in_flight = False

    def __str__(self):
        return "OP2 Arg: dat %s, map %s, index %s, access %s" % \
            (self._dat, self._map, self._idx, self._access)

    def __repr__(self):
        return "Arg(%r, %r, %r, %r)" % (self._dat, self._map, self._idx, self._access)

    @property
    def ctype(self):
        return self.data.ctype

    @property
    def dtype(self):
        return self.data.dtype

    @property
    def map(self):
        return self._map

    @property
    def idx(self):
        return self._idx

    @property
    def access(self):
        return self._access

    @property
    def data(self):
        return self._dat

    @property
    def _is_soa(self):
        return self._is_dat and self._dat.soa

    @property
    def _is_vec_map(self):
        return self._is_indirect and self._idx is None

    @property
    def _is_mat(self):
        return isinstance(self._dat, Mat)

    @property
    def _is_global(self):
        return isinstance(self._dat, Global)

    @property
    def _is_global_reduction(self):
        return self._is_global and self._access in [INC, MIN, MAX]

    @property
    def _is_dat(self):
        return isinstance(self._dat, Dat)

    @property
    def _is_INC(self):
        return self._access is INC

    @property
    def _is_MIN(self):
        return self._access is MIN

    @property
    def _is_MAX(self):
        return self._access is MAX

    @property
    def _is_direct(self):
        return isinstance(self._dat, Dat) and self._map is IdentityMap

    @property
    def _is_indirect(self):
        return isinstance(self._dat, Dat) and self._map not in [None, IdentityMap]

    @property
    def _is_indirect_and_not_read(self):
        return self._is_indirect and self._access is not READ

    @property
    def _is_indirect_reduction(self):
        return self._is_indirect and self._access is INC

    @property
    def _uses_itspace(self):
        return self._is_mat or isinstance(self.idx, IterationIndex)

    def halo_exchange_begin(self):
        """Halo exchanges are a no-op on one process."""
        assert self._is_dat, "Doing halo exchanges only makes sense for Dats"

    def halo_exchange_end(self):
        assert self._is_dat, "Doing halo exchanges only makes sense for Dats"

    def reduction_begin(self):
        """Start the cross-process reduction of a Global (identity on one
        process, reference: base.py:216-240)."""
        assert self._is_global, "Doing global reduction only makes sense for Globals"
        assert not self._in_flight, "Reduction already in flight for Arg %s" % self
        if self.access is not READ:
            self._in_flight = True
            MPI.comm.Allreduce(self.data._data, self.data._buf)

    def reduction_end(self):
        assert self._is_global, "Doing global reduction only makes sense for Globals"
        if self.access is not READ and self._in_flight:
            self._in_flight = False
            self.data._data = np.copy(self.data._buf)


class Set(object):
    """An op2 set (reference: base.py:261-400).

    ``size`` is an int, or four ints (core, owned, exec halo, non-exec halo)
    for a partitioned set; ``dim`` is the shape of the data a Dat on this set
    holds per element."""

    _globalcount = 0

    CORE_SIZE = 0
    OWNED_SIZE = 1
    IMPORT_EXEC_SIZE = 2
    IMPORT_NON_EXEC_SIZE = 3

    @classmethod
    def fromhdf5(cls, f, name, dim=1):
        """The Set named ``name`` in HDF5 data ``f`` (an h5py-like mapping,
        e.g. ``pyop2_amd.hdf5.File``): a one-element dataset holding the
        size, with an optional ``dim`` attribute (reference:
        base.py:386-394)."""
        slot = f[name]
        if tuple(slot.shape) != (1,):
            raise SizeTypeError("Shape of %s is incorrect" % name)
        size = np.asarray(slot[()]).astype(np.int64)
        dim = slot.attrs.get('dim', dim)
        dim = tuple(int(d) for d in np.atleast_1d(dim)) if np.ndim(dim) else int(dim)
        return cls(int(size[0]), dim, name)

    @validate_type(('size', (int, tuple, list), SizeTypeError),
                   ('name', str, NameTypeError))
    def __init__(self, size=None, dim=1, name=None, halo=None):
        if isinstance(size, (int, np.integer)):
            size = [int(size)] * 4
        size = as_tuple(size, int, 4)
        assert size[Set.CORE_SIZE] <= size[Set.OWNED_SIZE] <= \
            size[Set.IMPORT_EXEC_SIZE] <= size[Set.IMPORT_NON_EXEC_SIZE], \
            "Set received invalid sizes: %s" % (size,)
        self._core_size = size[Set.CORE_SIZE]
        self._size = size[Set.OWNED_SIZE]
        self._ieh_size = size[Set.IMPORT_EXEC_SIZE]
        self._inh_size = size[Set.IMPORT_NON_EXEC_SIZE]
        self._dim = as_tuple(dim, int)
        self._cdim = int(np.prod(self._dim))
        self._name = name or "set_%d" % Set._globalcount
        self._halo = halo
        if self.halo:
            self.halo.verify(self)
        Set._globalcount += 1

    def __call__(self, *dims):
        return IterationSpace(self, dims)

    @property
    def core_size(self):
        return self._core_size

    @property
    def size(self):
        return self._size

    @property
    def exec_size(self):
        return self._ieh_size

    @property
    def total_size(self):
        return self._inh_size

    @property
    def sizes(self):
        return self._core_size, self._size, self._ieh_size, self._inh_size

    @property
    def dim(self):
        return self._dim

    @property
    def cdim(self):
        return self._cdim

    @property
    def name(self):
        return self._name

    @property
    def halo(self):
        return self._halo

    def __str__(self):
        return "OP2 Set: %s with size %s, dim %s" % (self._name, self._size, self._dim)

    def __repr__(self):
        return "Set(%r, %r, %r)" % (self._size, self._dim, self._name)


class Halo(object):
    """Send/receive lists of a partitioned set (reference: base.py:402-505).

    Only the single-process case runs here: every list must be empty for
    the local rank."""

    def __init__(self, sends, receives, comm=None, gnn2unn=None):
        self._sends = tuple(np.asarray(x, dtype=np.int32) for x in sends)
        self._receives = tuple(np.asarray(x, dtype=np.int32) for x in receives)
        self._global_to_petsc_numbering = gnn2unn
        self._comm = comm if comm is not None else MPI.comm
        rank = self._comm.rank
        size = self._comm.size
        assert len(self._sends) == size, \
            "Invalid number of sends for Halo, got %d, wanted %d" % (len(self._sends), size)
        assert len(self._receives) == size, \
            "Invalid number of receives for Halo, got %d, wanted %d" % (len(self._receives), size)
        assert self._sends[rank].size == 0, "Halo was specified with self-sends on rank %d" % rank
        assert self._receives[rank].size == 0, "Halo was specified with self-receives on rank %d" % rank

    @property
    def sends(self):
        return self._sends

    @property
    def receives(self):
        return self._receives

    @property
    def global_to_petsc_numbering(self):
        return self._global_to_petsc_numbering

    @property
    def comm(self):
        return self._comm

    def verify(self, s):
        for dest, sends in enumerate(self.sends):
            assert (sends >= 0).all() and (sends < s.size).all(), \
                "Halo send to %d is invalid (outside owned elements)" % dest
        for source, receives in enumerate(self.receives):
            assert (receives >= s.size).all() and (receives < s.total_size).all(), \
                "Halo receive from %d is invalid (not in halo elements)" % source


class IterationSpace(object):
    """A set plus the extents of a local iteration space per element
    (reference: base.py:507-564)."""

    @validate_type(('iterset', Set, SetTypeError))
    def __init__(self, iterset, extents=()):
        self._iterset = iterset
        self._extents = as_tuple(extents, int)

    @property
    def iterset(self):
        return self._iterset

    @property
    def extents(self):
        return self._extents

    @property
    def name(self):
        return self._iterset.name

    @property
    def core_size(self):
        return self._iterset.core_size

    @property
    def size(self):
        return self._iterset.size

    @property
    def exec_size(self):
        return self._iterset.exec_size

    @property
    def total_size(self):
        return self._iterset.total_size

    @property
    def _extent_ranges(self):
        return [e for e in self.extents]

    def __str__(self):
        return "OP2 Iteration Space: %s with extents %s" % (self._iterset, self._extents)

    def __repr__(self):
        return "IterationSpace(%r, %r)" % (self._iterset, self._extents)


_CTYPES = {"bool": "unsigned char", "int": "int", "int8": "char", "int16": "short",
           "int32": "int", "int64": "long long", "uint8": "unsigned char",
           "uint16": "unsigned short", "uint32": "unsigned int",
           "uint64": "unsigned long", "float": "double", "float32": "float",
           "float64": "double"}


class DataCarrier(object):
    """Base of Const/Global (rank 0), Dat (rank 1) and Mat (rank 2)."""

    @property
    def dtype(self):
        return self._data.dtype

    @property
    def ctype(self):
        """The C type of the data (reference: base.py:573-592)."""
        return _CTYPES[self.dtype.name]

    @property
    def name(self):
        return self._name

    @property
    def dim(self):
        return self._dim

    @property
    def cdim(self):
        return self._cdim


class Dat(DataCarrier):
    """Data on every element of a set, stored array-of-structs with shape
    ``(total_size,) + dataset.dim`` (reference: base.py:612-855)."""

    _globalcount = 0
    _modes = [READ, WRITE, RW, INC]

    @classmethod
    def fromhdf5(cls, dataset, f, name):
        """The Dat named ``name`` in HDF5 data ``f`` on ``dataset``; its
        ``type`` attribute marks structure-of-arrays data with ':soa'
        (reference: base.py:842-849)."""
        slot = f[name]
        t = slot.attrs['type']
        if isinstance(t, bytes):
            t = t.decode('utf-8')
        return cls(dataset, np.asarray(slot[()]), name=name, soa=str(t).find(':soa') > 0)

    @validate_type(('dataset', Set, SetTypeError), ('name', str, NameTypeError))
    def __init__(self, dataset, data=None, dtype=None, name=None, soa=None, uid=None):
        if data is None:
            try:
                dt = np.dtype(dtype if dtype is not None else np.float64)
            except TypeError:
                raise DataTypeError("Invalid data type: %s" % dtype)
            data = np.zeros(dataset.total_size * dataset.cdim, dtype=dt)
        self._dataset = dataset
        self._data = verify_reshape(data, dtype, (dataset.total_size,) + dataset.dim,
                                    allow_none=True)
        self._soa = bool(soa)
        self._needs_halo_update = False
        if uid is None:
            self._id = Dat._globalcount
            Dat._globalcount += 1
        else:
            self._id = uid
        self._name = name or "dat_%d" % self._id

    _argmemo = None

    def __call__(self, path, access):
        # a time loop builds the same arguments every step: one Arg per
        # (Map or Map index, access), kept with the objects it was made from
        memo = self._argmemo
        if memo is not None:
            hit = memo.get((id(path), id(access)))
            if hit is not None and hit[0] is path and hit[1] is access:
                return hit[2]
        a = self._arg(canonical(path), canonical(access))
        if isinstance(path, Map) or getattr(path, '_template', False):
            if memo is None:
                memo = self._argmemo = {}
            memo[(id(path), id(access))] = (path, access, a)
        return a

    @validate_in(('access', _modes, ModeValueError))
    def _arg(self, path, access):
        if isinstance(path, Map):
            if path._dataset != self._dataset and path != IdentityMap:
                raise MapValueError("Dataset of Map does not match Dataset of Dat.")
            return _new_arg(self, data=self, map=path, access=access)
        return _new_arg(self, data=self, map=path._map, idx=path._idx, access=access)

    @property
    def dataset(self):
        return self._dataset

    @property
    def dim(self):
        return self.dataset.dim

    @property
    def cdim(self):
        return self.dataset.cdim

    @property
    def soa(self):
        return self._soa

    # Backends with device storage override these two hooks.
    def _sync_to_host(self):
        """Make the host array current."""

    def _host_written(self):
        """Note that the host array may have been modified."""

    @property
    def data(self):
        """Numpy array of the values (writable)."""
        if self.dataset.total_size > 0 and self._data.size == 0:
            raise RuntimeError("Illegal access: no data associated with this Dat!")
        self._sync_to_host()
        maybe_setflags(self._data, write=True)
        self._host_written()
        self.needs_halo_update = True
        return self._data

    @property
    def data_ro(self):
        """Numpy array of the values (read-only)."""
        if self.dataset.total_size > 0 and self._data.size == 0:
            raise RuntimeError("Illegal access: no data associated with this Dat!")
        self._sync_to_host()
        maybe_setflags(self._data, write=False)
        return self._data

    @property
    def needs_halo_update(self):
        return self._needs_halo_update

    @needs_halo_update.setter
    def needs_halo_update(self, val):
        self._needs_halo_update = val

    def zero(self):
        """Zero the data with a WRITE par_loop (reference: base.py:726-736)."""
        if not hasattr(self, '_zero_kernel'):
            k = """void zero(%(t)s *dat) {
                for (int n = 0; n < %(dim)s; ++n) {
                    dat[n] = (%(t)s)0;
                }
            }""" % {'t': self.ctype, 'dim': self.cdim}
            self._zero_kernel = _make_object('Kernel', k, 'zero')
        _make_object('ParLoop', self._zero_kernel, self.dataset,
                     self(IdentityMap, WRITE)).compute()

    def __str__(self):
        return "OP2 Dat: %s on (%s) with datatype %s" \
            % (self._name, self._dataset, self._data.dtype.name)

    def __repr__(self):
        return "Dat(%r, '%s', None, '%s')" % (self._dataset, self._data.dtype, self._name)

    def _check_shape(self, other):
        pass

    def _op(self, other, op):
        if np.isscalar(other):
            return _make_object('Dat', self.dataset,
                                op(self.data_ro, as_type(other, self.dtype)), self.dtype)
        self._check_shape(other)
        return _make_object('Dat', self.dataset,
                            op(self.data_ro, as_type(other.data_ro, self.dtype)), self.dtype)

    _BINARY = {operator.iadd: operator.add, operator.isub: operator.sub,
               operator.imul: operator.mul, operator.itruediv: operator.truediv}

    def _iop(self, other, op):
        if np.isscalar(other):
            val = as_type(other, self.dtype)
        else:
            self._check_shape(other)
            val = as_type(other.data_ro, self.dtype)
        tgt = self.data
        res = self._BINARY.get(op, op)(tgt, val)
        if res.shape != tgt.shape:
            raise ValueError("operands of shapes %s and %s do not combine in place"
                             % (tgt.shape, np.shape(val)))
        # the Dat keeps its dtype (integer division truncates)
        np.copyto(tgt, res, casting='unsafe')
        return self

    def __add__(self, other):
        return self._op(other, operator.add)

    def __sub__(self, other):
        return self._op(other, operator.sub)

    def __mul__(self, other):
        return self._op(other, operator.mul)

    def __truediv__(self, other):
        return self._op(other, operator.truediv)

    __div__ = __truediv__

    def __iadd__(self, other):
        return self._iop(other, operator.iadd)

    def __isub__(self, other):
        return self._iop(other, operator.isub)

    def __imul__(self, other):
        return self._iop(other, operator.imul)

    def __itruediv__(self, other):
        return self._iop(other, operator.itruediv)

    __idiv__ = __itruediv__

    def halo_exchange_begin(self):
        pass

    def halo_exchange_end(self):
        pass

    @property
    def norm(self):
        """The L2 norm of the flattened vector."""
        return np.linalg.norm(self.data_ro)


class Const(DataCarrier):
    """Data constant over every element of every set, visible to kernels as a
    global C variable (reference: base.py:857-926)."""

    class NonUniqueNameError(ValueError):
        """Const names are globally scoped and must be unique."""

    _defs = set()
    # bumped when a Const is declared or removed; the sorted list is cached
    _defs_version = 0
    _sorted = (-1, [])
    _globalcount = 0

    @classmethod
    def fromhdf5(cls, f, name):
        """The Const named ``name`` in HDF5 data ``f``; its dim is the
        dataset's shape (reference: base.py:918-926)."""
        slot = f[name]
        dim = tuple(slot.shape)
        if len(dim) < 1:
            raise DimTypeError("Invalid dimension value %s" % (dim,))
        return cls(dim, np.asarray(slot[()]), name)

    @validate_type(('name', str, NameTypeError))
    def __init__(self, dim, data=None, name=None, dtype=None):
        self._dim = as_tuple(dim, int)
        self._cdim = int(np.prod(self._dim))
        self._data = verify_reshape(data, dtype, self._dim, allow_none=True)
        self._name = name or "const_%d" % Const._globalcount
        if any(self._name == const._name for const in Const._defs):
            raise Const.NonUniqueNameError(
                "OP2 Constants are globally scoped, %s is already in use" % self._name)
        Const._defs.add(self)
        Const._defs_version += 1
        Const._globalcount += 1
        self._version = 0

    @property
    def data(self):
        if len(self._data) == 0:
            raise RuntimeError("Illegal access: No data associated with this Const!")
        return self._data

    @data.setter
    def data(self, value):
        self._data = verify_reshape(value, self.dtype, self.dim)
        self._version += 1

    def __str__(self):
        return "OP2 Const: %s of dim %s and type %s with value %s" \
            % (self._name, self._dim, self._data.dtype.name, self._data)

    def __repr__(self):
        return "Const(%s, %s, '%s')" % (self._dim, self._data, self._name)

    @classmethod
    def _definitions(cls):
        if Const._sorted[0] != Const._defs_version:
            Const._sorted = (Const._defs_version, sorted(Const._defs, key=lambda c: c.name))
        return list(Const._sorted[1])

    def remove_from_namespace(self):
        """Remove this Const so the name can be declared again."""
        Const._defs.discard(self)
        Const._defs_version += 1

    def _format_declaration(self):
        d = {'type': self.ctype, 'name': self.name, 'dim': self.cdim}
        if self.cdim == 1:
            return "static %(type)s %(name)s;" % d
        return "static %(type)s %(name)s[%(dim)s];" % d


class Global(DataCarrier):
    """A global value a kernel reads or reduces into
    (reference: base.py:928-977)."""

    _globalcount = 0
    _modes = [READ, INC, MIN, MAX]

    @validate_type(('name', str, NameTypeError))
    def __init__(self, dim, data=None, dtype=None, name=None):
        self._dim = as_tuple(dim, int)
        self._cdim = int(np.prod(self._dim))
        self._data = verify_reshape(data, dtype, self._dim, allow_none=True)
        self._buf = np.empty_like(self._data)
        self._name = name or "global_%d" % Global._globalcount
        Global._globalcount += 1

    def __call__(self, access):
        return self._arg(canonical(access))

    @validate_in(('access', _modes, ModeValueError))
    def _arg(self, access):
        return _new_arg(self, data=self, access=access)

    def __str__(self):
        return "OP2 Global Argument: %s with dim %s and value %s" \
            % (self._name, self._dim, self._data)

    def __repr__(self):
        return "Global('%s', %r, %r)" % (self._name, self._dim, self._data)

    def _sync_to_host(self):
        """Make the host value current (device backends override)."""

    def _host_written(self):
        """Note that the host value may have changed."""

    @property
    def data(self):
        self._sync_to_host()
        if len(self._data) == 0:
            raise RuntimeError("Illegal access: No data associated with this Global!")
        self._host_written()
        return self._data

    @data.setter
    def data(self, value):
        self._sync_to_host()
        self._data = verify_reshape(value, self.dtype, self.dim)
        self._host_written()

    @property
    def soa(self):
        return False


class IterationIndex(object):
    """Index into a local iteration space; use ``op2.i[n]``
    (reference: base.py:981-1016)."""

    def __init__(self, index=None):
        assert index is None or isinstance(index, int), "i must be an int"
        self._index = index

    def __str__(self):
        return "OP2 IterationIndex: %s" % self._index

    def __repr__(self):
        return "IterationIndex(%s)" % self._index

    @property
    def index(self):
        return self._index

    def __getitem__(self, idx):
        return IterationIndex(idx)

    def __iter__(self):
        yield self


i = IterationIndex()


class Map(object):
    """A relation from each element of ``iterset`` to ``dim`` elements of
    ``dataset``, stored as int32 of shape ``(iterset.total_size, dim)``
    (reference: base.py:1018-1125)."""

    _globalcount = 0

    @classmethod
    def fromhdf5(cls, iterset, dataset, f, name):
        """The Map named ``name`` in HDF5 data ``f``: a two-dimensional
        dataset whose second extent is the map's dim (reference:
        base.py:1117-1125)."""
        slot = f[name]
        dim = tuple(slot.shape)[1:]
        if len(dim) != 1:
            raise DimTypeError("Unrecognised dimension value %s" % (dim,))
        return cls(iterset, dataset, int(dim[0]), np.asarray(slot[()]), name)

    @validate_type(('iterset', Set, SetTypeError), ('dataset', Set, SetTypeError),
                   ('dim', int, DimTypeError), ('name', str, NameTypeError))
    def __init__(self, iterset, dataset, dim, values=None, name=None):
        self._iterset = iterset
        self._dataset = dataset
        self._dim = dim
        self._values = verify_reshape(values, np.int32, (iterset.total_size, dim),
                                      allow_none=True)
        self._name = name or "map_%d" % Map._globalcount
        self._uid = Map._globalcount
        Map._globalcount += 1

    _imemo = None

    def __getitem__(self, index):
        # one (Map, index) template per int or op2.i[n] index, so that the
        # Dat call it is passed to can reuse the Arg it made from it
        memo = self._imemo
        if memo is not None and type(index) is int:
            a = memo.get(index)
            if a is not None:
                return a
        t = type(index)
        key = index if t is int else ('i', index._index) if t is IterationIndex else None
        if key is not None:
            if memo is None:
                memo = self._imemo = {}
            a = memo.get(key)
            if a is None:
                a = self._index(index)
                a._template = True
                memo[key] = a
            return a
        if isinstance(index, slice):
            raise NotImplementedError("Slicing maps is not allowed")
        return self._index(canonical(index))

    @validate_type(('index', (int, IterationIndex), IndexTypeError))
    def _index(self, index):
        if isinstance(index, int) and not (0 <= index < self._dim):
            raise IndexValueError("Index must be in interval [0,%d]" % (self._dim - 1))
        if isinstance(index, IterationIndex) and index.index not in [0, 1]:
            raise IndexValueError("IterationIndex must be in interval [0,1]")
        return _new_arg(self, map=self, idx=index)

    def __iter__(self):
        yield self

    @property
    def iterset(self):
        return self._iterset

    @property
    def dataset(self):
        return self._dataset

    @property
    def dim(self):
        return self._dim

    @property
    def values(self):
        return self._values

    @property
    def name(self):
        return self._name

    def __str__(self):
        return "OP2 Map: %s from (%s) to (%s) with dim %s" \
            % (self._name, self._iterset, self._dataset, self._dim)

    def __repr__(self):
        return "Map(%r, %r, %s, None, '%s')" \
            % (self._iterset, self._dataset, self._dim, self._name)

    def __eq__(self, o):
        try:
            return (self._iterset == o._iterset and self._dataset == o._dataset and
                    self._dim == o.dim and self._name == o.name)
        except AttributeError:
            return False

    def __ne__(self, o):
        return not self.__eq__(o)

    def __hash__(self):
        return hash((id(self._iterset), id(self._dataset), self._dim, self._name))

    def __lt__(self, o):
        return self._uid < o._uid


IdentityMap = Map(Set(0), Set(0), 1, [], 'identity')

_ACCESS = {a._mode: a for a in (READ, WRITE, RW, INC, MIN, MAX)}


def canonical(obj):
    """This module's descriptor for one made by another implementation of
    the op2 API; other objects are returned unchanged.

    A backend plugged into the reference's own front end receives the
    reference's descriptors: its op2 module exports READ ... MAX,
    IdentityMap and i from its base module (pyop2/op2.py:43, defined at
    pyop2/base.py:56-90, 981-1011, 1127), and user code passes them to the
    Dats, Maps and Mats this backend made.  They are matched by role, not
    identity: an access descriptor by its mode name, an iteration index by
    its index, and the identity map as a map named 'identity' of arity 1
    from an empty Set.  Everything past this point compares descriptors by
    identity against this module's singletons."""
    if obj is None or isinstance(obj, (Access, IterationIndex, Map, Arg, int)):
        return obj
    kind = type(obj).__name__
    if kind == 'Access':
        return _ACCESS.get(getattr(obj, '_mode', None), obj)
    if kind == 'IterationIndex' and hasattr(obj, 'index'):
        idx = obj.index
        return i if idx is None else IterationIndex(idx)
    if kind == 'Map' and getattr(obj, 'name', None) == 'identity' and \
            getattr(obj, 'dim', None) == 1 and \
            getattr(getattr(obj, 'iterset', None), 'size', None) == 0:
        return IdentityMap
    return obj


class Sparsity(Cached):
    """The CSR structure of the union of outer products of (row map,
    column map) pairs (reference: base.py:1130-1312).

    The backend fills ``_rowptr``, ``_colidx``, ``_d_nnz`` and ``_d_nz`` in
    :meth:`_build` (reference: op_lib_core.pyx:544-596)."""

    _cache = {}
    _globalcount = 0

    @classmethod
    @validate_type(('maps', (Map, tuple), MapTypeError),)
    def _process_args(cls, maps, name=None, *args, **kwargs):
        assert not name or isinstance(name, str), "Name must be of type str"
        maps = (maps, maps) if isinstance(maps, Map) else maps
        maps = (maps,) if isinstance(maps[0], Map) else maps
        for pair in maps:
            for m in pair:
                if not isinstance(m, Map):
                    raise MapTypeError("All maps must be of type map, not type %r" % type(m))
                if len(m.values) == 0:
                    raise MapValueError("Unpopulated map values when trying to build sparsity.")
        return [tuple(sorted(maps, key=lambda p: (p[0]._uid, p[1]._uid))), name], {}

    @classmethod
    def _cache_key(cls, maps, *args, **kwargs):
        return (cls,) + tuple((id(a), id(b)) for a, b in maps)

    def __init__(self, maps, name=None):
        if self._initialized:
            return
        self._rmaps, self._cmaps = zip(*maps)
        self._maps_keepalive = maps
        for pair in maps:
            if pair[0].iterset is not pair[1].iterset:
                raise RuntimeError("Iterset of both maps in a pair must be the same")
        if not all(m.dataset is self._rmaps[0].dataset for m in self._rmaps):
            raise RuntimeError("Dataset of all row maps must be the same")
        if not all(m.dataset is self._cmaps[0].dataset for m in self._cmaps):
            raise RuntimeError("Dataset of all column maps must be the same")
        self._nrows = self._rmaps[0].dataset.size
        self._ncols = self._cmaps[0].dataset.size
        self._dims = (self._rmaps[0].dataset.cdim, self._cmaps[0].dataset.cdim)
        self._name = name or "sparsity_%d" % Sparsity._globalcount
        Sparsity._globalcount += 1
        self._build()
        self._initialized = True

    def _build(self):
        raise NotImplementedError("Sparsity construction is implemented by the backend")

    @property
    def _nmaps(self):
        return len(self._rmaps)

    @property
    def maps(self):
        return list(zip(self._rmaps, self._cmaps))

    @property
    def cmaps(self):
        return self._cmaps

    @property
    def rmaps(self):
        return self._rmaps

    @property
    def dims(self):
        return self._dims

    @property
    def nrows(self):
        return self._nrows

    @property
    def ncols(self):
        return self._ncols

    @property
    def name(self):
        return self._name

    def __str__(self):
        return "OP2 Sparsity: rmaps %s, cmaps %s, name %s" % (self._rmaps, self._cmaps, self._name)

    def __repr__(self):
        return "Sparsity(%r, %r)" % (tuple(self.maps), self.name)

    @property
    def rowptr(self):
        return self._rowptr

    @property
    def colidx(self):
        return self._colidx

    @property
    def nnz(self):
        return self._d_nnz

    @property
    def onnz(self):
        return self._o_nnz

    @property
    def nz(self):
        return int(self._d_nz)

    @property
    def onz(self):
        return int(self._o_nz)


class Mat(DataCarrier):
    """Matrix values on a :class:`Sparsity` (reference: base.py:1314-1395)."""

    _globalcount = 0
    _modes = [WRITE, INC]

    @validate_type(('sparsity', Sparsity, SparsityTypeError), ('name', str, NameTypeError))
    def __init__(self, sparsity, dtype=None, name=None):
        self._sparsity = sparsity
        self._datatype = np.dtype(dtype)
        self._name = name or "mat_%d" % Mat._globalcount
        Mat._globalcount += 1

    def __call__(self, path, access):
        return self._arg(path, canonical(access))

    @validate_in(('access', _modes, ModeValueError))
    def _arg(self, path, access):
        path = as_tuple(path, Arg, 2)
        path_maps = [arg.map for arg in path]
        path_idxs = [arg.idx for arg in path]
        if tuple(path_maps) not in [tuple(p) for p in self.sparsity.maps]:
            raise MapValueError("Path maps not in sparsity maps")
        return _new_arg(self, data=self, map=path_maps, access=access, idx=path_idxs)

    @property
    def dims(self):
        return self._sparsity._dims

    @property
    def sparsity(self):
        return self._sparsity

    @property
    def _is_scalar_field(self):
        return np.prod(self.dims) == 1

    @property
    def _is_vector_field(self):
        return not self._is_scalar_field

    @property
    def values(self):
        raise NotImplementedError("Abstract base Mat does not implement values()")

    @property
    def dtype(self):
        return self._datatype

    def __str__(self):
        return "OP2 Mat: %s, sparsity (%s), datatype %s" \
            % (self._name, self._sparsity, self._datatype.name)

    def __repr__(self):
        return "Mat(%r, '%s', '%s')" % (self._sparsity, self._datatype, self._name)


class Kernel(Cached):
    """A user kernel: C source plus the name of the function to call
    (reference: base.py:1399-1436).  Cached by md5(code + name)."""

    _globalcount = 0
    _cache = {}

    @classmethod
    @validate_type(('name', str, NameTypeError))
    def _cache_key(cls, code, name):
        return md5((code + name).encode()).hexdigest()

    def __init__(self, code, name):
        if self._initialized:
            return
        self._name = name or "kernel_%d" % Kernel._globalcount
        self._code = preprocess(code)
        Kernel._globalcount += 1
        self._initialized = True

    @property
    def name(self):
        return self._name

    @property
    def code(self):
        return self._code

    def __str__(self):
        return "OP2 Kernel: %s" % self._name

    def __repr__(self):
        return 'Kernel("""%s""", "%s")' % (self._code, self._name)


class JITModule(Cached):
    """Cached compiled loop for one kernel and argument signature
    (key as reference: base.py:1438-1471)."""

    _cache = {}

    @classmethod
    def _cache_key(cls, kernel, itspace_extents, *args, **kwargs):
        key = (cls, kernel.cache_key, itspace_extents)
        for arg in args:
            if arg._is_global:
                key += (arg.data.dim, arg.data.dtype, arg.access)
            elif arg._is_dat:
                if isinstance(arg.idx, IterationIndex):
                    idx = (arg.idx.__class__, arg.idx.index)
                else:
                    idx = arg.idx
                map_dim = None if arg.map is IdentityMap else arg.map.dim
                key += (arg.data.dim, arg.data.dtype, map_dim, idx, arg.access, arg._is_soa)
            elif arg._is_mat:
                idxs = (arg.idx[0].__class__, arg.idx[0].index, arg.idx[1].index)
                map_dims = (arg.map[0].dim, arg.map[1].dim)
                key += (arg.data.dims, arg.data.dtype, idxs, map_dims, arg.access)
        for c in Const._definitions():
            key += (c.name, c.dtype, c.cdim)
        return key


class ParLoop(object):
    """A kernel, its iteration space and its arguments
    (reference: base.py:1473-1589)."""

    def __init__(self, kernel, itspace, *args):
        self._actual_args = args
        self._kernel = kernel
        self._it_space = itspace if isinstance(itspace, IterationSpace) else IterationSpace(itspace)
        self.check_args()

    def compute(self):
        raise RuntimeError('Must select a backend')

    def halo_exchange_begin(self):
        if self.is_direct:
            return
        for arg in self.args:
            if arg._is_dat:
                arg.halo_exchange_begin()

    def halo_exchange_end(self):
        if self.is_direct:
            return
        for arg in self.args:
            if arg._is_dat:
                arg.halo_exchange_end()

    def reduction_begin(self):
        for arg in self.args:
            if arg._is_global_reduction:
                arg.reduction_begin()

    def reduction_end(self):
        for arg in self.args:
            if arg._is_global_reduction:
                arg.reduction_end()

    def maybe_set_halo_update_needed(self):
        for arg in self.args:
            if arg._is_dat and arg.access in [INC, WRITE, RW]:
                arg.data.needs_halo_update = True

    def check_args(self):
        """Each map must start at the loop's iteration set and end at its
        Dat's set; a :class:`MapValueError` names the offending argument
        (reference: base.py:1528-1550)."""
        iterset = self._it_space._iterset
        for i, arg in enumerate(self._actual_args):
            if arg._is_global or arg._map is IdentityMap:
                continue
            for j, m in enumerate(as_tuple(arg._map)):
                if m._iterset != iterset:
                    raise MapValueError(
                        "Iterset of arg %s map %s doesn't match ParLoop iterset." % (i, j))
                if arg._is_mat:
                    continue
                if m._dataset != arg.data._dataset:
                    raise MapValueError(
                        "Dataset of arg %s map %s doesn't match the set of its Dat." % (i, j))

    @property
    def it_space(self):
        return self._it_space

    @property
    def is_direct(self):
        return all(a.map in [None, IdentityMap] for a in self.args)

    @property
    def is_indirect(self):
        return not self.is_direct

    @property
    def needs_exec_halo(self):
        return any(arg._is_indirect_and_not_read or arg._is_mat for arg in self.args)

    @property
    def kernel(self):
        return self._kernel

    @property
    def args(self):
        return self._actual_args

    @property
    def _has_soa(self):
        return any(a._is_soa for a in self._actual_args)


DEFAULT_SOLVER_PARAMETERS = {'linear_solver': 'cg',
                             'preconditioner': 'jacobi',
                             'relative_tolerance': 1.0e-7,
                             'absolute_tolerance': 1.0e-50,
                             'divergence_tolerance': 1.0e+4,
                             'maximum_iterations': 1000,
                             'monitor_convergence': False,
                             'plot_convergence': False,
                             'plot_prefix': '',
                             'error_on_nonconvergence': True,
                             'gmres_restart': 30}


class Solver(object):
    """Linear solver parameters (reference: base.py:1591-1657); defaults are
    those of PETSc 3.3 (cg, jacobi, rtol 1e-7, atol 1e-50, dtol 1e4,
    max_it 1000)."""

    def __init__(self, parameters=None, **kwargs):
        self.parameters = DEFAULT_SOLVER_PARAMETERS.copy()
        if parameters and kwargs:
            raise RuntimeError("Solver options are set either by parameters or kwargs")
        if parameters:
            self.parameters.update(parameters)
        else:
            self.parameters.update(kwargs)

    def update_parameters(self, parameters):
        self.parameters.update(parameters)

    def solve(self, A, x, b):
        raise NotImplementedError("solve must be implemented by backend")
