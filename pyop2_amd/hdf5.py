"""A read-only HDF5 reader for mesh ingestion (``Set/Map/Dat/Const.fromhdf5``).

The reference reads meshes with h5py (``demo/airfoil.py:45-76``; the
``fromhdf5`` constructors, ``pyop2/base.py:386-394, 842-849, 918-926,
1117-1125``), which this environment does not have.  This module reads the
subset of the HDF5 file format (HDF Group, "HDF5 File Format
Specification", versions 2 and 3) that the HDF5 library's default settings
write, and offers the part of h5py's interface those constructors use:
``File(path)[name]`` gives a ``Group`` or a ``Dataset`` with ``shape``,
``dtype``, ``value`` / ``[()]`` and ``attrs``.

Supported: superblock versions 0-3; object headers version 1 and 2 with
continuation blocks; groups stored as symbol tables (B-tree version 1 and
local heap) or as compact link messages; dataspaces (scalar and simple);
fixed-point, floating-point and fixed-length string datatypes, either byte
order, and variable-length strings and numeric sequences (global heap; as
h5py stores Python strings); contiguous, compact and chunked (B-tree
version 1) layouts, with the deflate, shuffle and Fletcher-32 filters;
attributes with those types.  Anything else (dense link or attribute
storage, compound or nested variable-length types, chunked variable-length
data, other filters) raises NotImplementedError when it is reached.
"""

import mmap
import zlib

import numpy as np

_SIGNATURE = b'\x89HDF\r\n\x1a\n'

# object header message types
_MSG_DATASPACE = 0x0001
_MSG_LINK_INFO = 0x0002
_MSG_DATATYPE = 0x0003
_MSG_LINK = 0x0006
_MSG_LAYOUT = 0x0008
_MSG_FILTERS = 0x000B
_MSG_ATTRIBUTE = 0x000C
_MSG_CONTINUATION = 0x0010
_MSG_SYMBOL_TABLE = 0x0011
_MSG_ATTRIBUTE_INFO = 0x0015


class _Reader(object):
    """Little-endian field access into the mapped file."""

    def __init__(self, buf, offset_size, length_size, base=0):
        self.buf = buf
        self.so = offset_size
        self.sl = length_size
        self.base = base            # file addresses are relative to the superblock's base
        self.undef = (1 << (8 * offset_size)) - 1
        self._heaps = {}

    def heap_object(self, addr, index):
        """Bytes of object ``index`` of the global heap collection at
        ``addr`` (variable-length data)."""
        objs = self._heaps.get(addr)
        if objs is None:
            if self.bytes(addr, 4) != b'GCOL':
                raise ValueError("HDF5: bad global heap collection")
            end = addr + self.length(addr + 8)
            p = addr + 8 + self.sl
            objs = {}
            while p + 8 + self.sl <= end:
                idx = self.u(p, 2)
                if idx == 0:                 # free space: the rest of the collection
                    break
                size = self.length(p + 8)
                objs[idx] = (p + 8 + self.sl, size)
                p += 8 + self.sl + _pad8(size)
            self._heaps[addr] = objs
        pos, size = objs[index]
        return self.bytes(pos, size)

    def u(self, pos, size):
        return int.from_bytes(self.buf[pos:pos + size], 'little')

    def off(self, pos):
        v = self.u(pos, self.so)
        return v if v == self.undef else v + self.base

    def length(self, pos):
        return self.u(pos, self.sl)

    def bytes(self, pos, size):
        return bytes(self.buf[pos:pos + size])

    def cstr(self, pos):
        end = self.buf.find(b'\0', pos)
        return bytes(self.buf[pos:end]).decode('utf-8', 'replace')


def _pad8(n):
    return (n + 7) & ~7


# -- datatypes and dataspaces ---------------------------------------------------------

class _Datatype(object):
    """A decoded datatype message: ``dtype`` (numpy) and, for strings, the
    padding rule."""

    def __init__(self, r, pos):
        cv = r.u(pos, 1)
        self.cls, self.version = cv & 0x0F, cv >> 4
        bits = r.u(pos + 1, 3)
        self.size = r.u(pos + 4, 4)
        self.string = False
        self.vlen = False
        if self.cls == 0:                      # fixed-point
            order = '>' if bits & 1 else '<'
            kind = 'i' if bits & 0x8 else 'u'
            self.dtype = np.dtype('%s%s%d' % (order, kind, self.size))
        elif self.cls == 1:                    # floating-point
            if bits & 0x40:
                raise NotImplementedError("HDF5: VAX-order floating point")
            order = '>' if bits & 1 else '<'
            if self.size not in (2, 4, 8):
                raise NotImplementedError("HDF5: %d-byte floating point" % self.size)
            self.dtype = np.dtype('%sf%d' % (order, self.size))
        elif self.cls == 3:                    # fixed-length string
            self.dtype = np.dtype('S%d' % self.size)
            self.string = True
        elif self.cls == 9:                    # variable-length sequence or string
            self.vlen_string = (bits & 0x0F) == 1
            self.base = None if self.vlen_string else _Datatype(r, pos + 8)
            if self.base is not None and (self.base.vlen or self.base.string):
                raise NotImplementedError("HDF5: nested variable-length datatype")
            self.dtype = np.dtype(object)
            self.vlen = True
        else:
            names = {2: 'time', 4: 'bitfield', 5: 'opaque', 6: 'compound', 7: 'reference',
                     8: 'enumerated', 10: 'array'}
            raise NotImplementedError("HDF5: %s datatype" % names.get(self.cls, self.cls))


def _dataspace(r, pos):
    """Shape of a dataspace message (``()`` for a scalar, None for null)."""
    version = r.u(pos, 1)
    ndims = r.u(pos + 1, 1)
    if version == 1:
        p = pos + 8
        kind = 1 if ndims else 0
    elif version == 2:
        kind = r.u(pos + 3, 1)
        p = pos + 4
    else:
        raise NotImplementedError("HDF5: dataspace version %d" % version)
    if kind == 2:
        return None
    return tuple(r.length(p + k * r.sl) for k in range(ndims))


def _decode(r, raw, dt, shape):
    """numpy array (native byte order) from raw bytes of the given type;
    variable-length elements (a 4-byte length and a global heap reference
    each) become an object array of str or of 1-D arrays."""
    count = int(np.prod(shape)) if shape else 1
    if dt.vlen:
        out = np.empty(count, dtype=object)
        step = 4 + r.so + 4
        for k in range(count):
            q = k * step
            n = int.from_bytes(raw[q:q + 4], 'little')
            addr = int.from_bytes(raw[q + 4:q + 4 + r.so], 'little')
            idx = int.from_bytes(raw[q + 4 + r.so:q + step], 'little')
            data = b'' if n == 0 or addr in (0, r.undef) else \
                r.heap_object(addr + r.base, idx)
            if dt.vlen_string:
                out[k] = data[:n].rstrip(b'\0').decode('utf-8', 'replace')
            else:
                out[k] = np.frombuffer(data, dtype=dt.base.dtype, count=n).astype(
                    dt.base.dtype.newbyteorder('='))
        return out.reshape(shape)
    a = np.frombuffer(raw, dtype=dt.dtype, count=count)
    if dt.string:
        return a.reshape(shape)
    return a.astype(dt.dtype.newbyteorder('='), copy=True).reshape(shape)


# -- object headers -----------------------------------------------------------------------

class _Message(object):
    __slots__ = ('type', 'pos', 'size', 'flags')

    def __init__(self, mtype, pos, size, flags):
        self.type, self.pos, self.size, self.flags = mtype, pos, size, flags


def _object_header(r, addr):
    """Messages of the object header at ``addr`` (continuations followed)."""
    msgs = []
    if r.bytes(addr, 4) == b'OHDR':
        version = r.u(addr + 4, 1)
        if version != 2:
            raise NotImplementedError("HDF5: object header version %d" % version)
        flags = r.u(addr + 5, 1)
        p = addr + 6
        if flags & 0x20:
            p += 16
        if flags & 0x10:
            p += 4
        width = 1 << (flags & 3)
        size0 = r.u(p, width)
        p += width
        blocks = [(p, p + size0)]
        order = bool(flags & 0x04)
        while blocks:
            start, end = blocks.pop(0)
            q = start
            while q + 4 <= end:
                mtype = r.u(q, 1)
                msize = r.u(q + 1, 2)
                mflags = r.u(q + 3, 1)
                q += 4 + (2 if order else 0)
                if mtype == _MSG_CONTINUATION:
                    caddr, clen = r.off(q), r.length(q + r.so)
                    if r.bytes(caddr, 4) != b'OCHK':
                        raise ValueError("HDF5: bad continuation block")
                    blocks.append((caddr + 4, caddr + clen - 4))
                elif mtype != 0:
                    msgs.append(_Message(mtype, q, msize, mflags))
                q += msize
        return msgs
    version = r.u(addr, 1)
    if version != 1:
        raise NotImplementedError("HDF5: object header version %d" % version)
    nmsgs = r.u(addr + 2, 2)
    size0 = r.u(addr + 8, 4)
    blocks = [(addr + 16, addr + 16 + size0)]
    while blocks and len(msgs) < nmsgs:
        start, end = blocks.pop(0)
        q = start
        while q + 8 <= end:
            mtype = r.u(q, 2)
            msize = r.u(q + 2, 2)
            mflags = r.u(q + 4, 1)
            q += 8
            if mtype == _MSG_CONTINUATION:
                blocks.append((r.off(q), r.off(q) + r.length(q + r.so)))
            elif mtype != 0:
                msgs.append(_Message(mtype, q, msize, mflags))
            q += msize
    return msgs


def _resolve_shared(r, msg, mtype):
    """Position of the message body, following a shared-message reference
    (a committed datatype) to the object header that holds it."""
    if not msg.flags & 0x02:
        return msg.pos
    version = r.u(msg.pos, 1)
    if version == 1:
        addr = r.off(msg.pos + 8)
    elif version == 2:
        addr = r.off(msg.pos + 2)
    elif version == 3 and r.u(msg.pos + 1, 1) == 2:
        addr = r.off(msg.pos + 2)
    else:
        raise NotImplementedError("HDF5: shared message in the shared-message heap")
    for m in _object_header(r, addr):
        if m.type == mtype:
            return _resolve_shared(r, m, mtype)
    raise ValueError("HDF5: shared message not found")


# -- groups ---------------------------------------------------------------------------------

def _symbol_table_links(r, btree, heap):
    """name -> object header address of an old-style (symbol table) group."""
    if r.bytes(heap, 4) != b'HEAP':
        raise ValueError("HDF5: bad local heap")
    data = r.off(heap + 8 + 2 * r.sl)
    out = {}

    def node(addr):
        if r.bytes(addr, 4) != b'TREE':
            raise ValueError("HDF5: bad group B-tree node")
        level = r.u(addr + 5, 1)
        used = r.u(addr + 6, 2)
        p = addr + 8 + 2 * r.so + r.sl       # after the first key
        for _ in range(used):
            child = r.off(p)
            p += r.so + r.sl
            if level > 0:
                node(child)
            else:
                snod(child)

    def snod(addr):
        if r.bytes(addr, 4) != b'SNOD':
            raise ValueError("HDF5: bad symbol table node")
        n = r.u(addr + 6, 2)
        p = addr + 8
        for _ in range(n):
            # cache type 2 is a soft link (a path, not an object): not followed
            if r.u(p + 2 * r.so, 4) != 2:
                out[r.cstr(data + r.off(p))] = r.off(p + r.so)
            p += 2 * r.so + 24

    node(btree)
    return out


def _link(r, pos):
    """(name, object header address or None) of a link message."""
    flags = r.u(pos + 1, 1)
    p = pos + 2
    ltype = 0
    if flags & 0x08:
        ltype = r.u(p, 1)
        p += 1
    if flags & 0x04:
        p += 8
    if flags & 0x10:
        p += 1
    width = 1 << (flags & 3)
    nlen = r.u(p, width)
    p += width
    name = r.bytes(p, nlen).decode('utf-8')
    p += nlen
    return name, (r.off(p) if ltype == 0 else None)


# -- objects --------------------------------------------------------------------------------

class Attributes(object):
    """The attributes of an object, decoded on access (h5py's ``attrs``)."""

    def __init__(self, r, msgs):
        self._r = r
        self._pos = {}
        for m in msgs:
            if m.type == _MSG_ATTRIBUTE:
                self._pos[self._name(m.pos)] = m.pos
            elif m.type == _MSG_ATTRIBUTE_INFO:
                if r.off(m.pos + 2 + (2 if r.u(m.pos + 1, 1) & 1 else 0)) != r.undef:
                    raise NotImplementedError("HDF5: dense attribute storage")

    def _name(self, pos):
        r = self._r
        version = r.u(pos, 1)
        nsize = r.u(pos + 2, 2)
        start = pos + (8 if version == 1 else (8 if version == 2 else 9))
        return r.bytes(start, nsize).rstrip(b'\0').decode('utf-8')

    def __getitem__(self, name):
        r = self._r
        pos = self._pos[name]
        version = r.u(pos, 1)
        nsize, tsize, ssize = r.u(pos + 2, 2), r.u(pos + 4, 2), r.u(pos + 6, 2)
        if version == 1:
            p = pos + 8 + _pad8(nsize)
            tpos = p
            p += _pad8(tsize)
            spos = p
            p += _pad8(ssize)
        elif version in (2, 3):
            p = pos + (8 if version == 2 else 9) + nsize
            tpos = p
            spos = p + tsize
            p = spos + ssize
        else:
            raise NotImplementedError("HDF5: attribute version %d" % version)
        dt = _Datatype(r, tpos)
        shape = _dataspace(r, spos)
        if shape is None:
            return None
        count = int(np.prod(shape)) if shape else 1
        a = _decode(r, r.bytes(p, count * dt.size), dt, shape)
        if dt.vlen:
            return a[()] if a.shape == () else a
        if dt.string:
            a = np.char.decode(np.char.rstrip(a.astype('S'), b'\0 '), 'utf-8')
            return str(a[()]) if a.shape == () else a
        return a[()] if a.shape == () else a

    def get(self, name, default=None):
        return self[name] if name in self._pos else default

    def __contains__(self, name):
        return name in self._pos

    def keys(self):
        return list(self._pos)

    def __iter__(self):
        return iter(self._pos)

    def __len__(self):
        return len(self._pos)


class Dataset(object):
    """An HDF5 dataset: ``shape``, ``dtype``, ``attrs``; ``value`` or
    ``ds[()]`` reads it whole into a numpy array."""

    def __init__(self, r, name, msgs):
        self._r = r
        self.name = name
        self._msgs = {}
        for m in msgs:
            self._msgs.setdefault(m.type, m)
        self.attrs = Attributes(r, msgs)
        self._dt = _Datatype(r, _resolve_shared(r, self._msgs[_MSG_DATATYPE], _MSG_DATATYPE))
        self.shape = _dataspace(r, self._msgs[_MSG_DATASPACE].pos)
        self.dtype = self._dt.dtype.newbyteorder('=') if not self._dt.string else self._dt.dtype

    @property
    def ndim(self):
        return len(self.shape)

    @property
    def size(self):
        return int(np.prod(self.shape)) if self.shape else 1

    def __len__(self):
        if not self.shape:
            raise TypeError("len() of a scalar dataset")
        return self.shape[0]

    def __getitem__(self, key):
        return self.value[key]

    def __array__(self, dtype=None):
        v = self.value
        return v if dtype is None else v.astype(dtype)

    @property
    def value(self):
        r = self._r
        if self.shape is None:
            return None
        lay = self._msgs[_MSG_LAYOUT].pos
        version = r.u(lay, 1)
        nbytes = self.size * self._dt.size
        if version in (1, 2):
            ndims, cls = r.u(lay + 1, 1), r.u(lay + 2, 1)
            p = lay + 8
            if cls == 0:
                p += 4 * ndims
                raw = r.bytes(p + 4, r.u(p, 4))
            elif cls == 1:
                addr = r.off(p)
                raw = self._contiguous(addr, nbytes)
            else:
                addr = r.off(p)
                dims = [r.u(p + r.so + 4 * k, 4) for k in range(ndims)]
                return self._chunked(addr, dims)
        elif version in (3, 4):
            cls = r.u(lay + 1, 1)
            if cls == 0:
                raw = r.bytes(lay + 4, r.u(lay + 2, 2))
            elif cls == 1:
                raw = self._contiguous(r.off(lay + 2), nbytes)
            elif cls == 2 and version == 3:
                ndims = r.u(lay + 2, 1)
                addr = r.off(lay + 3)
                dims = [r.u(lay + 3 + r.so + 4 * k, 4) for k in range(ndims)]
                return self._chunked(addr, dims)
            else:
                raise NotImplementedError("HDF5: layout version %d class %d" % (version, cls))
        else:
            raise NotImplementedError("HDF5: layout version %d" % version)
        return _decode(r, raw, self._dt, self.shape)

    def _contiguous(self, addr, nbytes):
        if addr == self._r.undef:
            return b'\0' * nbytes
        return self._r.bytes(addr, nbytes)

    def _filters(self):
        m = self._msgs.get(_MSG_FILTERS)
        if m is None:
            return []
        r = self._r
        version, n = r.u(m.pos, 1), r.u(m.pos + 1, 1)
        p = m.pos + (8 if version == 1 else 2)
        out = []
        for _ in range(n):
            fid = r.u(p, 2)
            p += 2
            nlen = 0
            if version == 1 or fid >= 256:
                nlen = r.u(p, 2)
                p += 2
            p += 2                              # flags
            nvals = r.u(p, 2)
            p += 2
            p += _pad8(nlen) if version == 1 else nlen
            vals = [r.u(p + 4 * k, 4) for k in range(nvals)]
            p += 4 * nvals
            if version == 1 and nvals % 2:
                p += 4
            out.append((fid, vals))
        return out

    def _chunked(self, btree, dims):
        """Read a chunked dataset through its B-tree (version 1, type 1)."""
        r = self._r
        if self._dt.vlen:
            raise NotImplementedError("HDF5: chunked variable-length dataset")
        cdims = dims[:-1]
        rank = len(cdims)
        esize = self._dt.size
        out = np.zeros(self.shape, dtype=self._dt.dtype)
        filters = self._filters()
        if btree == r.undef:
            return _decode(r, out.tobytes(), self._dt, self.shape)
        ksize = 8 + 8 * (rank + 1)

        def node(addr):
            if r.bytes(addr, 4) != b'TREE' or r.u(addr + 4, 1) != 1:
                raise ValueError("HDF5: bad chunk B-tree node")
            level = r.u(addr + 5, 1)
            used = r.u(addr + 6, 2)
            p = addr + 8 + 2 * r.so
            for _ in range(used):
                csize, mask = r.u(p, 4), r.u(p + 4, 4)
                offs = [r.u(p + 8 + 8 * k, 8) for k in range(rank)]
                child = r.off(p + ksize)
                p += ksize + r.so
                if level > 0:
                    node(child)
                else:
                    chunk(r.bytes(child, csize), mask, offs)

        def chunk(raw, mask, offs):
            for k in range(len(filters) - 1, -1, -1):
                fid, vals = filters[k]
                if mask & (1 << k):
                    continue
                if fid == 1:
                    raw = zlib.decompress(raw)
                elif fid == 2:
                    n = len(raw) // esize
                    raw = np.frombuffer(raw, np.uint8)[:n * esize].reshape(esize, n).T.tobytes()
                elif fid == 3:
                    raw = raw[:-4]
                else:
                    raise NotImplementedError("HDF5: filter %d" % fid)
            block = np.frombuffer(raw, dtype=self._dt.dtype,
                                  count=int(np.prod(cdims))).reshape(cdims)
            dst = tuple(slice(o, min(o + c, s)) for o, c, s in zip(offs, cdims, self.shape))
            src = tuple(slice(0, d.stop - d.start) for d in dst)
            out[dst] = block[src]

        node(btree)
        if self._dt.string:
            return out
        return out.astype(self._dt.dtype.newbyteorder('='))


class Group(object):
    """An HDF5 group: ``g[name]`` (paths with '/' allowed), ``keys()``,
    ``attrs``."""

    def __init__(self, r, name, msgs, cache=None):
        self._r = r
        self.name = name
        self.attrs = Attributes(r, msgs)
        links = {}
        for m in msgs:
            if m.type == _MSG_SYMBOL_TABLE:
                links.update(_symbol_table_links(r, r.off(m.pos), r.off(m.pos + r.so)))
            elif m.type == _MSG_LINK:
                lname, addr = _link(r, m.pos)
                if addr is not None:
                    links[lname] = addr
            elif m.type == _MSG_LINK_INFO:
                flags = r.u(m.pos + 1, 1)
                heap = r.off(m.pos + 2 + (8 if flags & 1 else 0))
                if heap != r.undef:
                    raise NotImplementedError("HDF5: dense link storage")
        self._links = links

    def keys(self):
        return sorted(self._links)

    def __contains__(self, name):
        try:
            self[name]
            return True
        except KeyError:
            return False

    def __iter__(self):
        return iter(self.keys())

    def __getitem__(self, path):
        parts = [p for p in path.split('/') if p]
        obj = self
        for k, part in enumerate(parts):
            if not isinstance(obj, Group) or part not in obj._links:
                raise KeyError("no object %r in the HDF5 file" % path)
            msgs = _object_header(self._r, obj._links[part])
            full = '/' + '/'.join(parts[:k + 1])
            types = set(m.type for m in msgs)
            if _MSG_LAYOUT in types and _MSG_DATATYPE in types:
                obj = Dataset(self._r, full, msgs)
            else:
                obj = Group(self._r, full, msgs)
        return obj


class File(Group):
    """An HDF5 file opened read-only (``with File(path) as f: f['x'].value``)."""

    def __init__(self, path, mode='r'):
        if mode != 'r':
            raise ValueError("HDF5 files are read-only here (mode 'r')")
        self._fh = open(path, 'rb')
        try:
            self._mm = mmap.mmap(self._fh.fileno(), 0, access=mmap.ACCESS_READ)
        except ValueError:
            self._fh.close()
            raise ValueError("%s is empty, not an HDF5 file" % path)
        base = 0
        while base < len(self._mm) and self._mm[base:base + 8] != _SIGNATURE:
            base = 512 if base == 0 else base * 2
        if base >= len(self._mm):
            self.close()
            raise ValueError("%s is not an HDF5 file" % path)
        version = self._mm[base + 8]
        if version in (0, 1):
            so, sl = self._mm[base + 13], self._mm[base + 14]
            p = base + 24 + (4 if version == 1 else 0)
            r = _Reader(self._mm, so, sl, _Reader(self._mm, so, sl).u(p, so))
            p += 4 * so                      # base, free-space, EOF, driver addresses
            root = r.off(p + so)             # root symbol table entry: header address
        elif version in (2, 3):
            so, sl = self._mm[base + 9], self._mm[base + 10]
            r = _Reader(self._mm, so, sl, _Reader(self._mm, so, sl).u(base + 12, so))
            root = r.off(base + 12 + 3 * so)
        else:
            self.close()
            raise NotImplementedError("HDF5: superblock version %d" % version)
        Group.__init__(self, r, '/', _object_header(r, root))
        self.filename = path

    def close(self):
        if getattr(self, '_mm', None) is not None:
            self._mm.close()
            self._mm = None
        if getattr(self, '_fh', None) is not None:
            self._fh.close()
            self._fh = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
        return False
