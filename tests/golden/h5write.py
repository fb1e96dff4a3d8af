"""A minimal HDF5 writer for test fixtures (test infrastructure only).

Writes what the HDF5 library's defaults write for a flat file of
datasets: superblock version 0, the root group as a symbol table (a
version-1 B-tree over symbol-table nodes of at most 8 entries, names in a
local heap), version-1 object headers, contiguous little-endian datasets of
fixed-point or IEEE floating-point type, and fixed-length string, numeric
or variable-length string attributes (``VlenStr``, as h5py stores Python
strings; the strings go to one global heap collection) (HDF Group, "HDF5
File Format Specification", version 2).
The reader it feeds, pyop2_amd.hdf5, is checked separately against files
written by the HDF5 library.
"""

import struct

import numpy as np

UNDEF = 0xFFFFFFFFFFFFFFFF
LEAF_K = 4        # group leaf node K: 2K = 8 entries per symbol-table node
INTERNAL_K = 16   # group internal node K: 2K = 32 children per B-tree node


def _pad8(b):
    return b + b'\0' * (-len(b) % 8)


class VlenStr(str):
    """An attribute value stored as a variable-length string."""


_VLEN_STR_TYPE = struct.pack('<B3sI', 0x19, bytes([0x01, 0, 0]), 16) + \
    struct.pack('<B3sI', 0x10, bytes([0, 0, 0]), 1) + struct.pack('<HH', 0, 8)


def _global_heap(strings):
    """A global heap collection holding ``strings`` as objects 1..n, at
    least 4096 bytes as the HDF5 library makes them."""
    objs = b''
    for k, t in enumerate(strings):
        data = t.encode('utf-8')
        objs += struct.pack('<HH4xQ', k + 1, 1, len(data)) + _pad8(data)
    size = max(4096, 16 + len(objs) + 16)
    free = size - 16 - len(objs)
    return b'GCOL' + struct.pack('<B3xQ', 1, size) + objs + \
        struct.pack('<HH4xQ', 0, 0, free) + b'\0' * (free - 16)


def _datatype(dtype):
    dt = np.dtype(dtype)
    if dt.kind in 'iu':
        bits = 0x08 if dt.kind == 'i' else 0x00
        return struct.pack('<B3sI', 0x10, bytes([bits, 0, 0]), dt.itemsize) + \
            struct.pack('<HH', 0, 8 * dt.itemsize)
    if dt.kind == 'f':
        # IEEE: mantissa normalisation 2 (implied leading 1), sign bit position
        if dt.itemsize == 8:
            sign, eloc, esize, msize, bias = 63, 52, 11, 52, 1023
        else:
            sign, eloc, esize, msize, bias = 31, 23, 8, 23, 127
        return struct.pack('<B3sI', 0x11, bytes([0x20, sign, 0]), dt.itemsize) + \
            struct.pack('<HHBBBBI', 0, 8 * dt.itemsize, eloc, esize, 0, msize, bias)
    if dt.kind == 'S':
        return struct.pack('<B3sI', 0x13, bytes([0, 0, 0]), dt.itemsize)
    raise TypeError("unsupported dtype %s" % dt)


def _dataspace(shape):
    return struct.pack('<BBBB4x', 1, len(shape), 0, 0) + \
        b''.join(struct.pack('<Q', d) for d in shape)


def _message(mtype, body):
    body = _pad8(body)
    return struct.pack('<HHB3x', mtype, len(body), 0) + body


def _attribute(name, value, heap=None):
    if isinstance(value, VlenStr):
        addr, index = heap[str(value)]
        nm = name.encode('utf-8') + b'\0'
        ds = _dataspace(())
        data = struct.pack('<IQI', len(value.encode('utf-8')), addr, index)
        body = struct.pack('<BBHHH', 1, 0, len(nm), len(_VLEN_STR_TYPE), len(ds)) + \
            _pad8(nm) + _pad8(_VLEN_STR_TYPE) + _pad8(ds) + data
        return _message(0x000C, body)
    if isinstance(value, str):
        raw = value.encode('utf-8') + b'\0'
        a = np.array(raw, dtype='S%d' % len(raw))
    else:
        a = np.ascontiguousarray(np.asarray(value))
        if a.dtype.kind in 'iuf':
            a = a.astype(a.dtype.newbyteorder('<'))
    nm = name.encode('utf-8') + b'\0'
    dt = _datatype(a.dtype)
    ds = _dataspace(a.shape)
    body = struct.pack('<BBHHH', 1, 0, len(nm), len(dt), len(ds)) + \
        _pad8(nm) + _pad8(dt) + _pad8(ds) + a.tobytes()
    return _message(0x000C, body)


def _object_header(messages):
    body = b''.join(messages)
    return struct.pack('<BBHII4x', 1, 0, len(messages), 1, len(body)) + body


def write(path, datasets):
    """Write ``datasets``: name -> (array, {attribute name: value})."""
    names = sorted(datasets)
    out = bytearray(b'\0' * 96)            # superblock, filled in last

    def place(b):
        pos = len(out)
        out.extend(_pad8(bytes(b)))
        return pos

    # local heap data: "" at 0, then the names
    heap = bytearray(b'\0' * 8)
    name_off = {}
    for n in names:
        name_off[n] = len(heap)
        heap.extend(_pad8(n.encode('utf-8') + b'\0'))
    # variable-length strings of the attributes: one global heap collection
    vlen = sorted(set(str(v) for n in names for v in datasets[n][1].values()
                      if isinstance(v, VlenStr)))
    heap_refs = {}
    if vlen:
        gcol = place(_global_heap(vlen))
        heap_refs = {t: (gcol, k + 1) for k, t in enumerate(vlen)}
    # datasets: raw data, then their object headers
    headers = {}
    for n in names:
        arr, attrs = datasets[n]
        a = np.ascontiguousarray(arr)
        a = a.astype(a.dtype.newbyteorder('<'))
        data_pos = place(a.tobytes()) if a.nbytes else UNDEF
        msgs = [_message(0x0001, _dataspace(a.shape)),
                _message(0x0003, _datatype(a.dtype)),
                _message(0x0008, struct.pack('<BBQQ', 3, 1, data_pos, a.nbytes))]
        msgs += [_attribute(k, v, heap_refs) for k, v in sorted(attrs.items())]
        headers[n] = place(_object_header(msgs))
    # symbol-table nodes of at most 2 * LEAF_K entries
    groups = [names[k:k + 2 * LEAF_K] for k in range(0, len(names), 2 * LEAF_K)]
    if len(groups) > 2 * INTERNAL_K:
        raise ValueError("too many datasets for a one-level group B-tree")
    snods = []
    for g in groups:
        ents = b''.join(struct.pack('<QQII16x', name_off[n], headers[n], 0, 0) for n in g)
        ents += b'\0' * (40 * (2 * LEAF_K - len(g)))
        snods.append(place(b'SNOD' + struct.pack('<BBH', 1, 0, len(g)) + ents))
    heap_data = place(heap)
    heap_hdr = place(b'HEAP' + struct.pack('<B3xQQQ', 0, len(heap), UNDEF, heap_data))
    # the group B-tree (level 0): key 0 is "", key i the last name of node i-1
    keys = [0] + [name_off[g[-1]] for g in groups]
    tree = b'TREE' + struct.pack('<BBHQQ', 0, 0, len(groups), UNDEF, UNDEF)
    for k, s in enumerate(snods):
        tree += struct.pack('<QQ', keys[k], s)
    tree += struct.pack('<Q', keys[-1])
    tree += b'\0' * (16 * (2 * INTERNAL_K - len(groups)))
    btree = place(tree)
    root = place(_object_header([_message(0x0011, struct.pack('<QQ', btree, heap_hdr))]))
    sb = b'\x89HDF\r\n\x1a\n' + struct.pack('<BBBBBBBBHHI', 0, 0, 0, 0, 0, 8, 8, 0,
                                            LEAF_K, INTERNAL_K, 0)
    sb += struct.pack('<QQQQ', 0, UNDEF, len(out), UNDEF)
    sb += struct.pack('<QQII', 0, root, 1, 0) + struct.pack('<QQ', btree, heap_hdr)
    assert len(sb) == 96
    out[0:96] = sb
    with open(path, 'wb') as f:
        f.write(bytes(out))
