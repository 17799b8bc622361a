"""A writer of JLD2 files (the format of Oceananigans' JLD2Writer) for
numeric arrays and scalars, in plain Python.

A JLD2 file is an HDF5 file behind a 512-byte text header ("HDF5-based
Julia Data Format, version ..."): the HDF5 superblock (version 2) sits at
byte 512 and is the base address that every file address is relative to.
Groups are new-style compact groups (link-info, group-info and hard-link
messages in a version-2 object header); datasets have version-2 object
headers with a dataspace, an IEEE little-endian float or a signed integer
datatype, a fill-value message and a contiguous layout.  A Julia array is
column-major, so an array of Julia size (n1, n2, ..., nd) is stored with its
bytes in that order (first index fastest) and the HDF5 dataspace lists the
dimensions reversed, (nd, ..., n1), which is how JLD2 itself writes it.

The file is written incrementally: dataset payloads and dataset headers are
appended once; a group whose links change gets a new header appended at the
next flush and the superblock (rewritten in place) points at the new root.
Superseded group headers stay in the file as unreferenced bytes.

Checksums are Bob Jenkins' lookup3 hashlittle (initial value 0), which
HDF5 uses for version-2 superblocks and object headers.
"""
import os
import struct

import numpy as np

JLD2_HEADER = b"HDF5-based Julia Data Format, version 0.1.1"
HEADER_LENGTH = 512
UNDEF = 0xFFFFFFFFFFFFFFFF
SIGNATURE = b"\x89HDF\r\n\x1a\n"


def lookup3(data, initval=0):
    """Bob Jenkins' lookup3 hashlittle of a byte string (the HDF5 checksum)."""
    M = 0xFFFFFFFF

    def rot(x, k):
        return ((x << k) | (x >> (32 - k))) & M

    n = len(data)
    a = b = c = (0xDEADBEEF + n + initval) & M
    i = 0
    while n > 12:
        a = (a + int.from_bytes(data[i:i + 4], "little")) & M
        b = (b + int.from_bytes(data[i + 4:i + 8], "little")) & M
        c = (c + int.from_bytes(data[i + 8:i + 12], "little")) & M
        a = (a - c) & M; a ^= rot(c, 4); c = (c + b) & M
        b = (b - a) & M; b ^= rot(a, 6); a = (a + c) & M
        c = (c - b) & M; c ^= rot(b, 8); b = (b + a) & M
        a = (a - c) & M; a ^= rot(c, 16); c = (c + b) & M
        b = (b - a) & M; b ^= rot(a, 19); a = (a + c) & M
        c = (c - b) & M; c ^= rot(b, 4); b = (b + a) & M
        n -= 12
        i += 12
    if n == 0:
        return c
    tail = data[i:i + n] + bytes(12 - n)
    a = (a + int.from_bytes(tail[0:4], "little")) & M
    b = (b + int.from_bytes(tail[4:8], "little")) & M
    c = (c + int.from_bytes(tail[8:12], "little")) & M
    c ^= b; c = (c - rot(b, 14)) & M
    a ^= c; a = (a - rot(c, 11)) & M
    b ^= a; b = (b - rot(a, 25)) & M
    c ^= b; c = (c - rot(b, 16)) & M
    a ^= c; a = (a - rot(c, 4)) & M
    b ^= a; b = (b - rot(a, 14)) & M
    c ^= b; c = (c - rot(b, 24)) & M
    return c


def _message(mtype, body, flags=0):
    return struct.pack("<BHB", mtype, len(body), flags) + body


def _object_header(messages):
    """Version-2 object header: OHDR, version 2, flags (4-byte chunk size,
    no times, no attribute phase change), chunk size, messages, checksum."""
    body = b"".join(messages)
    head = b"OHDR" + struct.pack("<BBI", 2, 2, len(body)) + body
    return head + struct.pack("<I", lookup3(head))


def _dataspace(dims):
    if len(dims) == 0:
        return _message(0x01, struct.pack("<BBBB", 2, 0, 0, 0))
    return _message(0x01, struct.pack("<BBBB", 2, len(dims), 0, 1) + b"".join(struct.pack("<Q", d) for d in dims))


def _datatype(dtype):
    dt = np.dtype(dtype)
    if dt == np.float64:
        props = struct.pack("<HHBBBBI", 0, 64, 52, 11, 0, 52, 1023)
        return _message(0x03, struct.pack("<B3BI", 0x11, 0x20, 63, 0, 8) + props, flags=1)
    if dt == np.float32:
        props = struct.pack("<HHBBBBI", 0, 32, 23, 8, 0, 23, 127)
        return _message(0x03, struct.pack("<B3BI", 0x11, 0x20, 31, 0, 4) + props, flags=1)
    if dt.kind == "i":
        bits = 8 * dt.itemsize
        return _message(0x03, struct.pack("<B3BI", 0x10, 0x08, 0, 0, dt.itemsize) + struct.pack("<HH", 0, bits),
                        flags=1)
    raise TypeError(f"unsupported element type {dt}")


def _fill_value():
    # version 3, late space allocation, fill written only if set, none defined
    return _message(0x05, struct.pack("<BB", 3, 2 | (2 << 2)))


def _layout(address, size):
    return _message(0x08, struct.pack("<BBQQ", 3, 1, address, size))


def _group_header(links):
    msgs = [_message(0x02, struct.pack("<BBQQ", 0, 0, UNDEF, UNDEF)),  # link info: compact storage
            _message(0x0A, struct.pack("<BB", 0, 0))]                  # group info
    for name, address in links.items():
        n = name.encode("utf-8")
        if not 0 < len(n) < 256:
            raise ValueError(f"link name {name!r} must have 1 to 255 bytes")
        flags = 0x10  # a character-set field follows: UTF-8
        msgs.append(_message(0x06, struct.pack("<BBB", 1, flags, 1) + struct.pack("<B", len(n)) + n +
                             struct.pack("<Q", address)))
    return _object_header(msgs)


class _Group:
    def __init__(self):
        self.children = {}   # name -> _Group or dataset header address (relative)
        self.address = None  # relative address of its current header
        self.dirty = True


class JLD2File:
    """An open JLD2 file for writing.  `f[path] = array` stores a numeric
    array (Julia dimension order, see the module docstring) or a scalar at a
    '/'-separated path, creating the groups on the way; `flush()` makes the
    file complete on disk; `close()` flushes and closes."""

    def __init__(self, filename, overwrite_existing=True):
        if os.path.exists(filename) and not overwrite_existing:
            raise FileExistsError(f"{filename} exists and overwrite_existing is false")
        self.filename = filename
        self._f = open(filename, "w+b")
        self._f.write(JLD2_HEADER + bytes(HEADER_LENGTH - len(JLD2_HEADER)))
        self._f.write(bytes(48))  # superblock, written by flush
        self._root = _Group()
        self.flush()

    # relative file addresses (the superblock is the base address)
    def _append(self, data):
        self._f.seek(0, os.SEEK_END)
        address = self._f.tell() - HEADER_LENGTH
        self._f.write(data)
        return address

    def __setitem__(self, path, value):
        parts = [p for p in path.split("/") if p]
        if not parts:
            raise ValueError("empty dataset path")
        g = self._root
        for p in parts[:-1]:
            child = g.children.get(p)
            if child is None:
                child = g.children[p] = _Group()
                g.dirty = True
            elif not isinstance(child, _Group):
                raise ValueError(f"{p} in {path} is a dataset, not a group")
            g = child
        if parts[-1] in g.children:
            raise ValueError(f"{path} exists")
        arr = np.asarray(value)
        if arr.dtype.kind == "b":
            arr = arr.astype(np.int8)
        if arr.dtype.kind not in "fi":
            raise TypeError(f"{path}: only numeric arrays and scalars are written")
        arr = arr.astype(arr.dtype.newbyteorder("<"), copy=False)
        raw = np.asfortranarray(arr).tobytes(order="F")
        data_address = self._append(raw) if raw else UNDEF
        dims = tuple(reversed(arr.shape))
        header = _object_header([_dataspace(dims), _datatype(arr.dtype), _fill_value(),
                                 _layout(data_address, len(raw))])
        g.children[parts[-1]] = self._append(header)
        g.dirty = True

    def _write_group(self, g):
        for child in g.children.values():
            if isinstance(child, _Group):
                if self._write_group(child):
                    g.dirty = True
        if not g.dirty:
            return False
        links = {name: (c.address if isinstance(c, _Group) else c) for name, c in g.children.items()}
        g.address = self._append(_group_header(links))
        g.dirty = False
        return True

    def flush(self):
        self._write_group(self._root)
        self._f.seek(0, os.SEEK_END)
        eof = self._f.tell()  # the end-of-file address is absolute; the others are relative to the base
        sb = SIGNATURE + struct.pack("<BBBB", 2, 8, 8, 0) + struct.pack("<QQQQ", HEADER_LENGTH, UNDEF, eof,
                                                                        self._root.address)
        self._f.seek(HEADER_LENGTH)
        self._f.write(sb + struct.pack("<I", lookup3(sb)))
        self._f.flush()

    def close(self):
        if self._f is not None:
            self.flush()
            self._f.close()
            self._f = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
