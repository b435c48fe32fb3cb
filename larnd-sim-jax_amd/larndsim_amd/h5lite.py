"""A small read-only HDF5 reader for the reference's segment files (numpy only).

The reference reads its inputs with h5py (``f['segments'][:]``, optimize/dataio.py:113-114), which
is not installed here.  This module reads what those files use, following the HDF5 file-format
specification (version 0 superblock, symbol-table groups with version 1 B-trees and local heaps,
version 1 object headers with continuation blocks, version 1/2 dataspaces, fixed-point,
floating-point and compound datatypes, contiguous and compact data layouts) and returns a numpy
structured array.  Anything else (chunked or filtered datasets, other superblock versions, variable-length
types) raises NotImplementedError instead of guessing.
"""
from __future__ import annotations

import struct

import numpy as np

_SIG = b"\x89HDF\r\n\x1a\n"
_UNDEF = 0xFFFFFFFFFFFFFFFF


class _File:
    def __init__(self, data: bytes):
        self.d = data
        if data[:8] != _SIG:
            raise ValueError("not an HDF5 file")
        ver = data[8]
        if ver not in (0, 1):
            raise NotImplementedError(f"HDF5 superblock version {ver}")
        self.so, self.sl = data[13], data[14]
        if (self.so, self.sl) != (8, 8):
            raise NotImplementedError("only 8-byte offsets and lengths")
        p = 24 if ver == 0 else 28
        self.base = self.u("Q", p)
        root = p + 32  # base, free-space, end-of-file, driver-info addresses
        self.root_header = self.u("Q", root + 8)
        cache = self.u("I", root + 16)
        self.root_btree, self.root_heap = (self.u("Q", root + 24), self.u("Q", root + 32)) if cache == 1 else (None, None)

    def u(self, fmt, off):
        v = struct.unpack_from("<" + fmt, self.d, off)
        return v[0] if len(v) == 1 else v

    # ---- object headers (version 1) ----
    def messages(self, addr):
        """[(type, flags, payload offset, size)] of an object header, following continuations."""
        ver = self.d[addr]
        if ver != 1:
            raise NotImplementedError(f"object header version {ver}")
        nmsg = self.u("H", addr + 2)
        size = self.u("I", addr + 8)
        blocks = [(addr + 16, size)]
        out = []
        while blocks and len(out) < nmsg:
            p, n = blocks.pop(0)
            end = p + n
            while p + 8 <= end and len(out) < nmsg:
                mt, ms = self.u("HH", p)
                fl = self.d[p + 4]
                out.append((mt, fl, p + 8, ms))
                if mt == 0x10:  # continuation: offset, length
                    blocks.append(self.u("QQ", p + 8))
                p += 8 + ms
        return out

    # ---- groups (symbol tables) ----
    def heap_name(self, heap, off):
        if self.d[heap:heap + 4] != b"HEAP":
            raise ValueError("bad local heap")
        data = self.u("Q", heap + 24)
        s = data + off
        e = self.d.index(b"\0", s)
        return self.d[s:e].decode()

    def group_entries(self, btree, heap):
        """{name: object header address} of a symbol-table group."""
        out = {}
        stack = [btree]
        while stack:
            node = stack.pop()
            if self.d[node:node + 4] != b"TREE":
                raise ValueError("bad B-tree node")
            ntype, level, used = self.d[node + 4], self.d[node + 5], self.u("H", node + 6)
            if ntype != 0:
                raise ValueError("not a group B-tree")
            p = node + 24  # signature, type, level, entries, left and right siblings
            p += 8  # key 0
            for _ in range(used):
                child = self.u("Q", p)
                p += 16  # child, next key
                if level > 0:
                    stack.append(child)
                    continue
                if self.d[child:child + 4] != b"SNOD":
                    raise ValueError("bad symbol-table node")
                n = self.u("H", child + 6)
                q = child + 8
                for _ in range(n):
                    name_off, hdr = self.u("QQ", q)
                    out[self.heap_name(heap, name_off)] = hdr
                    q += 40
        return out

    def find(self, path):
        """Object header address of a slash-separated path from the root group."""
        addr = self.root_header
        for part in [p for p in path.split("/") if p]:
            btree, heap = None, None
            for mt, _, q, _ in self.messages(addr):
                if mt == 0x11:  # symbol table message: B-tree, local heap
                    btree, heap = self.u("QQ", q)
            if btree is None:
                if addr == self.root_header and self.root_btree is not None:
                    btree, heap = self.root_btree, self.root_heap
                else:
                    raise KeyError(part)
            entries = self.group_entries(btree, heap)
            if part not in entries:
                raise KeyError(part)
            addr = entries[part]
        return addr

    # ---- datatypes ----
    def dtype(self, q):
        """(numpy dtype, bytes consumed) of the datatype message at q."""
        cls, ver = self.d[q] & 0x0F, self.d[q] >> 4
        bits = self.d[q + 1] | (self.d[q + 2] << 8) | (self.d[q + 3] << 16)
        size = self.u("I", q + 4)
        p = q + 8
        order = ">" if bits & 1 else "<"
        if cls == 0:  # fixed point: byte order, sign; offset, precision
            signed = bool(bits & 8)
            return np.dtype(f"{order}{'i' if signed else 'u'}{size}"), 8 + 4
        if cls == 1:  # floating point: properties of 12 bytes
            return np.dtype(f"{order}f{size}"), 8 + 12
        if cls == 6:  # compound
            nmem = bits & 0xFFFF
            names, formats, offsets = [], [], []
            for _ in range(nmem):
                e = self.d.index(b"\0", p)
                name = self.d[p:e].decode()
                if ver < 3:
                    p += ((e - p) // 8 + 1) * 8  # name padded to a multiple of 8, with its terminator
                else:
                    p = e + 1
                if ver == 1:
                    off = self.u("I", p)
                    ndim = self.d[p + 4]
                    if ndim:
                        raise NotImplementedError("array members of a version 1 compound")
                    p += 4 + 4 + 4 + 4 + 16
                elif ver == 2:
                    off = self.u("I", p)
                    p += 4
                else:
                    nb = max(1, (size.bit_length() + 7) // 8)
                    off = int.from_bytes(self.d[p:p + nb], "little")
                    p += nb
                dt, used = self.dtype(p)
                p += used
                names.append(name)
                formats.append(dt)
                offsets.append(off)
            return np.dtype(dict(names=names, formats=formats, offsets=offsets, itemsize=size)), p - q
        raise NotImplementedError(f"HDF5 datatype class {cls}")

    def dataset(self, addr):
        shape, dt, data = None, None, None
        for mt, _, q, ms in self.messages(addr):
            if mt == 0x1:  # dataspace
                ver, ndim = self.d[q], self.d[q + 1]
                p = q + (8 if ver == 1 else 4)
                shape = tuple(self.u("Q", p + 8 * i) for i in range(ndim))
            elif mt == 0x3:
                dt, _ = self.dtype(q)
            elif mt == 0x8:  # data layout
                ver = self.d[q]
                if ver != 3:
                    raise NotImplementedError(f"data layout version {ver}")
                lc = self.d[q + 1]
                if lc == 1:  # contiguous: address, size
                    a, n = self.u("QQ", q + 2)
                    data = (a, n)
                elif lc == 0:  # compact: size, raw data
                    n = self.u("H", q + 2)
                    data = (q + 4, n)
                else:
                    raise NotImplementedError("chunked HDF5 datasets")
            elif mt == 0xB:
                raise NotImplementedError("filtered HDF5 datasets")
        if shape is None or dt is None or data is None:
            raise ValueError("incomplete dataset header")
        count = int(np.prod(shape)) if shape else 1
        a, n = data
        if a == _UNDEF:
            return np.zeros(shape, dt)
        if n < count * dt.itemsize:
            raise ValueError("dataset storage smaller than its shape")
        return np.frombuffer(self.d, dtype=dt, count=count, offset=self.base + a).reshape(shape).copy()


def read_dataset(filename, path):
    """The dataset at `path` (e.g. "segments") of an HDF5 file as a numpy array (structured for a
    compound type), like h5py's ``f[path][:]``."""
    with open(filename, "rb") as fh:
        f = _File(fh.read())
    return f.dataset(f.find(path))
