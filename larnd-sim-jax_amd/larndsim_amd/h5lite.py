"""A small HDF5 reader and writer for the reference's files (numpy only).

The reference reads its inputs with h5py (``f['segments'][:]``, optimize/dataio.py:113-114) and
writes its hits as per-event groups (optimize/simulate.py:135-155); h5py is not installed here.
This module follows the HDF5 file-format specification for what those files use: the version 0
superblock, symbol-table groups (version 1 B-trees, symbol-table nodes, local heaps), version 1
object headers with continuation blocks, version 1/2 dataspaces, fixed-point, floating-point and
compound datatypes, and contiguous and compact layouts.  Reading returns numpy arrays (structured
for a compound type); anything else (chunked or filtered datasets, other superblock versions,
variable-length types) raises NotImplementedError instead of guessing.  Writing produces the
same structures as the h5py files the reference ships (superblock version 0 with group node
sizes K = 4 and 16, one symbol-table message per group, datasets with dataspace, datatype, fill
value and contiguous layout messages) for groups of numeric datasets.
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


def _tree(f, addr):
    out = {}
    btree = heap = None
    for mt, _, q, _ in f.messages(addr):
        if mt == 0x11:
            btree, heap = f.u("QQ", q)
    if btree is None and addr == f.root_header:
        btree, heap = f.root_btree, f.root_heap
    for name, hdr in sorted(f.group_entries(btree, heap).items()):
        is_group = any(mt == 0x11 for mt, _, _, _ in f.messages(hdr))
        out[name] = _tree(f, hdr) if is_group else f.dataset(hdr)
    return out


def read_tree(filename):
    """Every group and dataset of an HDF5 file as nested dicts {name: dict | numpy array}."""
    with open(filename, "rb") as fh:
        f = _File(fh.read())
    return _tree(f, f.root_header)


# ---- writing ----

_LEAF_K, _NODE_K = 4, 16  # symbol-table node and group B-tree node sizes (the library's defaults)
_SNOD_CAP, _BTREE_CAP = 2 * _LEAF_K, 2 * _NODE_K
_FREE_NULL = 1  # end of a local heap's free list


def _pad8(n):
    return (n + 7) & ~7


def _dtype_message(dt):
    """Datatype message (version 1) of a little-endian numeric numpy dtype."""
    dt = np.dtype(dt)
    if dt.byteorder == ">" or (dt.byteorder == "=" and np.little_endian is False):
        raise NotImplementedError("big-endian data")
    if dt.kind in "iu":
        bits = 0x08 if dt.kind == "i" else 0x00
        return struct.pack("<BBBBIHH", 0x10, bits, 0, 0, dt.itemsize, 0, 8 * dt.itemsize)
    if dt.kind == "f" and dt.itemsize in (4, 8):
        if dt.itemsize == 4:
            sign, exp_loc, exp_size, man_size, bias = 31, 23, 8, 23, 127
        else:
            sign, exp_loc, exp_size, man_size, bias = 63, 52, 11, 52, 1023
        return struct.pack("<BBBBIHHBBBBI", 0x11, 0x20, sign, 0, dt.itemsize, 0, 8 * dt.itemsize, exp_loc, exp_size,
                           0, man_size, bias)
    raise NotImplementedError(f"HDF5 writing of dtype {dt}")


class _Writer:
    def __init__(self):
        self.buf = bytearray(96)  # the superblock, written last

    def alloc(self, n):
        a = _pad8(len(self.buf))
        self.buf.extend(b"\0" * (a + n - len(self.buf)))
        return a

    def put(self, addr, data):
        self.buf[addr:addr + len(data)] = data

    def header(self, messages):
        """Version 1 object header holding [(type, flags, payload)]; returns its address."""
        body = b""
        for mt, fl, payload in messages:
            payload = payload + b"\0" * (_pad8(len(payload)) - len(payload))
            body += struct.pack("<HHB3x", mt, len(payload), fl) + payload
        a = self.alloc(16 + len(body))
        self.put(a, struct.pack("<BBHII4x", 1, 0, len(messages), 1, len(body)) + body)
        return a

    def dataset(self, arr):
        arr = np.ascontiguousarray(arr)
        if arr.dtype.byteorder == ">":
            arr = arr.astype(arr.dtype.newbyteorder("<"))
        shape = arr.shape if arr.ndim else (1,)
        space = struct.pack("<BBBB4x", 1, len(shape), 1, 0) + struct.pack(f"<{len(shape)}Q", *shape) * 2
        nbytes = arr.size * arr.dtype.itemsize
        data = self.alloc(nbytes) if nbytes else _UNDEF
        if nbytes:
            self.put(data, arr.tobytes())
        layout = struct.pack("<BBQQ", 3, 1, data, nbytes)
        fill = struct.pack("<BBBBI", 2, 2, 2, 1, 0)
        return self.header([(0x1, 0, space), (0x3, 1, _dtype_message(arr.dtype)), (0x5, 1, fill), (0x8, 0, layout)])

    def group(self, tree):
        """Objects of a group first, then its local heap, symbol-table nodes and B-tree; returns
        (object header address, B-tree address, local heap address)."""
        names = sorted(tree, key=lambda n: n.encode())
        objs = {}
        for n in names:
            v = tree[n]
            objs[n] = self.group(v) if isinstance(v, dict) else (self.dataset(np.asarray(v)), None, None)
        # local heap: "" at offset 0, the names, then one free block (as the library leaves it)
        heap_data = bytearray(8)
        offs = {}
        for n in names:
            b = n.encode()
            if b"\0" in b or b"/" in b or not b:
                raise ValueError(f"bad HDF5 link name {n!r}")
            offs[n] = len(heap_data)
            heap_data += b + b"\0" * (_pad8(len(b) + 1) - len(b))
        free = len(heap_data)
        heap_data += struct.pack("<QQ", _FREE_NULL, 16)
        heap = self.alloc(32)
        data = self.alloc(len(heap_data))
        self.put(data, bytes(heap_data))
        self.put(heap, b"HEAP" + struct.pack("<B3xQQQ", 0, len(heap_data), free, data))
        # symbol-table nodes of at most 2 K_leaf entries, in name order
        leaves = []
        for i in range(0, len(names), _SNOD_CAP):
            chunk = names[i:i + _SNOD_CAP]
            a = self.alloc(8 + _SNOD_CAP * 40)
            ent = b""
            for n in chunk:
                hdr, bt, hp = objs[n]
                if bt is None:
                    ent += struct.pack("<QQI4x16x", offs[n], hdr, 0)
                else:
                    ent += struct.pack("<QQI4xQQ", offs[n], hdr, 1, bt, hp)
            self.put(a, b"SNOD" + struct.pack("<BxH", 1, len(chunk)) + ent)
            leaves.append((a, offs[chunk[-1]]))
        # B-tree levels of at most 2 K_node children; key i + 1 is the last name under child i
        level, nodes = 0, leaves
        while True:
            parents = []
            for i in range(0, max(len(nodes), 1), _BTREE_CAP):
                chunk = nodes[i:i + _BTREE_CAP]
                a = self.alloc(24 + (2 * _BTREE_CAP + 1) * 8)
                body = struct.pack("<Q", 0)
                for child, last in chunk:
                    body += struct.pack("<QQ", child, last)
                self.put(a, b"TREE" + struct.pack("<BBHQQ", 0, level, len(chunk), _UNDEF, _UNDEF) + body)
                parents.append((a, chunk[-1][1] if chunk else 0))
            for (a, _), (b, _) in zip(parents, parents[1:]):  # sibling links of the level
                self.put(a + 16, struct.pack("<Q", b))
                self.put(b + 8, struct.pack("<Q", a))
            if len(parents) == 1:
                btree = parents[0][0]
                break
            level, nodes = level + 1, parents
        hdr = self.header([(0x11, 0, struct.pack("<QQ", btree, heap))])
        return hdr, btree, heap

    def finish(self, root):
        hdr, btree, heap = root
        eof = _pad8(len(self.buf))
        self.buf.extend(b"\0" * (eof - len(self.buf)))
        sb = _SIG + struct.pack("<8B", 0, 0, 0, 0, 0, 8, 8, 0) + struct.pack("<HHI", _LEAF_K, _NODE_K, 0)
        sb += struct.pack("<QQQQ", 0, _UNDEF, eof, _UNDEF)
        sb += struct.pack("<QQI4xQQ", 0, hdr, 1, btree, heap)
        self.put(0, sb)
        return bytes(self.buf)


def write_tree(filename, tree):
    """Write nested dicts {name: dict (a group) | numeric numpy array (a dataset)} as an HDF5 file
    (replacing it); read_tree(filename) gives the same tree back."""
    w = _Writer()
    data = w.finish(w.group(tree))
    tmp = str(filename) + ".tmp"
    with open(tmp, "wb") as fh:
        fh.write(data)
    import os

    os.replace(tmp, filename)


def read_dataset(filename, path):
    """The dataset at `path` (e.g. "segments") of an HDF5 file as a numpy array (structured for a
    compound type), like h5py's ``f[path][:]``."""
    with open(filename, "rb") as fh:
        f = _File(fh.read())
    return f.dataset(f.find(path))
