"""A small HDF5 file writer and reader, written from the published HDF5 file
format specification (version 0 superblock, symbol-table groups, contiguous
datasets), for the GridDataset files of IO/gridsIO/write_grids.jl.

The reference writes those files through HDF5.jl (libhdf5); neither libhdf5 nor
h5py is available to this package, so it carries its own encoder of the subset
such a file needs:

  * superblock version 0 (8-byte offsets and lengths), root group entry with
    its B-tree and local heap addresses cached;
  * groups: version 1 object header holding a symbol table message, a local
    heap of the member names, one version 1 B-tree leaf node (group node type)
    and one symbol table node whose entries are sorted by name;
  * datasets: version 1 object header with a version 1 dataspace (rank 0 for a
    scalar), a datatype, a version 2 fill value and a version 3 contiguous
    layout message; data stored little-endian, in the array's memory order
    (a Fortran-ordered array of shape (nx, ny, nz) is written with the
    dataspace (nz, ny, nx), as HDF5.jl writes a Julia array);
  * datatypes: IEEE float32 / float64, signed and unsigned 8/16/32/64-bit
    integers, fixed-length UTF-8 strings (null padded) and booleans as an
    8-bit enumeration {FALSE = 0, TRUE = 1}.

`read` parses exactly that subset back (and rejects anything else), so files
written here round-trip; version 1 object headers and version 0 superblocks
carry no checksums.
"""
import struct

import numpy as np

UNDEF = 0xFFFFFFFFFFFFFFFF
SIGNATURE = b"\x89HDF\r\n\x1a\n"
_LEAF_K_MIN = 4       # group leaf node K: a symbol table node holds 2K entries
_INTERNAL_K = 16      # group internal node K: a B-tree node has room for 2K children
_ENTRY = 40           # symbol table entry size with 8-byte offsets
# The end of a local heap's free list on disk.  The specification names the
# undefined address; the HDF5 library writes 1 (no block starts at an odd
# offset) and rejects other offsets past the data segment, so a heap without
# free blocks is written with 1 and either value is read.
_FREE_NULL = 1


class HDF5FormatError(ValueError):
    pass


class Group:
    """An HDF5 group: an ordered mapping of names to groups and datasets."""

    def __init__(self):
        self.members = {}

    def create_group(self, name):
        _check_name(name, self)
        g = Group()
        self.members[name] = g
        return g

    def write(self, name, value, column_major=False):
        """Creates dataset `name` holding `value` (an array, a number, a bool or
        a string, or a sequence of strings).  column_major: store an array as a
        Julia array of that shape (dataspace dimensions reversed, first index
        fastest), whatever its memory order."""
        _check_name(name, self)
        if column_major and np.ndim(value) > 1:
            value = np.asfortranarray(value)
            d = _dataset(value)
            if not _fortran(value):  # also C-contiguous (a singleton dimension): force the layout
                le = value.dtype.newbyteorder("<")
                d = _Dataset(le.str, value.astype(le).tobytes(order="F"), tuple(reversed(value.shape)))
            self.members[name] = d
            return
        self.members[name] = _dataset(value)

    def __getitem__(self, name):
        return self.members[name]

    def __contains__(self, name):
        return name in self.members

    def keys(self):
        return list(self.members.keys())


def _check_name(name, g):
    if not isinstance(name, str) or not name or "/" in name or "\0" in name:
        raise ValueError("invalid HDF5 link name %r" % (name,))
    if name in g.members:
        raise ValueError("link %r already exists" % (name,))


class _Dataset:
    def __init__(self, kind, data, shape, strlen=0):
        self.kind = kind      # numpy dtype string, "str" or "bool"
        self.data = data      # raw bytes in file order
        self.shape = shape    # dataspace dims (C order, slowest first); () for a scalar
        self.strlen = strlen


def _dataset(value):
    if isinstance(value, (bool, np.bool_)):
        return _Dataset("bool", bytes([1 if value else 0]), ())
    if isinstance(value, str):
        raw = value.encode("utf-8")
        n = max(len(raw), 1)
        return _Dataset("str", raw.ljust(n, b"\0"), (), n)
    if isinstance(value, (list, tuple)) and value and all(isinstance(v, str) for v in value):
        raws = [v.encode("utf-8") for v in value]
        n = max(1, max(len(r) for r in raws))
        return _Dataset("str", b"".join(r.ljust(n, b"\0") for r in raws), (len(raws),), n)
    a = np.asarray(value)
    if a.dtype == np.bool_:
        u = a.astype(np.uint8, order="A")
        return _Dataset("bool", u.tobytes(order="F" if _fortran(u) else "C"), _dims(u))
    if a.dtype.kind not in "fiu" or a.dtype.itemsize not in (1, 2, 4, 8) or (a.dtype.kind == "f" and
                                                                                a.dtype.itemsize < 4):
        raise TypeError("unsupported dataset type %s" % a.dtype)
    le = a.dtype.newbyteorder("<")
    if a.ndim == 0:
        return _Dataset(le.str, a.astype(le).tobytes(), ())
    if _fortran(a):
        # column-major memory: the dataspace lists the dimensions reversed
        return _Dataset(le.str, np.asfortranarray(a, dtype=le).tobytes(order="F"), tuple(reversed(a.shape)))
    return _Dataset(le.str, np.ascontiguousarray(a, dtype=le).tobytes(order="C"), tuple(a.shape))


def _fortran(a):
    return a.ndim > 1 and a.flags.f_contiguous and not a.flags.c_contiguous


def _dims(a):
    if a.ndim == 0:
        return ()
    return tuple(reversed(a.shape)) if _fortran(a) else tuple(a.shape)


# ---- encoding ----

def _pad8(b):
    return b + b"\0" * (-len(b) % 8)


def _datatype_message(d):
    if d.kind == "str":  # class 3, version 1: null pad (1), UTF-8 (1 << 4)
        return struct.pack("<B3BI", 0x13, 0x11, 0, 0, d.strlen)
    if d.kind == "bool":  # class 8 enumeration of two members over int8
        base = struct.pack("<B3BIHH", 0x10, 0x08, 0, 0, 1, 0, 8)
        names = _pad8(b"FALSE\0") + _pad8(b"TRUE\0")
        return struct.pack("<B3BI", 0x18, 2, 0, 0, 1) + base + names + bytes([0, 1])
    dt = np.dtype(d.kind)
    size = dt.itemsize
    if dt.kind == "f":  # class 1, version 1: little-endian, implied mantissa msb
        if size == 8:
            return struct.pack("<B3BIHHBBBBI", 0x11, 0x20, 63, 0, 8, 0, 64, 52, 11, 0, 52, 1023)
        return struct.pack("<B3BIHHBBBBI", 0x11, 0x20, 31, 0, 4, 0, 32, 23, 8, 0, 23, 127)
    signed = 0x08 if dt.kind == "i" else 0x00  # class 0, version 1
    return struct.pack("<B3BIHH", 0x10, signed, 0, 0, size, 0, 8 * size)


def _message(mtype, body):
    body = _pad8(body)
    return struct.pack("<HHB3x", mtype, len(body), 0) + body


def _object_header(messages):
    blob = b"".join(messages)
    return struct.pack("<BBHII4x", 1, 0, len(messages), 1, len(blob)) + blob


class _Writer:
    def __init__(self, leaf_k):
        self.buf = bytearray(96)  # the superblock, filled in last
        self.leaf_k = leaf_k

    def append(self, data):
        self.buf.extend(b"\0" * (-len(self.buf) % 8))
        addr = len(self.buf)
        self.buf.extend(data)
        return addr

    def dataset(self, d):
        if d.data:
            addr, size = self.append(d.data), len(d.data)
        else:
            addr, size = UNDEF, 0
        space = struct.pack("<BBBB4x", 1, len(d.shape), 0, 0) + b"".join(struct.pack("<Q", n) for n in d.shape)
        fill = struct.pack("<BBBB", 2, 2, 2, 0)  # version 2, late allocation, fill if set, no fill value
        layout = struct.pack("<BBQQ", 3, 1, addr, size)  # version 3, contiguous
        return self.append(_object_header([_message(0x0001, space), _message(0x0003, _datatype_message(d)),
                                           _message(0x0005, fill), _message(0x0008, layout)]))

    def group(self, g):
        """Writes g's members, then g; returns (object header, B-tree, heap) addresses."""
        entries = []
        for name in sorted(g.members, key=lambda s: s.encode("utf-8")):
            m = g.members[name]
            if isinstance(m, Group):
                oh, bt, hp = self.group(m)
                entries.append((name, oh, 1, bt, hp))
            else:
                entries.append((name, self.dataset(m), 0, 0, 0))
        if len(entries) > 2 * self.leaf_k:
            raise HDF5FormatError("group of %d members exceeds one symbol table node" % len(entries))
        # local heap: "" at offset 0, then the names, each null-terminated and 8-byte padded
        heap, offsets = bytearray(8), []
        for e in entries:
            offsets.append(len(heap))
            heap.extend(_pad8(e[0].encode("utf-8") + b"\0"))
        heap_data = self.append(bytes(heap))
        heap_addr = self.append(b"HEAP" + struct.pack("<B3xQQQ", 0, len(heap), _FREE_NULL, heap_data))
        # symbol table node with room for 2K entries
        snod = bytearray(b"SNOD" + struct.pack("<BBH", 1, 0, len(entries)) + b"\0" * (_ENTRY * 2 * self.leaf_k))
        for k, (e, off) in enumerate(zip(entries, offsets)):
            _, oh, cache, bt, hp = e
            struct.pack_into("<QQIIQQ", snod, 8 + _ENTRY * k, off, oh, cache, 0, bt, hp)
        snod_addr = self.append(bytes(snod))
        # B-tree leaf node (type 0: group): key 0 = "", child 0, key 1 = the largest name
        tree = bytearray(b"TREE" + struct.pack("<BBHQQ", 0, 0, 1, UNDEF, UNDEF) +
                         b"\0" * (8 * (2 * _INTERNAL_K + 1) + 8 * 2 * _INTERNAL_K))
        struct.pack_into("<QQQ", tree, 24, 0, snod_addr, offsets[-1] if offsets else 0)
        tree_addr = self.append(bytes(tree))
        oh = self.append(_object_header([_message(0x0011, struct.pack("<QQ", tree_addr, heap_addr))]))
        return oh, tree_addr, heap_addr

    def finish(self, root):
        oh, bt, hp = self.group(root)
        self.buf.extend(b"\0" * (-len(self.buf) % 8))
        sb = SIGNATURE + struct.pack("<8BHHI", 0, 0, 0, 0, 0, 8, 8, 0, self.leaf_k, _INTERNAL_K, 0)
        sb += struct.pack("<QQQQ", 0, UNDEF, len(self.buf), UNDEF)
        sb += struct.pack("<QQII", 0, oh, 1, 0) + struct.pack("<QQ", bt, hp)
        assert len(sb) == 96
        self.buf[0:96] = sb
        return bytes(self.buf)


def _max_members(g):
    n = len(g.members)
    for m in g.members.values():
        if isinstance(m, Group):
            n = max(n, _max_members(m))
    return n


def encode(root):
    """The bytes of an HDF5 file holding the tree under `root` (a Group)."""
    leaf_k = max(_LEAF_K_MIN, (_max_members(root) + 1) // 2)
    if leaf_k > 0xFFFF:
        raise HDF5FormatError("group too large")
    return _Writer(leaf_k).finish(root)


def write(filename, root):
    with open(filename, "wb") as f:
        f.write(encode(root))


# ---- decoding (the subset above) ----

class _Reader:
    def __init__(self, buf):
        self.b = buf
        if buf[:8] != SIGNATURE:
            raise HDF5FormatError("not an HDF5 file (signature)")
        (sv, fsv, rv, _, shv, so, sl, _, self.leaf_k, self.int_k, _) = struct.unpack_from("<8BHHI", buf, 8)
        if (sv, so, sl) != (0, 8, 8):
            raise HDF5FormatError("only version 0 superblocks with 8-byte offsets are read")
        base, _, eof, _ = struct.unpack_from("<QQQQ", buf, 24)
        if base != 0 or eof != len(buf):
            raise HDF5FormatError("base address or end-of-file address mismatch")
        _, self.root_oh, _, _ = struct.unpack_from("<QQII", buf, 56)

    def check(self, addr, n, what):
        if addr == UNDEF or addr + n > len(self.b):
            raise HDF5FormatError("%s at %#x outside the file" % (what, addr))

    def messages(self, addr):
        self.check(addr, 16, "object header")
        ver, _, nmsg, _, size = struct.unpack_from("<BBHII", self.b, addr)
        if ver != 1:
            raise HDF5FormatError("only version 1 object headers are read")
        self.check(addr + 16, size, "object header messages")
        out, p = [], addr + 16
        for _ in range(nmsg):
            mtype, msize, _ = struct.unpack_from("<HHB", self.b, p)
            out.append((mtype, self.b[p + 8:p + 8 + msize]))
            p += 8 + msize
        if p > addr + 16 + size:
            raise HDF5FormatError("object header messages overrun the header")
        return out

    def group(self, addr):
        msgs = dict(self.messages(addr))
        if 0x0011 not in msgs:
            raise HDF5FormatError("object at %#x is not a symbol-table group" % addr)
        tree, heap = struct.unpack_from("<QQ", msgs[0x0011])
        self.check(heap, 32, "local heap")
        if self.b[heap:heap + 4] != b"HEAP":
            raise HDF5FormatError("local heap signature")
        hsize, free, hdata = struct.unpack_from("<QQQ", self.b, heap + 8)
        self.check(hdata, hsize, "local heap data")
        if free not in (_FREE_NULL, UNDEF) and free >= hsize:
            raise HDF5FormatError("local heap free list outside its data segment")
        self.check(tree, 24, "B-tree node")
        if self.b[tree:tree + 4] != b"TREE":
            raise HDF5FormatError("B-tree signature")
        ntype, level, used = struct.unpack_from("<BBH", self.b, tree + 4)
        if ntype != 0 or level != 0:
            raise HDF5FormatError("only single-level group B-trees are read")
        g = Group()
        for c in range(used):
            snod = struct.unpack_from("<Q", self.b, tree + 24 + 8 + 16 * c)[0]
            self.check(snod, 8, "symbol table node")
            if self.b[snod:snod + 4] != b"SNOD":
                raise HDF5FormatError("symbol table node signature")
            nsym = struct.unpack_from("<H", self.b, snod + 6)[0]
            for k in range(nsym):
                off, oh, _, _ = struct.unpack_from("<QQII", self.b, snod + 8 + _ENTRY * k)
                end = self.b.index(b"\0", hdata + off)
                name = self.b[hdata + off:end].decode("utf-8")
                sub = dict(self.messages(oh))
                g.members[name] = self.group(oh) if 0x0011 in sub else self.dataset(sub)
        return g

    def dataset(self, msgs):
        space, dtype, layout = msgs.get(0x0001), msgs.get(0x0003), msgs.get(0x0008)
        if space is None or dtype is None or layout is None:
            raise HDF5FormatError("dataset without dataspace, datatype or layout")
        ver, rank = space[0], space[1]
        if ver != 1:
            raise HDF5FormatError("only version 1 dataspaces are read")
        shape = tuple(struct.unpack_from("<%dQ" % rank, space, 8)) if rank else ()
        if layout[0] != 3 or layout[1] != 1:
            raise HDF5FormatError("only contiguous version 3 layouts are read")
        addr, size = struct.unpack_from("<QQ", layout, 2)
        raw = b""
        if size:
            self.check(addr, size, "dataset data")
            raw = self.b[addr:addr + size]
        cls, bits, tsize = dtype[0] & 0x0F, dtype[1], struct.unpack_from("<I", dtype, 4)[0]
        count = int(np.prod(shape)) if shape else 1
        if size != count * tsize:
            raise HDF5FormatError("dataset size does not match its dataspace and datatype")
        if cls == 3:
            vals = [raw[i * tsize:(i + 1) * tsize].rstrip(b"\0").decode("utf-8") for i in range(count)]
            return vals[0] if not shape else vals
        if cls == 8:
            v = np.frombuffer(raw, dtype=np.uint8).astype(bool)
            return bool(v[0]) if not shape else v.reshape(shape)
        if cls == 1:
            dt = np.dtype("<f%d" % tsize)
        elif cls == 0:
            dt = np.dtype("<%s%d" % ("i" if bits & 0x08 else "u", tsize))
        else:
            raise HDF5FormatError("datatype class %d is not read" % cls)
        a = np.frombuffer(raw, dtype=dt).copy()
        return a[0] if not shape else a.reshape(shape)


def decode(buf):
    """The tree of a file written by `encode`: Groups of datasets, each dataset
    a numpy array in the dataspace's C order (scalars as numpy scalars), a
    string or list of strings, or a bool."""
    r = _Reader(bytes(buf))
    return r.group(r.root_oh)


def read(filename):
    with open(filename, "rb") as f:
        return decode(f.read())
