"""MAT-file version 7.3 writer and reader for the snapshot dictionaries that the reference
saves with MAT.jl's `matwrite` (simulation.jl:118-138, scripts/Main.jl:138-152; MAT.jl
writes version 7.3 unless asked otherwise).

A 7.3 MAT-file is an HDF5 file behind a 512-byte user block whose first 128 bytes are the
MATLAB header (text, subsystem offset, version 0x0200, endian tag "IM").  The writer emits
the part of the HDF5 format (version 0 superblock, version 1 object headers) that a
dictionary of real double arrays needs, with no HDF5 library:

* the root group as a symbol table: a version-1 group B-tree whose children are symbol
  table nodes (at most 2 x 4 entries each, sorted by name) and a local heap of the names;
* one dataset per variable: dataspace, IEEE float64 little-endian datatype, fill value,
  contiguous layout, and the attribute MATLAB_class = "double" that MATLAB and MAT.jl read
  the class from.

MATLAB arrays are column-major and HDF5 dataspaces row-major: an array of size
(n1, ..., nk) is stored as its column-major bytes under the dataspace (nk, ..., n1), as
MAT.jl and MATLAB store it; a scalar is a 1 x 1 array, a vector of length n an n x 1 one.

`read_mat73` reads that subset back, whoever wrote it (contiguous float64 datasets in a
symbol-table root group, version 1 object headers with continuation blocks), e.g. the
HDF5 library's own files with the 512-byte user block; chunked or compressed datasets,
which MATLAB itself may write for large arrays, are refused with an error.
"""
from __future__ import annotations

import os
import struct
import time
from typing import Dict

import numpy as np

_SIG = b"\x89HDF\r\n\x1a\n"
_UNDEF = 0xFFFFFFFFFFFFFFFF     # the undefined address (8-byte offsets)
_USERBLOCK = 512
_LEAF_K = 4                     # symbol table nodes hold at most 2 * _LEAF_K entries
_INTERNAL_K = 16                # group B-tree nodes hold at most 2 * _INTERNAL_K children
_HEAP_FREE_NONE = 1             # local heap free-list offset meaning "no free block"

# object header message types
_M_DATASPACE, _M_DATATYPE, _M_FILL, _M_LAYOUT, _M_ATTR, _M_CONT, _M_STAB = 0x1, 0x3, 0x5, 0x8, 0xC, 0x10, 0x11

# IEEE float64, little-endian: class 1 version 1; bit field: byte order 0, mantissa
# normalisation 2 (implied leading 1), sign bit 63; properties: offset 0, precision 64,
# exponent at bit 52 of 11 bits, mantissa at bit 0 of 52 bits, bias 1023
_F64 = struct.pack("<B3BI", 0x11, 0x20, 63, 0, 8) + struct.pack("<HHBBBBI", 0, 64, 52, 11, 0, 52, 1023)


def _pad8(b: bytes) -> bytes:
    return b + b"\0" * (-len(b) % 8)


def _dataspace(dims) -> bytes:
    """Version 1 dataspace: rank, no maximum dimensions; rank 0 is a scalar."""
    return struct.pack("<BBB5x", 1, len(dims), 0) + b"".join(struct.pack("<Q", int(d)) for d in dims)


def _string_type(n: int) -> bytes:
    """Fixed-length ASCII string of n bytes, null-terminated padding (class 3, version 1)."""
    return struct.pack("<B3BI", 0x13, 0, 0, 0, n)


def _msg(mtype: int, data: bytes, flags: int = 0) -> bytes:
    data = _pad8(data)
    return struct.pack("<HHB3x", mtype, len(data), flags) + data


def _object_header(msgs) -> bytes:
    """Version 1 object header: 16-byte prefix (version, message count, reference count,
    size of the message data, alignment) followed by the messages."""
    body = b"".join(msgs)
    return struct.pack("<BBHII4x", 1, 0, len(msgs), 1, len(body)) + body


def _attr_matlab_class(cls: str) -> bytes:
    name = b"MATLAB_class\0"
    dt = _string_type(len(cls))
    ds = _dataspace(())
    return (struct.pack("<BBHHH", 1, 0, len(name), len(dt), len(ds)) + _pad8(name) + _pad8(dt) + _pad8(ds)
            + cls.encode("ascii"))


def _entry(name_off: int, ohdr: int, cache: int = 0, scratch: bytes = b"\0" * 16) -> bytes:
    """Symbol table entry (40 bytes with 8-byte offsets)."""
    return struct.pack("<QQII", name_off, ohdr, cache, 0) + scratch


def matlab_header(created: float | None = None) -> bytes:
    """The 128-byte MATLAB header that opens the user block of a 7.3 MAT-file."""
    stamp = time.strftime("%a %b %d %H:%M:%S %Y", time.gmtime(created if created is not None else time.time()))
    text = f"MATLAB 7.3 MAT-file, Platform: GLNXA64, Created on: {stamp} HDF5 schema 1.00 .".encode("ascii")
    return text.ljust(116, b" ") + b"\0" * 8 + struct.pack("<H", 0x0200) + b"IM"


def _as_matlab(v) -> np.ndarray:
    a = np.asarray(v, dtype=np.float64)
    if a.ndim == 0:
        a = a.reshape(1, 1)
    elif a.ndim == 1:
        a = a.reshape(-1, 1)
    return a


def write_mat73(path: os.PathLike, variables: Dict[str, object]):
    """Writes `variables` (name -> real array or scalar) as a 7.3 MAT-file."""
    names = sorted(variables)  # symbol table order: byte-wise name comparison
    if not names:
        raise ValueError("write_mat73: nothing to write")
    if len(names) > 2 * _LEAF_K * 2 * _INTERNAL_K:
        raise ValueError("write_mat73: too many variables for one B-tree node")
    for n in names:
        if not n or "\0" in n or "/" in n or not n.isascii():
            raise ValueError(f"write_mat73: bad variable name {n!r}")
    arrays = [_as_matlab(variables[n]) for n in names]

    # local heap data segment: "" at offset 0 (the root's name), then each name
    heap = bytearray(b"\0" * 8)
    name_off = []
    for n in names:
        name_off.append(len(heap))
        heap += _pad8(n.encode("ascii") + b"\0")
    nodes = [list(range(i, min(i + 2 * _LEAF_K, len(names)))) for i in range(0, len(names), 2 * _LEAF_K)]

    # sizes, then addresses relative to the superblock (the base address; the end-of-file
    # address alone is absolute, the HDF5 library reads it as the file's length)
    sb_size = 96
    root_oh = _object_header([_msg(_M_STAB, b"\0" * 16)])
    heap_hdr_size = 32
    btree_size = 24 + (2 * _INTERNAL_K + 1) * 8 + 2 * _INTERNAL_K * 8
    snod_size = 8 + 2 * _LEAF_K * 40

    def dset_header(a, addr):
        return _object_header([
            _msg(_M_DATASPACE, _dataspace(a.shape[::-1])),
            _msg(_M_DATATYPE, _F64, flags=1),                     # constant
            _msg(_M_FILL, struct.pack("<BBBB", 2, 2, 2, 0), flags=1),  # late, if set, undefined
            _msg(_M_LAYOUT, struct.pack("<BBQQ", 3, 1, addr, a.size * 8)),
            _msg(_M_ATTR, _attr_matlab_class("double")),
        ])

    addr = sb_size
    a_root = addr
    addr += len(root_oh)
    a_heap = addr
    a_heap_data = a_heap + heap_hdr_size
    addr = a_heap_data + len(heap)
    a_btree = addr
    addr += btree_size
    a_snod = []
    for _ in nodes:
        a_snod.append(addr)
        addr += snod_size
    a_oh = []
    for a in arrays:
        a_oh.append(addr)
        addr += len(dset_header(a, 0))
    a_data = []
    for a in arrays:
        a_data.append(addr)
        addr += a.size * 8
    eof = addr

    out = bytearray(matlab_header().ljust(_USERBLOCK, b"\0"))
    root_entry = _entry(0, a_root, 1, struct.pack("<QQ", a_btree, a_heap))
    sb = (_SIG + struct.pack("<8BHHI", 0, 0, 0, 0, 0, 8, 8, 0, _LEAF_K, _INTERNAL_K, 0)
          + struct.pack("<QQQQ", _USERBLOCK, _UNDEF, _USERBLOCK + eof, _UNDEF) + root_entry)
    assert len(sb) == sb_size
    out += sb
    out += _object_header([_msg(_M_STAB, struct.pack("<QQ", a_btree, a_heap))])
    out += b"HEAP" + struct.pack("<B3xQQQ", 0, len(heap), _HEAP_FREE_NONE, a_heap_data) + heap
    # B-tree (level 0): key 0 = "", child i = symbol table node i, key i+1 = its last name
    bt = b"TREE" + struct.pack("<BBHQQ", 0, 0, len(nodes), _UNDEF, _UNDEF) + struct.pack("<Q", 0)
    for i, nd in enumerate(nodes):
        bt += struct.pack("<QQ", a_snod[i], name_off[nd[-1]])
    out += bt.ljust(btree_size, b"\0")
    for i, nd in enumerate(nodes):
        sn = b"SNOD" + struct.pack("<BBH", 1, 0, len(nd)) + b"".join(_entry(name_off[j], a_oh[j]) for j in nd)
        out += sn.ljust(snod_size, b"\0")
    for a, d in zip(arrays, a_data):
        out += dset_header(a, d)
    for a in arrays:
        out += a.tobytes(order="F")
    assert len(out) == _USERBLOCK + eof
    with open(path, "wb") as f:
        f.write(bytes(out))


def is_mat73(path: os.PathLike) -> bool:
    """True when the MATLAB header says version 7.3 (0x0200) and HDF5 data follow at 512."""
    with open(path, "rb") as f:
        head = f.read(_USERBLOCK + 8)
    return len(head) == _USERBLOCK + 8 and head[124:128] == struct.pack("<H", 0x0200) + b"IM" \
        and head[_USERBLOCK:] == _SIG


class _Reader:
    def __init__(self, buf: bytes):
        self.b = buf
        if buf[_USERBLOCK:_USERBLOCK + 8] != _SIG:
            raise ValueError("not an HDF5 file with a 512-byte user block")
        sb = _USERBLOCK
        ver, = struct.unpack_from("<B", buf, sb + 8)
        if ver != 0:
            raise ValueError(f"superblock version {ver} is not read here")
        so, sl = struct.unpack_from("<BB", buf, sb + 13)
        if (so, sl) != (8, 8):
            raise ValueError("offsets and lengths other than 8 bytes are not read here")
        self.leaf_k, self.int_k = struct.unpack_from("<HH", buf, sb + 16)
        base, _, eof, _ = struct.unpack_from("<QQQQ", buf, sb + 24)
        self.base = sb  # addresses are relative to the superblock
        if eof > len(buf):  # the end-of-file address counts the user block
            raise ValueError("truncated file")
        _, root_oh, cache, _ = struct.unpack_from("<QQII", buf, sb + 56)
        self.root_oh = root_oh

    def at(self, a: int) -> int:
        return self.base + a

    def messages(self, oh: int):
        """The messages of a version 1 object header, following continuation blocks
        (message 0x10: address and length of a further block of messages)."""
        p = self.at(oh)
        ver, _, nm, _, size = struct.unpack_from("<BBHII", self.b, p)
        if ver != 1:
            raise ValueError(f"object header version {ver} is not read here")
        chunks = [(p + 16, p + 16 + size)]
        seen = 0
        while chunks and seen < nm:
            q, end = chunks.pop(0)
            while q + 8 <= end and seen < nm:
                t, sz = struct.unpack_from("<HH", self.b, q)
                data = self.b[q + 8:q + 8 + sz]
                seen += 1
                if t == _M_CONT:
                    ca, cl = struct.unpack_from("<QQ", data)
                    chunks.append((self.at(ca), self.at(ca) + cl))
                else:
                    yield t, data
                q += 8 + sz

    def group(self, oh: int) -> Dict[str, int]:
        stab = [d for t, d in self.messages(oh) if t == _M_STAB]
        if not stab:
            raise ValueError("the root group has no symbol table")
        bt, hp = struct.unpack_from("<QQ", stab[0])
        p = self.at(hp)
        if self.b[p:p + 4] != b"HEAP":
            raise ValueError("bad local heap")
        seg_size, _, seg = struct.unpack_from("<QQQ", self.b, p + 8)
        heap = self.b[self.at(seg):self.at(seg) + seg_size]

        def name(off):
            return heap[off:heap.index(b"\0", off)].decode("ascii")

        out = {}

        def walk(node):
            p = self.at(node)
            if self.b[p:p + 4] != b"TREE":
                raise ValueError("bad group B-tree node")
            typ, level, used = struct.unpack_from("<BBH", self.b, p + 4)
            for i in range(used):
                child, = struct.unpack_from("<Q", self.b, p + 24 + 8 + 16 * i)
                if level > 0:
                    walk(child)
                    continue
                s = self.at(child)
                if self.b[s:s + 4] != b"SNOD":
                    raise ValueError("bad symbol table node")
                n, = struct.unpack_from("<H", self.b, s + 6)
                for j in range(n):
                    off, ohdr = struct.unpack_from("<QQ", self.b, s + 8 + 40 * j)
                    out[name(off)] = ohdr

        walk(bt)
        return out

    def dataset(self, oh: int):
        dims = layout = dtype = None
        cls = None
        for t, d in self.messages(oh):
            if t == _M_DATASPACE:
                ver, rank = struct.unpack_from("<BB", d)
                off = 8 if ver == 1 else 4
                dims = struct.unpack_from(f"<{rank}Q", d, off)
            elif t == _M_DATATYPE:
                dtype = d[:20]
            elif t == _M_LAYOUT:
                ver, lc = struct.unpack_from("<BB", d)
                if ver != 3 or lc != 1:
                    raise ValueError("only contiguous datasets are read here")
                layout = struct.unpack_from("<QQ", d, 2)
            elif t == _M_ATTR:
                ver, _, nsz, tsz, ssz = struct.unpack_from("<BBHHH", d)
                nm = d[8:8 + nsz].rstrip(b"\0").decode("ascii")
                if nm == "MATLAB_class":
                    p = 8 + len(_pad8(b"\0" * nsz)) + len(_pad8(b"\0" * tsz)) + len(_pad8(b"\0" * ssz))
                    n, = struct.unpack_from("<I", d, 8 + len(_pad8(b"\0" * nsz)) + 4)
                    cls = d[p:p + n].rstrip(b"\0").decode("ascii")
        if dims is None or layout is None or dtype is None:
            raise ValueError("incomplete dataset header")
        if dtype[:1] != b"\x11" or struct.unpack_from("<I", dtype, 4)[0] != 8 or dtype[1] & 1:
            raise ValueError("only little-endian float64 datasets are read here")
        a_addr, nbytes = layout
        n = int(np.prod(dims)) if dims else 1
        if nbytes != 8 * n:
            raise ValueError("layout size does not match the dataspace")
        p = self.at(a_addr)
        a = np.frombuffer(self.b, dtype="<f8", count=n, offset=p).reshape(tuple(dims)[::-1], order="F")
        return a.copy(), cls


def read_mat73(path: os.PathLike) -> Dict[str, np.ndarray]:
    """Reads the variables of a 7.3 MAT-file written as `write_mat73` writes them; arrays
    come back in MATLAB's shape (the dataspace reversed), scalars as 1 x 1 arrays."""
    with open(path, "rb") as f:
        r = _Reader(f.read())
    out = {}
    for name, oh in r.group(r.root_oh).items():
        a, cls = r.dataset(oh)
        if cls not in (None, "double"):
            raise ValueError(f"{name}: MATLAB class {cls} is not read here")
        out[name] = a
    return out
