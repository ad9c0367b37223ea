"""MATLAB .mat files: v5 through scipy, v7.3 (HDF5) through a small reader of the
HDF5 subset MATLAB writes.

The reference's inputs come in both formats. The physics tables are v5, and
scipy.io.loadmat reads them. The measured diffraction patterns
(open_source_data/*.mat, read by XRD_builder.m:37,87,137), the HD source spectrum
(Spectrum_80kVp/spectral_data_HD.mat, spectrum_builder.m:9) and
PE_and_Compton_coefs.mat are v7.3. That is HDF5 with a 512-byte MATLAB user
block, and this image has no HDF5 library.

The reader follows the public HDF5 file format specification, version 0 superblock
layout:

  * superblock v0/v1 (8-byte offsets and lengths), addresses relative to the
    superblock, which sits after the user block;
  * groups as symbol tables: a v1 B-tree of type 0 over SNOD nodes, names in a local
    heap;
  * v1 object headers with continuation blocks;
  * dataspace v1; little-endian fixed-point and floating-point data;
  * data layout v3: compact, contiguous, and chunked through a v1 B-tree of type 1
    with the deflate filter;
  * v1 attribute messages (MATLAB_class: double, single, integer and logical arrays).

That subset covers every v7.3 file the reference ships. Their chunked datasets
are single chunks: assembling several chunks (and walking chunk B-trees deeper than
one level) follows the specification but no shipped file exercises it. Anything
outside the subset (complex or sparse data, cell and char arrays, other filters,
v2 object headers) raises ValueError naming what was found. Every B-tree, symbol-table and heap
signature is checked, and every read is bounds-checked against the file.

MATLAB stores an array of size [d1, d2, ...] as an HDF5 dataset of dims [..., d2, d1]
(column-major as row-major, reversed). load_mat returns arrays in MATLAB's shape,
so D(:, 1) of an N x 2 matrix is out["D"][:, 0]. A struct becomes a dict of its
fields.
"""
from __future__ import annotations

import zlib
from typing import Dict, Optional, Sequence

import numpy as np

_HDF5_SIG = b"\x89HDF\r\n\x1a\n"
_UNDEF = 0xFFFFFFFFFFFFFFFF


def mat_version(path: str) -> str:
    """'5' or '7.3', from the 128-byte MATLAB header."""
    with open(path, "rb") as f:
        head = f.read(128)
    if head.startswith(b"MATLAB 7.3 MAT-file"):
        return "7.3"
    if head.startswith(b"MATLAB 5.0 MAT-file"):
        return "5"
    raise ValueError(f"{path}: not a MATLAB v5 or v7.3 MAT-file")


def load_mat(path: str, names: Optional[Sequence[str]] = None) -> Dict[str, object]:
    """Variables of a MAT-file as numpy values. names restricts the result to those
    variables, as MATLAB's load(path, names...) does; a missing name raises KeyError."""
    if mat_version(path) == "5":
        import scipy.io
        raw = scipy.io.loadmat(path, squeeze_me=False, struct_as_record=False,
                               chars_as_strings=True)
        out = {k: _from_scipy(v) for k, v in raw.items() if not k.startswith("__")}
    else:
        with open(path, "rb") as f:
            out = _H5File(f.read()).load_root()
    if names is None:
        return out
    missing = [n for n in names if n not in out]
    if missing:
        raise KeyError(f"{path}: no variable(s) {missing}")
    return {n: out[n] for n in names}


def _from_scipy(v):
    from scipy.io.matlab import mat_struct
    if isinstance(v, mat_struct):
        return {k: _from_scipy(getattr(v, k)) for k in v._fieldnames}
    if isinstance(v, np.ndarray) and v.dtype == object:
        if v.size == 1 and isinstance(v.flat[0], mat_struct):
            return _from_scipy(v.flat[0])
        out = np.empty(v.shape, dtype=object)
        for i, x in enumerate(v.flat):
            out.flat[i] = _from_scipy(x)
        return out
    if isinstance(v, np.ndarray) and v.dtype.kind == "U" and v.size == 1:
        return str(v.flat[0])
    return v


# ---------------------------------------------------------------------------
# HDF5 reader

class _Reader:
    """Bounds-checked little-endian reads from the file image."""

    def __init__(self, buf: bytes, base: int):
        self.buf = buf
        self.base = base

    def at(self, addr: int, n: int) -> bytes:
        off = self.base + addr
        if addr == _UNDEF or off < 0 or off + n > len(self.buf):
            raise ValueError(f"HDF5: read of {n} bytes at address {addr:#x} is outside the file")
        return self.buf[off:off + n]

    def u(self, addr: int, n: int) -> int:
        return int.from_bytes(self.at(addr, n), "little")


class _Message:
    __slots__ = ("mtype", "data")

    def __init__(self, mtype: int, data: bytes):
        self.mtype = mtype
        self.data = data


class _H5File:
    def __init__(self, buf: bytes):
        sb = -1
        for off in (0, 512, 1024, 2048, 4096):
            if buf[off:off + 8] == _HDF5_SIG:
                sb = off
                break
        if sb < 0:
            raise ValueError("HDF5: no superblock signature at 0 or a power-of-two user block")
        ver = buf[sb + 8]
        if ver not in (0, 1):
            raise ValueError(f"HDF5: superblock version {ver} is outside the supported subset (0, 1)")
        so, sl = buf[sb + 13], buf[sb + 14]
        if so != 8 or sl != 8:
            raise ValueError(f"HDF5: offset/length sizes {so}/{sl}; the reader handles 8/8")
        self.r = _Reader(buf, sb)
        p = 24 + (4 if ver == 1 else 0)
        base_addr = self.r.u(p, 8)
        if base_addr not in (0, sb):
            raise ValueError(f"HDF5: base address {base_addr:#x} with the superblock at {sb:#x}")
        # addresses are relative to the superblock; the root group's symbol table entry
        # follows the four addresses (base, free space, end of file, driver info)
        self.root_entry = p + 32
        self.root_header = self.r.u(self.root_entry + 8, 8)

    # --- object headers -----------------------------------------------------
    def messages(self, addr: int):
        r = self.r
        if r.at(addr, 4) == b"OHDR":
            raise ValueError("HDF5: version 2 object headers are outside the supported subset")
        if r.u(addr, 1) != 1:
            raise ValueError(f"HDF5: object header version {r.u(addr, 1)} at {addr:#x}")
        nmsg = r.u(addr + 2, 2)
        size = r.u(addr + 8, 4)
        blocks = [(addr + 16, size)]
        out = []
        while blocks and len(out) < nmsg:
            start, length = blocks.pop(0)
            p, end = start, start + length
            while p + 8 <= end and len(out) < nmsg:
                mtype, msize = r.u(p, 2), r.u(p + 2, 2)
                data = r.at(p + 8, msize)
                if mtype == 0x10:           # continuation
                    blocks.append((int.from_bytes(data[0:8], "little"),
                                   int.from_bytes(data[8:16], "little")))
                out.append(_Message(mtype, data))
                p += 8 + msize
        return out

    # --- groups ---------------------------------------------------------------
    def heap_name(self, heap_addr: int, off: int) -> str:
        r = self.r
        if r.at(heap_addr, 4) != b"HEAP":
            raise ValueError(f"HDF5: no local heap at {heap_addr:#x}")
        seg_size = r.u(heap_addr + 8, 8)
        seg = r.u(heap_addr + 24, 8)
        if off >= seg_size:
            raise ValueError("HDF5: link name offset outside the local heap")
        raw = r.at(seg + off, seg_size - off)
        return raw[:raw.index(b"\0")].decode("utf-8")

    def group_links(self, btree: int, heap: int) -> Dict[str, int]:
        r = self.r
        links: Dict[str, int] = {}
        todo = [btree]
        while todo:
            node = todo.pop()
            if r.at(node, 4) != b"TREE" or r.u(node + 4, 1) != 0:
                raise ValueError(f"HDF5: no group B-tree node at {node:#x}")
            level, used = r.u(node + 5, 1), r.u(node + 6, 2)
            p = node + 24 + 8                      # past the header and key 0
            for _ in range(used):
                child = r.u(p, 8)
                p += 16                            # child address, next key
                if level > 0:
                    todo.append(child)
                    continue
                if r.at(child, 4) != b"SNOD":
                    raise ValueError(f"HDF5: no symbol table node at {child:#x}")
                nsym = r.u(child + 6, 2)
                e = child + 8
                for _ in range(nsym):
                    links[self.heap_name(heap, r.u(e, 8))] = r.u(e + 8, 8)
                    e += 40
        return links

    # --- datatype / dataspace ------------------------------------------------
    @staticmethod
    def dataspace(d: bytes):
        ver, rank = d[0], d[1]
        if ver != 1:
            raise ValueError(f"HDF5: dataspace version {ver} is outside the supported subset")
        return tuple(int.from_bytes(d[8 + 8 * i:16 + 8 * i], "little") for i in range(rank))

    @staticmethod
    def datatype(d: bytes):
        """('num', numpy dtype) for fixed/floating point, ('str', size) for strings."""
        cls, bits = d[0] & 0x0F, d[1]
        size = int.from_bytes(d[4:8], "little")
        if cls in (0, 1):
            if bits & 1:
                raise ValueError("HDF5: big-endian data is outside the supported subset")
            kind = "f" if cls == 1 else ("i" if bits & 0x08 else "u")
            if (kind == "f" and size not in (4, 8)) or size not in (1, 2, 4, 8):
                raise ValueError(f"HDF5: {size}-byte {kind} data is outside the supported subset")
            return "num", np.dtype(f"<{kind}{size}")
        if cls == 3:
            return "str", size
        raise ValueError(f"HDF5: datatype class {cls} is outside the supported subset "
                         "(complex, cell, char and sparse MATLAB data)")

    # --- layout and filters ---------------------------------------------------
    @staticmethod
    def check_filters(d: bytes) -> None:
        """The v1 filter pipeline may hold deflate only."""
        ver, n = d[0], d[1]
        if ver != 1:
            raise ValueError(f"HDF5: filter pipeline version {ver} is outside the supported subset")
        p = 8
        for _ in range(n):
            fid, nlen, _, nval = (int.from_bytes(d[p + 2 * i:p + 2 * i + 2], "little")
                                  for i in range(4))
            if fid != 1:
                raise ValueError(f"HDF5: filter {fid} is outside the supported subset (deflate)")
            p += 8 + (nlen + 7) // 8 * 8 + 4 * nval + (4 if nval % 2 else 0)

    def read_data(self, layout: bytes, dims, elem: int, deflate: bool) -> bytes:
        r = self.r
        ver, cls = layout[0], layout[1]
        if ver != 3:
            raise ValueError(f"HDF5: data layout version {ver} is outside the supported subset")
        total = int(np.prod(dims)) * elem
        if cls == 0:                               # compact: the data is in the message
            size = int.from_bytes(layout[2:4], "little")
            if size < total:
                raise ValueError("HDF5: compact data shorter than its dataspace")
            return layout[4:4 + total]
        if cls == 1:                               # contiguous
            addr = int.from_bytes(layout[2:10], "little")
            return bytes(total) if addr == _UNDEF else r.at(addr, total)
        if cls != 2:
            raise ValueError(f"HDF5: layout class {cls}")
        nd = layout[2]                             # rank + 1 (the element size)
        rank = nd - 1
        if rank != len(dims):
            raise ValueError("HDF5: chunk rank differs from the dataspace rank")
        btree = int.from_bytes(layout[3:11], "little")
        cshape = [int.from_bytes(layout[11 + 4 * i:15 + 4 * i], "little") for i in range(rank)]
        out = np.zeros((*dims, elem), np.uint8)
        todo = [] if btree == _UNDEF else [btree]
        while todo:
            node = todo.pop()
            if r.at(node, 4) != b"TREE" or r.u(node + 4, 1) != 1:
                raise ValueError(f"HDF5: no chunk B-tree node at {node:#x}")
            level, used = r.u(node + 5, 1), r.u(node + 6, 2)
            ksize = 8 + 8 * nd                     # chunk size, filter mask, offsets
            p = node + 24
            for _ in range(used):
                csize, mask = r.u(p, 4), r.u(p + 4, 4)
                offs = [r.u(p + 8 + 8 * i, 8) for i in range(rank)]
                child = r.u(p + ksize, 8)
                p += ksize + 8
                if level > 0:
                    todo.append(child)
                    continue
                raw = r.at(child, csize)
                if deflate and not mask & 1:
                    raw = zlib.decompress(raw)
                need = int(np.prod(cshape)) * elem
                if len(raw) != need:
                    raise ValueError("HDF5: chunk size differs from its declared shape")
                chunk = np.frombuffer(raw, np.uint8).reshape(*cshape, elem)
                if any(o >= dd for o, dd in zip(offs, dims)):
                    raise ValueError("HDF5: chunk offset outside the dataspace")
                sl_out = tuple(slice(o, min(o + c, dd)) for o, c, dd in zip(offs, cshape, dims))
                sl_in = tuple(slice(0, s.stop - s.start) for s in sl_out)
                out[sl_out] = chunk[sl_in]
        return out.tobytes()

    # --- objects ----------------------------------------------------------------
    def attributes(self, msgs) -> Dict[str, object]:
        """String and numeric v1 attributes by name. Variable-length attributes
        (MATLAB_fields of a struct group, whose field names are also the group's
        links) are skipped."""
        attrs = {}
        for m in msgs:
            if m.mtype != 0x0C:
                continue
            d = m.data
            if d[0] != 1:
                raise ValueError(f"HDF5: attribute message version {d[0]} is outside the "
                                 "supported subset")
            nsize, tsize, ssize = (int.from_bytes(d[2 + 2 * i:4 + 2 * i], "little") for i in range(3))
            pad = lambda n: (n + 7) // 8 * 8      # noqa: E731
            p = 8
            name = d[p:p + nsize].split(b"\0")[0].decode()
            p += pad(nsize)
            if d[p] & 0x0F == 9:
                continue
            kind, dt = self.datatype(d[p:p + tsize])
            p += pad(tsize)
            dims = self.dataspace(d[p:p + ssize])
            p += pad(ssize)
            raw = d[p:]
            if kind == "str":
                attrs[name] = raw[:dt].split(b"\0")[0].decode("utf-8", "replace")
            else:
                n = int(np.prod(dims))
                attrs[name] = np.frombuffer(raw[:n * dt.itemsize], dt).copy()
        return attrs

    def load_object(self, addr: int, depth: int = 0):
        if depth > 32:
            raise ValueError("HDF5: object nesting deeper than 32")
        msgs = self.messages(addr)
        by = {}
        for m in msgs:
            by.setdefault(m.mtype, m.data)
        if 0x11 in by:                             # group: a MATLAB struct
            st = by[0x11]
            links = self.group_links(int.from_bytes(st[0:8], "little"),
                                     int.from_bytes(st[8:16], "little"))
            return {k: self.load_object(v, depth + 1) for k, v in sorted(links.items())}
        if 0x06 in by or 0x02 in by:
            raise ValueError("HDF5: link-message groups are outside the supported subset")
        if 0x01 not in by or 0x03 not in by or 0x08 not in by:
            raise ValueError(f"HDF5: object at {addr:#x} is neither a group nor a dataset")
        attrs = self.attributes(msgs)
        mclass = attrs.get("MATLAB_class", "double")
        if mclass not in ("double", "single", "logical", "int8", "uint8", "int16", "uint16",
                          "int32", "uint32", "int64", "uint64") or "MATLAB_sparse" in attrs \
                or "MATLAB_empty" in attrs:
            raise ValueError(f"HDF5: MATLAB class {mclass!r} (or a sparse/empty array) is "
                             "outside the supported subset")
        dims = self.dataspace(by[0x01])
        kind, dt = self.datatype(by[0x03])
        if kind != "num":
            raise ValueError("HDF5: string datasets are outside the supported subset")
        if 0x0B in by:
            self.check_filters(by[0x0B])
        raw = self.read_data(by[0x08], dims, dt.itemsize, 0x0B in by)
        a = np.frombuffer(raw, dt).reshape(dims).copy().T   # HDF5 dims are MATLAB's, reversed
        return a.astype(bool) if mclass == "logical" else a

    def load_root(self) -> Dict[str, object]:
        msgs = self.messages(self.root_header)
        st = next((m.data for m in msgs if m.mtype == 0x11), None)
        if st is None:
            raise ValueError("HDF5: the root group has no symbol table message")
        links = self.group_links(int.from_bytes(st[0:8], "little"),
                                 int.from_bytes(st[8:16], "little"))
        # '#refs#' holds the targets of object references (cell arrays); a variable
        # that uses them has datatype class 7 and is rejected when read
        return {k: self.load_object(v) for k, v in sorted(links.items()) if k != "#refs#"}
