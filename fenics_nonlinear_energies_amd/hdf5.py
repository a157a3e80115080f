"""Minimal read-only HDF5 reader for the reference's problem files (no h5py needed).

Supports what h5py writes for ``data/meshes/*.h5``: superblock v0/v1, version-1 object
headers, old-style groups (symbol-table B-tree + local heap), contiguous / compact /
chunked layouts, the deflate and shuffle filters, fixed-point, floating-point and enum
datatypes.  ``load_problem_hdf5`` mirrors core/problem_data/hdf5.py:20-35 of the
reference: it returns ``(params, adjacency)`` with the optional ``adjacency`` group turned
into a ``scipy.sparse.coo_matrix``.
"""

from __future__ import annotations

import struct
import zlib
from dataclasses import dataclass, field

import numpy as np

_SIG = b"\x89HDF\r\n\x1a\n"
_UNDEF = 0xFFFFFFFFFFFFFFFF


class HDF5FormatError(ValueError):
    pass


@dataclass
class _Dataset:
    shape: tuple[int, ...] = ()
    dtype: np.dtype | None = None
    layout_class: int = -1
    address: int = _UNDEF
    size: int = 0
    compact: bytes = b""
    chunk_dims: tuple[int, ...] = ()
    filters: list[int] = field(default_factory=list)


class HDF5File:
    """Tiny HDF5 reader; ``keys()``, ``group(name)``, ``read(path)``."""

    def __init__(self, path: str):
        with open(path, "rb") as fh:
            self.buf = fh.read()
        self._parse_superblock()

    # --- primitive helpers -------------------------------------------------------------
    def _u(self, off: int, n: int) -> int:
        return int.from_bytes(self.buf[off : off + n], "little")

    def _parse_superblock(self) -> None:
        b = self.buf
        base = b.find(_SIG)
        if base != 0:
            raise HDF5FormatError("HDF5 signature not found at offset 0")
        version = b[8]
        if version not in (0, 1):
            raise HDF5FormatError(f"unsupported superblock version {version}")
        self.so = b[13]
        self.sl = b[14]
        if self.so != 8 or self.sl != 8:
            raise HDF5FormatError("only 8-byte offsets/lengths supported")
        off = 24 + (4 if version == 1 else 0)
        # base, free-space, eof, driver
        off += 32
        # root group symbol table entry
        self.root_header = self._u(off + 8, 8)

    # --- object headers ----------------------------------------------------------------
    def _messages(self, addr: int):
        b = self.buf
        version = b[addr]
        if version != 1:
            raise HDF5FormatError(f"unsupported object header version {version} at {addr}")
        nmsg = self._u(addr + 2, 2)
        hsize = self._u(addr + 8, 4)
        blocks = [(addr + 16, hsize)]
        out = []
        while blocks and len(out) < nmsg:
            start, length = blocks.pop(0)
            off = start
            end = start + length
            while off + 8 <= end and len(out) < nmsg:
                mtype = self._u(off, 2)
                msize = self._u(off + 2, 2)
                data_off = off + 8
                if mtype == 0x10:  # continuation
                    blocks.append((self._u(data_off, 8), self._u(data_off + 8, 8)))
                out.append((mtype, data_off, msize))
                off = data_off + msize
        return out

    def _group_entries(self, header_addr: int) -> dict[str, int]:
        btree = heap = None
        for mtype, off, _ in self._messages(header_addr):
            if mtype == 0x11:
                btree = self._u(off, 8)
                heap = self._u(off + 8, 8)
        if btree is None:
            raise HDF5FormatError("object is not an old-style group")
        if self.buf[heap : heap + 4] != b"HEAP":
            raise HDF5FormatError("bad local heap signature")
        heap_data = self._u(heap + 24, 8)
        entries: dict[str, int] = {}
        self._walk_group_btree(btree, heap_data, entries)
        return entries

    def _walk_group_btree(self, addr: int, heap_data: int, entries: dict[str, int]) -> None:
        b = self.buf
        if b[addr : addr + 4] != b"TREE":
            raise HDF5FormatError("bad B-tree signature")
        level = b[addr + 5]
        nused = self._u(addr + 6, 2)
        off = addr + 24 + 8  # skip key0
        for _ in range(nused):
            child = self._u(off, 8)
            off += 16  # child + next key
            if level > 0:
                self._walk_group_btree(child, heap_data, entries)
                continue
            if b[child : child + 4] != b"SNOD":
                raise HDF5FormatError("bad symbol node signature")
            nsym = self._u(child + 6, 2)
            eo = child + 8
            for _s in range(nsym):
                name_off = self._u(eo, 8)
                ohdr = self._u(eo + 8, 8)
                nstart = heap_data + name_off
                nend = b.index(b"\x00", nstart)
                entries[b[nstart:nend].decode()] = ohdr
                eo += 40

    # --- datasets ----------------------------------------------------------------------
    def _datatype(self, off: int) -> np.dtype:
        b = self.buf
        cls = b[off] & 0x0F
        bits0 = b[off + 1]
        size = self._u(off + 4, 4)
        order = ">" if (bits0 & 1) else "<"
        if cls == 0:  # fixed point
            signed = bool(bits0 & 0x08)
            return np.dtype(f"{order}{'i' if signed else 'u'}{size}")
        if cls == 1:
            return np.dtype(f"{order}f{size}")
        if cls == 8:  # enum: base type follows the 8-byte header
            return self._datatype(off + 8)
        raise HDF5FormatError(f"unsupported datatype class {cls}")

    def _dataset(self, header_addr: int) -> _Dataset:
        ds = _Dataset()
        b = self.buf
        for mtype, off, _ in self._messages(header_addr):
            if mtype == 0x01:  # dataspace
                ver = b[off]
                rank = b[off + 1]
                if ver == 1:
                    doff = off + 8
                elif ver == 2:
                    doff = off + 4
                else:
                    raise HDF5FormatError(f"dataspace version {ver}")
                ds.shape = tuple(self._u(doff + 8 * i, 8) for i in range(rank))
            elif mtype == 0x03:
                ds.dtype = self._datatype(off)
            elif mtype == 0x08:  # layout
                ver = b[off]
                if ver != 3:
                    raise HDF5FormatError(f"layout version {ver}")
                ds.layout_class = b[off + 1]
                if ds.layout_class == 0:
                    n = self._u(off + 2, 2)
                    ds.compact = b[off + 4 : off + 4 + n]
                elif ds.layout_class == 1:
                    ds.address = self._u(off + 2, 8)
                    ds.size = self._u(off + 10, 8)
                elif ds.layout_class == 2:
                    rank1 = b[off + 2]
                    ds.address = self._u(off + 3, 8)
                    ds.chunk_dims = tuple(self._u(off + 11 + 4 * i, 4) for i in range(rank1))
                else:
                    raise HDF5FormatError(f"layout class {ds.layout_class}")
            elif mtype == 0x0B:  # filter pipeline
                ver = b[off]
                nf = b[off + 1]
                p = off + (8 if ver == 1 else 2)
                for _ in range(nf):
                    fid = self._u(p, 2)
                    if ver == 1 or fid >= 256:
                        nlen = self._u(p + 2, 2)
                    else:
                        nlen = 0
                    ncd = self._u(p + 6, 2)
                    p += 8
                    if nlen:
                        p += (nlen + 7) // 8 * 8
                    p += 4 * ncd
                    if ver == 1 and ncd % 2:
                        p += 4
                    ds.filters.append(fid)
        return ds

    def _unfilter(self, raw: bytes, filters: list[int], mask: int, itemsize: int) -> bytes:
        for i, fid in reversed(list(enumerate(filters))):
            if mask & (1 << i):
                continue
            if fid == 1:
                raw = zlib.decompress(raw)
            elif fid == 2:  # shuffle
                arr = np.frombuffer(raw, dtype=np.uint8)
                n = arr.size // itemsize
                raw = arr[: n * itemsize].reshape(itemsize, n).T.tobytes() + raw[n * itemsize :]
            else:
                raise HDF5FormatError(f"unsupported filter id {fid}")
        return raw

    def _read_chunks(self, ds: _Dataset, out: np.ndarray, addr: int) -> None:
        b = self.buf
        if b[addr : addr + 4] != b"TREE":
            raise HDF5FormatError("bad chunk B-tree signature")
        level = b[addr + 5]
        nused = self._u(addr + 6, 2)
        rank1 = len(ds.chunk_dims)
        key_size = 8 + 8 * rank1
        off = addr + 24
        cdims = ds.chunk_dims[:-1]
        for _ in range(nused):
            csize = self._u(off, 4)
            fmask = self._u(off + 4, 4)
            offsets = [self._u(off + 8 + 8 * i, 8) for i in range(rank1 - 1)]
            child = self._u(off + key_size, 8)
            off += key_size + 8
            if level > 0:
                self._read_chunks(ds, out, child)
                continue
            raw = self._unfilter(b[child : child + csize], ds.filters, fmask, ds.dtype.itemsize)
            chunk = np.frombuffer(raw, dtype=ds.dtype, count=int(np.prod(cdims))).reshape(cdims)
            sl = tuple(
                slice(o, min(o + c, s)) for o, c, s in zip(offsets, cdims, ds.shape)
            )
            sub = tuple(slice(0, s.stop - s.start) for s in sl)
            out[sl] = chunk[sub]

    def read_dataset(self, header_addr: int) -> np.ndarray:
        ds = self._dataset(header_addr)
        if ds.dtype is None:
            raise HDF5FormatError("dataset without datatype")
        count = int(np.prod(ds.shape)) if ds.shape else 1
        if ds.layout_class == 0:
            arr = np.frombuffer(ds.compact, dtype=ds.dtype, count=count)
        elif ds.layout_class == 1:
            if ds.address == _UNDEF:
                arr = np.zeros(count, dtype=ds.dtype)
            else:
                arr = np.frombuffer(self.buf, dtype=ds.dtype, count=count, offset=ds.address)
        else:
            out = np.zeros(ds.shape, dtype=ds.dtype)
            if ds.address != _UNDEF:
                self._read_chunks(ds, out, ds.address)
            arr = out
        arr = np.array(arr, dtype=ds.dtype.newbyteorder("="), copy=True).reshape(ds.shape)
        return arr

    # --- public API --------------------------------------------------------------------
    def keys(self, header_addr: int | None = None) -> dict[str, int]:
        return self._group_entries(self.root_header if header_addr is None else header_addr)

    def is_group(self, header_addr: int) -> bool:
        return any(m == 0x11 for m, _, _ in self._messages(header_addr))


def load_problem_hdf5(filename: str):
    """Load a problem HDF5 file: ``(params, adjacency)`` like the reference's loader."""
    import scipy.sparse as sp

    f = HDF5File(filename)
    params: dict[str, object] = {}
    adjacency = None
    for name, addr in f.keys().items():
        if name == "adjacency" and f.is_group(addr):
            sub = f.keys(addr)
            data = f.read_dataset(sub["data"])
            row = f.read_dataset(sub["row"])
            col = f.read_dataset(sub["col"])
            shape = tuple(int(v) for v in f.read_dataset(sub["shape"]))
            adjacency = sp.coo_matrix((data, (row, col)), shape=shape)
            continue
        arr = f.read_dataset(addr)
        params[name] = arr[()] if arr.shape == () else arr
    return params, adjacency
