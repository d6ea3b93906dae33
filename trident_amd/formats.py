"""On-disk and wire formats of the reference, read and written on the host.

* `.bmtx` — the binary MatrixMarket files the reference ships under
  small_matrices/ (read there by the un-vendored dmmio).  Layout, inferred
  from the bytes: line 1 is a MatrixMarket banner ending in two integers
  `<index_bytes> <value_bytes>`, line 2 is `nrows ncols nnz`, then nnz packed
  little-endian records {row: index_bytes, col: index_bytes, val:
  value_bytes}, 0-based.
* tile wire format — include/tile_holder.cuh:74-89,120-127: a header
  {nnz, buf_size} of two index-type integers followed by one contiguous
  buffer [val[nnz] | idx[nnz] | ptr[ptrdim+1]].
"""
from __future__ import annotations

import numpy as np

from .gen import Csr

_INT = {1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64}
_FLT = {4: np.float32, 8: np.float64}


def read_bmtx(path: str) -> Csr:
    """Read a .bmtx file into a host CSR (rows sorted, fp64 values)."""
    with open(path, "rb") as f:
        data = f.read()
    e1 = data.index(b"\n")
    e2 = data.index(b"\n", e1 + 1)
    banner = data[:e1].decode("ascii").split()
    if not banner or not banner[0].startswith("%%MatrixMarket"):
        raise ValueError(f"{path}: not a MatrixMarket banner")
    ib, vb = int(banner[-2]), int(banner[-1])
    nrows, ncols, nnz = (int(x) for x in data[e1 + 1:e2].split())
    rec = np.dtype([("r", _INT[ib]), ("c", _INT[ib]), ("v", _FLT[vb])])
    body = data[e2 + 1:]
    if len(body) != nnz * rec.itemsize:
        raise ValueError(f"{path}: expected {nnz * rec.itemsize} payload bytes, found {len(body)}")
    arr = np.frombuffer(body, dtype=rec, count=nnz)
    rows = arr["r"].astype(np.int64)
    cols = arr["c"].astype(np.int64)
    vals = arr["v"].astype(np.float64)
    order = np.lexsort((cols, rows))
    rows, cols, vals = rows[order], cols[order], vals[order]
    ptr = np.zeros(nrows + 1, np.int64)
    np.cumsum(np.bincount(rows, minlength=nrows), out=ptr[1:])
    return Csr(nrows, ncols, ptr.astype(np.int32), cols.astype(np.int32), vals)


def write_bmtx(path: str, m: Csr, index_bytes: int = 4, value_bytes: int = 4) -> None:
    rec = np.dtype([("r", _INT[index_bytes]), ("c", _INT[index_bytes]), ("v", _FLT[value_bytes])])
    arr = np.zeros(m.nnz, dtype=rec)
    arr["r"] = m.row_ids()
    arr["c"] = m.idx
    arr["v"] = m.val
    with open(path, "wb") as f:
        f.write(f"%%MatrixMarket matrix coordinate real general {index_bytes} {value_bytes}\n".encode())
        f.write(f"{m.nrows} {m.ncols} {m.nnz}\n".encode())
        f.write(arr.tobytes())


def wire_header(nnz: int, buf_size: int, index_dtype=np.int32) -> bytes:
    """The tile's size message {nnz, buf_size}: two integers of the index type,
    as the reference sends them (`IT payload[2]`, include/tile_holder.cuh:121).
    Raises OverflowError when either value does not fit the index type (an
    int32 tile's buffer is limited to 2^31 - 1 bytes)."""
    info = np.iinfo(index_dtype)
    for name, v in (("nnz", nnz), ("buf_size", buf_size)):
        if v < 0 or v > info.max:
            raise OverflowError(f"wire tile: {name} = {v} does not fit {np.dtype(index_dtype).name}")
    return np.array([nnz, buf_size], dtype=index_dtype).tobytes()


def tile_to_wire(m: Csr, index_dtype=np.int32, value_dtype=np.float64) -> bytes:
    """Serialise a tile as header {nnz, buf_size} + [val | idx | ptr]
    (include/tile_holder.cuh:80-89 buffer order, :121 header)."""
    val = np.ascontiguousarray(m.val, value_dtype).tobytes()
    idx = np.ascontiguousarray(m.idx, index_dtype).tobytes()
    ptr = np.ascontiguousarray(m.ptr, index_dtype).tobytes()
    buf = val + idx + ptr
    return wire_header(m.nnz, len(buf), index_dtype) + buf


def tile_from_wire(data: bytes, nrows: int, ncols: int, index_dtype=np.int32, value_dtype=np.float64) -> Csr:
    """Inverse of tile_to_wire (same index and value types on both sides)."""
    isz = np.dtype(index_dtype).itemsize
    vsz = np.dtype(value_dtype).itemsize
    nnz, buf_size = (int(x) for x in np.frombuffer(data[:2 * isz], dtype=index_dtype))
    if len(data) < 2 * isz + buf_size:
        raise ValueError("wire tile: message shorter than its header says")
    if buf_size != nnz * (vsz + isz) + (nrows + 1) * isz:
        raise ValueError("wire tile: buffer size does not match nnz and nrows")
    buf = data[2 * isz:2 * isz + buf_size]
    val = np.frombuffer(buf[:nnz * vsz], dtype=value_dtype).astype(np.float64)
    idx = np.frombuffer(buf[nnz * vsz:nnz * (vsz + isz)], dtype=index_dtype).astype(np.int32)
    ptr = np.frombuffer(buf[nnz * (vsz + isz):], dtype=index_dtype).astype(np.int32)
    return Csr(nrows, ncols, ptr, idx, val)
