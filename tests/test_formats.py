"""The tile wire format of include/tile_holder.cuh (header {nnz, buf_size} of
two index-type integers, :121, then the contiguous [val | idx | ptr] buffer,
:80-89): host round trips for both index widths, the int32 header limit
checked without building a 2 GiB buffer, malformed messages rejected, and
(on the GPU) the library's own contiguous device tile is byte-identical to
the wire buffer."""
import ctypes

import numpy as np
import pytest

from trident_amd import formats, gen


def _same(X, Y):
    return (X.nrows == Y.nrows and np.array_equal(X.ptr, Y.ptr) and np.array_equal(X.idx, Y.idx)
            and np.array_equal(X.val.view(np.uint64), Y.val.view(np.uint64)))


@pytest.mark.parametrize("index_dtype", [np.int32, np.int64])
def test_wire_round_trip(index_dtype):
    m = gen.rmat(8, 8, seed=3)
    w = formats.tile_to_wire(m, index_dtype=index_dtype)
    isz = np.dtype(index_dtype).itemsize
    nnz, size = np.frombuffer(w[:2 * isz], dtype=index_dtype)
    assert nnz == m.nnz and size == m.nnz * (8 + isz) + (m.nrows + 1) * isz == len(w) - 2 * isz
    # the buffer is [val | idx | ptr] in that order
    assert w[2 * isz:2 * isz + 8 * m.nnz] == m.val.astype(np.float64).tobytes()
    assert w[-(m.nrows + 1) * isz:] == m.ptr.astype(index_dtype).tobytes()
    assert _same(formats.tile_from_wire(w, m.nrows, m.ncols, index_dtype=index_dtype), m)


def test_wire_empty_tile():
    z = gen.from_arrays(5, 7, np.zeros(6, np.int64), np.zeros(0, np.int32), np.zeros(0))
    w = formats.tile_to_wire(z)
    assert len(w) == 8 + 6 * 4
    assert _same(formats.tile_from_wire(w, 5, 7), z)


def test_wire_header_limits():
    assert np.frombuffer(formats.wire_header(3, 2**31 - 1), np.int32).tolist() == [3, 2**31 - 1]
    with pytest.raises(OverflowError):
        formats.wire_header(3, 2**31)  # an int32 tile's buffer above 2 GiB - 1 has no header
    with pytest.raises(OverflowError):
        formats.wire_header(2**31, 16)
    assert np.frombuffer(formats.wire_header(2**31, 2**40, np.int64), np.int64).tolist() == [2**31, 2**40]


def test_wire_rejects_malformed():
    m = gen.rmat(6, 4, seed=1)
    w = formats.tile_to_wire(m)
    with pytest.raises(ValueError):
        formats.tile_from_wire(w[:-4], m.nrows, m.ncols)       # truncated
    with pytest.raises(ValueError):
        formats.tile_from_wire(w, m.nrows + 1, m.ncols)        # wrong row count


@pytest.mark.gpu
def test_device_tile_buffer_is_the_wire_buffer(ctx):
    m = gen.rmat(10, 8, seed=2)
    d = ctx.upload(m)
    assert d.d.contig == 1 and d.d.buf
    w = formats.tile_to_wire(m)[8:]
    host = np.empty(len(w), np.uint8)
    ctx.check(ctx.lib.tr_memcpy_d2h(ctx.h, host.ctypes.data, ctypes.c_void_p(d.d.buf), len(w)))
    assert host.tobytes() == w
    assert _same(formats.tile_from_wire(formats.wire_header(m.nnz, len(w)) + host.tobytes(), m.nrows, m.ncols), m)
