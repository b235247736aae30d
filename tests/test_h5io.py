"""vector_dist save / load in the reference's HDF5 checkpoint layout (CPU).

The reference's own checkpoint (test_data/vector_dist_24.h5, written by 24
processors, committed as tests/golden/vector_dist_24.h5) loads with what
vector_dist_hdf5_load_test asserts (src/Vector/tests/
vector_dist_HDF5_chckpnt_restart_test.cpp:119-176): 32^3 particles, property
0 == (float)(0.51234 + x + y + z); the positions are the 32^3 grid of spacing
1/31 the save test placed them on (:48-65).  The writer is pinned against
the same file byte for byte: saving its particles gives one block whose bytes
are the file's 24 blocks' position parts, then their property parts, each
after its count.  Round trips keep every bit, padding included."""
import ctypes as C
import os

import numpy as np
import pytest

from openfpm_pdata_amd import _lib as L
from openfpm_pdata_amd import h5io

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "vector_dist_24.h5")


def raw_datasets(path):
    """(metadata int32[], vector_dist bytes) read with the HDF5 C library
    directly, independent of libofp_io."""
    h5 = C.CDLL("/opt/conda/lib/libhdf5.so")
    h5.H5open()
    hid = C.c_int64
    h5.H5Fopen.restype = hid
    h5.H5Fopen.argtypes = [C.c_char_p, C.c_uint, hid]
    h5.H5Dopen2.restype = hid
    h5.H5Dopen2.argtypes = [hid, C.c_char_p, hid]
    h5.H5Dget_space.restype = hid
    h5.H5Dget_space.argtypes = [hid]
    h5.H5Sget_simple_extent_dims.argtypes = [hid, C.c_void_p, C.c_void_p]
    h5.H5Dread.argtypes = [hid, hid, hid, hid, hid, C.c_void_p]
    for h in ("H5Dclose", "H5Sclose", "H5Fclose"):
        getattr(h5, h).argtypes = [hid]
    f = h5.H5Fopen(os.fsencode(path), 0, 0)
    out = []
    for name, tname, npd in ((b"metadata", "H5T_NATIVE_INT_g", np.int32), (b"vector_dist", "H5T_NATIVE_SCHAR_g",
                                                                          np.int8)):
        d = h5.H5Dopen2(f, name, 0)
        sp = h5.H5Dget_space(d)
        dim = C.c_uint64(0)
        h5.H5Sget_simple_extent_dims(sp, C.byref(dim), None)
        a = np.zeros(dim.value, npd)
        assert h5.H5Dread(d, hid.in_dll(h5, tname).value, 0, 0, 0, a.ctypes.data) >= 0
        h5.H5Sclose(sp)
        h5.H5Dclose(d)
        out.append(a)
    h5.H5Fclose(f)
    return out


def test_load_reference_checkpoint():
    pos, (p0, p1, p2) = h5io.load(GOLDEN, np.float32, [(np.float32, 3)])
    assert pos.shape == (3, 32 ** 3)
    x, y, z = pos.astype(np.float64)
    want = (((0.51234 + x) + y) + z).astype(np.float32)
    for p in (p0, p1, p2):
        np.testing.assert_array_equal(p, want)
    sp = np.float32(1.0) / np.float32(31)
    grid = np.arange(32, dtype=np.float32) * sp
    for k in range(3):
        np.testing.assert_array_equal(np.unique(pos[k]), grid)
    assert len(np.unique(pos.T, axis=0)) == 32 ** 3


def test_writer_layout_matches_reference_file(tmp_path):
    pos, props = h5io.load(GOLDEN, np.float32, [(np.float32, 3)])
    out = str(tmp_path / "one_block.h5")
    h5io.save(out, pos, props, [(np.float32, 3)])
    meta_ref, bytes_ref = raw_datasets(GOLDEN)
    meta, data = raw_datasets(out)
    assert meta.tolist() == [len(data)]
    n = pos.shape[1]
    # the reference's blocks: [u64 n][n x 12 B positions][u64 n][n x 12 B objects]
    pos_parts, obj_parts, at = [], [], 0
    for b in meta_ref:
        blk = bytes_ref[at:at + b].tobytes()
        m = int.from_bytes(blk[:8], "little")
        pos_parts.append(blk[8:8 + 12 * m])
        obj_parts.append(blk[16 + 12 * m:16 + 24 * m])
        at += b
    cnt = n.to_bytes(8, "little")
    assert data.tobytes() == cnt + b"".join(pos_parts) + cnt + b"".join(obj_parts)


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_save_load_round_trip(tmp_path, dt):
    """Mixed members (float, double[2], int32: 4 bytes of padding after the
    float, object of 32 bytes) keep every bit; an empty vector round-trips."""
    rng = np.random.default_rng(3)
    props = [(np.float32, 1), (np.float64, 2), (np.int32, 1)]
    planes, obj = h5io.layout(props)
    assert [o for o, _ in planes] == [0, 8, 16, 24] and obj == 32
    for n in (0, 1, 5000):
        pos = rng.random((3, n)).astype(dt)
        arrs = [rng.random(n).astype(np.float32), rng.random(n), rng.random(n) * 1e300,
                rng.integers(-2 ** 31, 2 ** 31, n).astype(np.int32)]
        f = str(tmp_path / f"rt_{n}.h5")
        h5io.save(f, pos, arrs, props)
        pos2, arrs2 = h5io.load(f, dt, props)
        np.testing.assert_array_equal(pos2, pos)
        for a, b in zip(arrs, arrs2):
            assert a.dtype == b.dtype
            np.testing.assert_array_equal(a, b)


def test_layout_mismatch_is_an_error(tmp_path):
    with pytest.raises(L.OfpError, match="layout"):
        h5io.load(GOLDEN, np.float64, [(np.float32, 3)])
    with pytest.raises(L.OfpError, match="layout"):
        h5io.load(GOLDEN, np.float32, [(np.float32, 2)])
    with pytest.raises(L.OfpError, match="cannot open"):
        h5io.load(str(tmp_path / "missing.h5"), np.float32, [(np.float32, 3)])
