"""The JLD2 writer (ripple_amd/jld2.py) against the format's published
facts: Bob Jenkins' lookup3 known answers (the HDF5 checksum), the JLD2 text
header and the superblock at byte 512, and files read back by the HDF5
project's h5ls/h5dump (an independent reader): groups, dimensions (Julia's
reversed), element types and values (CPU only)."""
import os

import numpy as np
import pytest

from ripple_amd import jld2
import h5dump_reader as h5


def test_lookup3_known_answers():
    # the test driver of Jenkins' lookup3.c
    assert jld2.lookup3(b"") == 0xDEADBEEF
    assert jld2.lookup3(b"Four score and seven years ago") == 0x17770551
    assert jld2.lookup3(b"Four score and seven years ago", 1) == 0xCD628161


def test_header_and_superblock(tmp_path):
    fn = str(tmp_path / "h.jld2")
    with jld2.JLD2File(fn) as f:
        f["x"] = np.arange(3.0)
    raw = open(fn, "rb").read()
    assert raw.startswith(b"HDF5-based Julia Data Format, version ")
    assert raw[512:520] == b"\x89HDF\r\n\x1a\n" and raw[520] == 2  # superblock version 2 at the base
    sb = raw[512:512 + 44]
    assert int.from_bytes(raw[556:560], "little") == jld2.lookup3(sb)
    assert int.from_bytes(raw[512 + 12:512 + 20], "little") == 512  # base address
    assert int.from_bytes(raw[512 + 28:512 + 36], "little") == len(raw)  # end of file


@pytest.mark.skipif(not h5.available(), reason="h5ls/h5dump not installed")
def test_read_back_by_hdf5_tools(tmp_path):
    fn = str(tmp_path / "t.jld2")
    rng = np.random.default_rng(0)
    a = np.asfortranarray(rng.standard_normal((5, 4, 3, 2)))
    b = rng.standard_normal((7,)).astype(np.float32)
    with jld2.JLD2File(fn) as f:
        f["timeseries/N/0"] = a
        f["timeseries/t/0"] = 0.0
        f.flush()  # a group gains links after a flush
        f["timeseries/N/10"] = 2 * a
        f["timeseries/t/10"] = 0.25
        f["grid/Nx"] = np.int64(5)
        f["b"] = b
        f["empty"] = np.zeros((0, 3))
    ls = h5.listing(fn)
    assert ls["/timeseries"] == "Group" and ls["/timeseries/N"] == "Group"
    assert ls["/timeseries/N/0"] == (2, 3, 4, 5)  # Julia's (5, 4, 3, 2), reversed
    assert ls["/timeseries/t/10"] == () and ls["/grid/Nx"] == ()
    assert np.array_equal(h5.julia_array(fn, "/timeseries/N/0"), a)
    assert np.array_equal(h5.julia_array(fn, "/timeseries/N/10"), 2 * a)
    assert h5.julia_array(fn, "/timeseries/t/10") == 0.25
    vals, dims, dtype = h5.dataset(fn, "/b")
    assert dtype == "H5T_IEEE_F32LE" and dims == (7,) and np.array_equal(vals.astype(np.float32), b)
    vals, dims, dtype = h5.dataset(fn, "/grid/Nx")
    assert dtype == "H5T_STD_I64LE" and vals[0] == 5
    assert ls["/empty"] == (3, 0)


def test_refuses_to_overwrite(tmp_path):
    fn = str(tmp_path / "o.jld2")
    jld2.JLD2File(fn).close()
    with pytest.raises(FileExistsError):
        jld2.JLD2File(fn, overwrite_existing=False)
    with jld2.JLD2File(fn) as f:
        f["a"] = 1.0
        with pytest.raises(ValueError):
            f["a"] = 2.0
        with pytest.raises(ValueError):
            f["a/b"] = 2.0
        with pytest.raises(TypeError):
            f["s"] = np.array(["x"])
