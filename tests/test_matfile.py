"""MAT v5 file I/O (OpenCLIPER::mat, include/OpenCLIPER/MatFile.hpp): the role libmatio plays for the reference's
Data::readMatlabVariablesFromFile and matlabSave (src/Data.cpp:1392-1486).  Checked against scipy.io (an independent
MAT v5 implementation) in both directions, compressed and uncompressed, and against the two libmatio-written files
the reference ships (recon_fullySampled.mat, recon_halfScan.mat).  Host-only: no device needed."""
import json
import os
import subprocess

import numpy as np
import pytest
import scipy.io as sio

PKG = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "opencliper_amd")
TOOL = os.path.join(PKG, "examples", "mat_tool")
REF_FILES = ["/root/reference/recon_fullySampled.mat", "/root/reference/recon_halfScan.mat"]


def _dump(path):
    r = subprocess.run([TOOL, "dump", path], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    return {d["name"]: d for d in (json.loads(l) for l in r.stdout.splitlines())}


def _copy(src, dst, compress):
    r = subprocess.run([TOOL, "copy", src, dst, str(int(compress))], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr


def _user_vars(d):
    return {k: v for k, v in d.items() if not k.startswith("__")}


def _variables(rng):
    """One variable per numeric class, real and complex, vectors, matrices, 4-D, scalars and empty-free small cases
    (a one-character name and a one-byte scalar are stored as 4-byte small data elements)."""
    return {
        "a": np.uint8(7) * np.ones((1, 1), np.uint8),
        "dbl": rng.standard_normal((5, 3)),
        "sgl": rng.standard_normal((4, 6)).astype(np.float32),
        "cplx64": (rng.standard_normal((3, 4, 5)) + 1j * rng.standard_normal((3, 4, 5))).astype(np.complex64),
        "cplx128": rng.standard_normal((2, 7)) + 1j * rng.standard_normal((2, 7)),
        "u8": rng.integers(0, 256, (9, 2), dtype=np.uint8),
        "i8": rng.integers(-128, 128, (3, 3), dtype=np.int8),
        "u16": rng.integers(0, 65536, (2, 5), dtype=np.uint16),
        "i16": rng.integers(-32768, 32768, (5, 2), dtype=np.int16),
        "u32": rng.integers(0, 2**32, (1, 11), dtype=np.uint32),
        "i32": rng.integers(-2**31, 2**31, (4, 1), dtype=np.int32),
        "i64": rng.integers(-2**52, 2**52, (2, 2), dtype=np.int64),
        "u64": rng.integers(0, 2**52, (3, 1), dtype=np.uint64),
        "four_d": rng.standard_normal((3, 2, 4, 5)).astype(np.float32),
    }


def test_tool_usage_error():
    r = subprocess.run([TOOL], capture_output=True, text=True)
    assert r.returncode == 2 and "usage" in r.stderr


@pytest.mark.parametrize("scipy_compress", [False, True])
@pytest.mark.parametrize("our_compress", [False, True])
def test_copy_through_reader_and_writer_matches_scipy(tmp_path, scipy_compress, our_compress):
    vars_in = _variables(np.random.default_rng(11))
    src, dst = str(tmp_path / "src.mat"), str(tmp_path / "dst.mat")
    sio.savemat(src, vars_in, do_compression=scipy_compress)
    _copy(src, dst, our_compress)
    back = _user_vars(sio.loadmat(dst))
    assert set(back) == set(vars_in)
    for k, v in vars_in.items():
        assert back[k].dtype == v.dtype, k
        assert back[k].shape == v.shape, k
        assert np.array_equal(back[k], v), k


def test_dump_matches_scipy(tmp_path):
    vars_in = _variables(np.random.default_rng(12))
    src = str(tmp_path / "src.mat")
    sio.savemat(src, vars_in, do_compression=True)
    d = _dump(src)
    classes = {np.float64: 6, np.float32: 7, np.int8: 8, np.uint8: 9, np.int16: 10, np.uint16: 11, np.int32: 12,
               np.uint32: 13, np.int64: 14, np.uint64: 15, np.complex64: 7, np.complex128: 6}
    for k, v in vars_in.items():
        assert d[k]["dims"] == list(v.shape), k
        assert d[k]["class"] == classes[v.dtype.type], k
        assert d[k]["complex"] == np.iscomplexobj(v), k
        assert d[k]["sum_re"] == pytest.approx(float(np.sum(v.real.astype(np.float64))), rel=1e-12, abs=1e-9)


def test_non_numeric_variables_are_skipped(tmp_path):
    src = str(tmp_path / "mixed.mat")
    sio.savemat(src, {"txt": "hello", "cell": np.array([[np.zeros(2), "x"]], dtype=object),
                      "st": {"f": np.ones(3)}, "num": np.arange(6.0).reshape(2, 3)}, do_compression=False)
    d = _dump(src)
    assert set(d) == {"num"} and d["num"]["dims"] == [2, 3] and d["num"]["sum_re"] == 15.0


def test_rejects_files_that_are_not_mat_v5(tmp_path):
    bad = tmp_path / "bad.mat"
    bad.write_bytes(b"not a mat file" * 20)
    r = subprocess.run([TOOL, "dump", str(bad)], capture_output=True, text=True)
    assert r.returncode == 1 and "not a little-endian MAT v5 file" in r.stderr
    r = subprocess.run([TOOL, "dump", str(tmp_path / "missing.mat")], capture_output=True, text=True)
    assert r.returncode == 1 and "cannot open" in r.stderr


@pytest.mark.parametrize("compress", [False, True])
def test_truncated_file_is_an_error_not_a_crash(tmp_path, compress):
    src = str(tmp_path / "src.mat")
    sio.savemat(src, {"big": np.random.default_rng(3).standard_normal((64, 64))}, do_compression=compress)
    raw = open(src, "rb").read()
    cut = tmp_path / "cut.mat"
    cut.write_bytes(raw[: len(raw) // 2])
    r = subprocess.run([TOOL, "dump", str(cut)], capture_output=True, text=True)
    assert r.returncode == 1 and "MAT-file" in r.stderr


@pytest.mark.skipif(not all(os.path.exists(f) for f in REF_FILES), reason="reference checkout not present")
@pytest.mark.parametrize("path", REF_FILES)
def test_reads_libmatio_files_of_the_reference(tmp_path, path):
    """The reference's own XData outputs (written by libmatio with zlib compression): Dims and XData read back as
    scipy reads them, and survive a copy through the writer bit-exactly."""
    ref = _user_vars(sio.loadmat(path))
    d = _dump(path)
    assert set(d) == {"Dims", "XData"}
    assert d["Dims"]["class"] == 13 and d["XData"]["class"] == 7 and d["XData"]["complex"]
    assert d["XData"]["dims"] == list(ref["XData"].shape)
    dims = ref["Dims"].ravel()
    # [nSpatial, allSizesEqual, nCoils, nTemporal, temporal..., spatial...] with nCoils = 0 for XData
    assert list(dims[:4]) == [2, 1, 0, 1] and dims[4] == ref["XData"].shape[2]
    for key in ("sum_re", "sum_im"):
        part = ref["XData"].real if key == "sum_re" else ref["XData"].imag
        assert d["XData"][key] == pytest.approx(float(np.sum(part.astype(np.float64))), rel=1e-12, abs=1e-6)
    dst = str(tmp_path / "copy.mat")
    _copy(path, dst, True)
    back = _user_vars(sio.loadmat(dst))
    for k in ref:
        assert back[k].dtype == ref[k].dtype and np.array_equal(back[k], ref[k]), k
