"""The MAT v5 reader (opencliper_amd/cpp/MatFile.cpp) parses untrusted files: build it with AddressSanitizer and
UndefinedBehaviorSanitizer (host code, g++) and feed it truncated, bit-flipped and length-corrupted variants of valid
files.  Every run must end with the parsed variables or a clean error (exit 0 or 1), never a sanitizer report."""
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BUILD = os.path.join(ROOT, "_build")
TOOL = os.path.join(BUILD, "mat_tool_san")

pytestmark = pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++ with the sanitizer runtimes")


@pytest.fixture(scope="module")
def tool():
    os.makedirs(BUILD, exist_ok=True)
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
           "-fno-omit-frame-pointer", "-I" + os.path.join(ROOT, "include"), "-o", TOOL,
           os.path.join(ROOT, "opencliper_amd", "examples", "mat_tool.cpp"),
           os.path.join(ROOT, "opencliper_amd", "cpp", "MatFile.cpp"), "-lz"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return TOOL


def _valid_files(tmp):
    import scipy.io
    rng = np.random.default_rng(7)
    vars_ = {"Dims": np.array([[2, 1, 3, 1, 4, 6, 5]], dtype=np.uint32),
             "KData": (rng.normal(size=(6, 5, 3, 4)) + 1j * rng.normal(size=(6, 5, 3, 4))).astype(np.complex64),
             "SamplingMasks": (rng.random((6, 5, 4)) < 0.5).astype(np.uint8),
             "d": rng.normal(size=(7,)), "i16": np.arange(-5, 5, dtype=np.int16), "s": np.float32(3.5)}
    out = []
    for comp in (False, True):
        p = os.path.join(tmp, f"valid_{int(comp)}.mat")
        scipy.io.savemat(p, vars_, do_compression=comp)
        out.append(p)
    return out


def _variants(data, rng, n):
    """Truncations, single-bit flips and 32-bit fields overwritten with extreme lengths."""
    L = len(data)
    cuts = sorted(set([0, 1, 127, 128, 129, 135, 136, L - 1] + list(rng.integers(1, L, n))))
    for c in cuts:
        yield f"cut{c}", data[:c]
    for k in range(n):
        b = bytearray(data)
        i = int(rng.integers(128, L))
        b[i] ^= 1 << int(rng.integers(0, 8))
        yield f"flip{i}", bytes(b)
    for k in range(n):
        b = bytearray(data)
        i = int(rng.integers(128, max(129, L - 4))) & ~3
        v = [0xFFFFFFFF, 0x7FFFFFFF, 0x80000000, 0xFFFFFFF8, 0x00010000][k % 5]
        b[i:i + 4] = int(v).to_bytes(4, "little")
        yield f"len{i}", bytes(b)


def test_malformed_files_end_cleanly_under_sanitizers(tool, tmp_path):
    rng = np.random.default_rng(11)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    runs = errors = 0
    for src in _valid_files(str(tmp_path)):
        r = subprocess.run([tool, "dump", src], capture_output=True, text=True, env=env, timeout=60)
        assert r.returncode == 0 and '"name": "KData"' in r.stdout, r.stderr
        data = open(src, "rb").read()
        for tag, blob in _variants(data, rng, 40):
            p = str(tmp_path / f"{tag}.mat")
            with open(p, "wb") as fh:
                fh.write(blob)
            r = subprocess.run([tool, "dump", p], capture_output=True, text=True, env=env, timeout=60)
            runs += 1
            errors += r.returncode == 1
            assert "AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr, (tag, r.stderr[-2000:])
            assert r.returncode in (0, 1), (tag, r.returncode, r.stderr[-2000:])
    print(f"{runs} malformed files, {errors} refused with an error", flush=True)
    assert errors > runs // 4              # most corruptions are caught as format errors, not silently accepted
