"""The server's block cache (opencliper_amd/csrc/ocl_bufcache.hpp), host-only: tests/bufcache_check.cpp drives it
over a fake device with seeded random admissions and completions, built with AddressSanitizer and
UndefinedBehaviorSanitizer, and checks its accounting after every step (live blocks = held, held + cached within the
budget, the device holding exactly what the cache accounts, reuse within 4/3 of the request)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BUILD = os.path.join(ROOT, "_build")
EXE = os.path.join(BUILD, "bufcache_check")

pytestmark = pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++ with the sanitizer runtimes")


@pytest.fixture(scope="module")
def exe():
    os.makedirs(BUILD, exist_ok=True)
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
           "-fno-omit-frame-pointer", "-I" + os.path.join(ROOT, "opencliper_amd", "csrc"), "-o", EXE,
           os.path.join(ROOT, "tests", "bufcache_check.cpp")]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return EXE


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_block_cache_accounting(exe, seed):
    r = subprocess.run([exe, "20000", str(seed)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.startswith("ok")
