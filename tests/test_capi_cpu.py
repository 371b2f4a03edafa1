"""CPU-side checks of the C ABI library: it loads, exports every symbol the header declares, and refuses to run
without a GPU instead of falling back to anything."""
import ctypes
import os
import subprocess

import pytest

from opencliper_amd import capi


def test_header_declares_entry_points():
    syms = capi.header_symbols()
    for s in ("ocl_ctx_create", "ocl_ttv", "ocl_fft2d", "ocl_nesta_run", "ocl_batch_run", "ocl_normest"):
        assert s in syms
    assert len(syms) >= 36


def test_library_exports_every_header_symbol():
    L = ctypes.CDLL(capi.LIB_PATH)
    missing = [s for s in capi.header_symbols() if not hasattr(L, s)]
    assert not missing, missing


def test_nm_shows_gfx950_code_object():
    out = subprocess.run(["nm", "-D", "--defined-only", capi.LIB_PATH], capture_output=True, text=True).stdout
    assert "ocl_nesta_run" in out
    blob = open(capi.LIB_PATH, "rb").read()
    assert b"gfx950" in blob


def test_null_arguments_are_rejected_without_gpu():
    L = capi.lib()
    assert L.ocl_ctx_sync(None) == -1
    assert b"null context" in L.ocl_last_error()
    assert L.ocl_ttv(None, None, None, 1, 1, 1) == -1
    assert L.ocl_batch_setup(None) == -1
    assert L.ocl_host_malloc(None, 16, None) == -1
    assert L.ocl_host_free(None, None) == 0          # no context needed; nothing to free
    assert L.ocl_ctx_set_group_threshold(None, 4) == -1


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="only meaningful without a GPU")
def test_context_creation_fails_loudly_without_gpu():
    with pytest.raises(capi.OclError):
        capi.Context(0)


def test_work_buffers_allocate_only_through_dev_malloc():
    """ocl_nesta_run_noalloc relies on every work-buffer allocation of the library going through dev_malloc (which
    carves from the static scratch during such a call): hipMalloc may appear only in dev_malloc itself, in context
    creation, in ocl_malloc and in the cached FFT twiddle tables (which that call bypasses)."""
    import re
    csrc = os.path.join(os.path.dirname(capi.LIB_PATH), "csrc")
    allowed = {("ocl_batch.cpp", "dev_malloc"), ("ocl_capi.cpp", "ocl_ctx_create"), ("ocl_capi.cpp", "ocl_malloc"),
               ("ocl_capi.cpp", "make_plan")}
    found = set()
    for fn in sorted(os.listdir(csrc)):
        if not fn.endswith((".cpp", ".hip", ".hpp")):
            continue
        func = None
        for line in open(os.path.join(csrc, fn)):
            m = re.match(r"^(?:extern \"C\" )?(?:static )?[\w:<>*\s]+?\b(\w+)\(", line)
            if m and not line.startswith((" ", "\t", "#", "//")):
                func = m.group(1)
            if re.search(r"\bhipMalloc\w*\(", line) and not line.lstrip().startswith("//"):
                found.add((fn, func))
    assert found <= allowed, found - allowed
