"""HBM byte-counter calibration for rocprofv3 --pmc runs.

Runs the library's device copy kernel (k_copy) over 256 MiB five times on the context's stream: each dispatch reads
256 MiB and writes 256 MiB, so FETCH_SIZE / WRITE_SIZE (KiB per dispatch) of these dispatches give the factors that
turn the same counters of any other kernel of this library into bytes.

    rocprofv3 --pmc FETCH_SIZE --output-format csv -d <dir> -- python3 opencliper_amd/calibrate.py
"""
from __future__ import annotations

import os
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
from opencliper_amd import capi  # noqa: E402

NBYTES = 256 << 20


def main():
    ctx = capi.Context(0)
    try:
        n = NBYTES // 8
        a = ctx.zeros((n,), np.complex64)
        b = ctx.empty((n,), np.complex64)
        for _ in range(5):
            ctx.copy(a, b)
        ctx.sync()
        print(f"copy kernel: {NBYTES} bytes read + {NBYTES} bytes written per dispatch, 5 dispatches")
    finally:
        ctx.close()


if __name__ == "__main__":
    main()
