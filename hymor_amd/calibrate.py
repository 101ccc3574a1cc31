"""Streaming calibration: achievable HBM bandwidth on this GPU (a copy, and a
read-only stream), and the reference kernel for scaling the FETCH_SIZE /
WRITE_SIZE profiler counters.

``python hymor_amd/calibrate.py [MiB] [reps]`` (or ``python -m
hymor_amd.calibrate``) prints one JSON line: a copy of S bytes reads S and
writes S, so its GB/s = 2*S / time; the read stream's GB/s = S / time.
"""
from __future__ import annotations

import ctypes as C
import json
import sys
from pathlib import Path

if __package__ in (None, ""):      # run as a script (e.g. under a profiler from another cwd)
    sys.path.insert(0, str(Path(__file__).resolve().parent.parent))
    from hymor_amd import _lib
else:
    from . import _lib


def copy_bandwidth(mib: int = 1024, reps: int = 20):
    nbytes = mib * 1024 * 1024
    ms = C.c_double()
    _lib.check(_lib.load().hymor_bench_copy(nbytes, reps, C.byref(ms)))
    return {"kernel": "k_copy", "bytes_each_way": nbytes, "ms_per_launch": ms.value,
            "GBps": 2.0 * nbytes / (ms.value * 1e-3) / 1e9, "reps": reps}


def read_bandwidth(mib: int = 1024, reps: int = 20):
    """Streaming read of `mib` MiB: GB/s = S / time."""
    nbytes = mib * 1024 * 1024
    ms = C.c_double()
    _lib.check(_lib.load().hymor_bench_read(nbytes, reps, C.byref(ms)))
    return {"kernel": "k_read", "bytes": nbytes, "ms_per_launch": ms.value,
            "GBps": nbytes / (ms.value * 1e-3) / 1e9, "reps": reps}


def mixed_bandwidth(read_bytes: int, write_bytes: int, reps: int = 20):
    """Streaming reads and writes in the proportion read_bytes : write_bytes
    (k_stream_mix): GB/s = bytes moved / time, the ceiling for a kernel
    whose HBM traffic has that read/write mix."""
    ms, moved = C.c_double(), C.c_double()
    _lib.check(_lib.load().hymor_bench_stream_mix(int(read_bytes), int(write_bytes), reps,
                                                  C.byref(ms), C.byref(moved)))
    return {"kernel": "k_stream_mix", "read_bytes": int(read_bytes),
            "write_bytes": int(write_bytes), "bytes_moved": moved.value,
            "ms_per_launch": ms.value, "GBps": moved.value / (ms.value * 1e-3) / 1e9,
            "reps": reps}


if __name__ == "__main__":
    mib = int(sys.argv[1]) if len(sys.argv) > 1 else 1024
    reps = int(sys.argv[2]) if len(sys.argv) > 2 else 20
    print(json.dumps({"copy": copy_bandwidth(mib, reps), "read": read_bandwidth(mib, reps)}),
          flush=True)
