"""HBM probe: times the library's plain copy kernel (read + write of N bytes).

    python -m padenti_amd.hbm_probe [--gib 4] [--reps 20]

Prints one JSON line with the copy bandwidth.  Run under
``rocprofv3 --pmc FETCH_SIZE`` / ``--pmc WRITE_SIZE`` it is also the
reference that scales those counters to bytes (a copy of N bytes reads N
and writes N).
"""
import argparse
import ctypes
import json

from . import _lib as L


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gib", type=float, default=4.0)
    ap.add_argument("--reps", type=int, default=20)
    a = ap.parse_args()
    nbytes = int(a.gib * (1 << 30)) // 4096 * 4096
    ms = ctypes.c_float()
    L.check(L.lib().pd_copy_bench(0, nbytes, a.reps, ctypes.byref(ms)))
    gbs = 2 * nbytes / (ms.value * 1e-3) / 1e9
    print(json.dumps({"copy_bytes": nbytes, "ms_per_copy": ms.value, "read_plus_write_GBps": gbs}), flush=True)


if __name__ == "__main__":
    main()
