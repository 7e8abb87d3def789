"""Texel-gather probe: the ceilings of the histogram kernel's scattered
uint16 gathers on this GPU.

    python -m padenti_amd.gather_probe [--iters 2048] [--reps 3]

For windows inside the vector L1 (16 KiB), inside L2 (2 MiB) and beyond it
(64 MiB) and for 1..64 distinct 128-byte lines per 64-lane load, prints CU
cycles per wave load (one JSON line per case, then a table).  The L2-window
rows give the cost of an L1 miss (a 128-byte L2 -> L1 fill) that bounds the
histogram kernel, whose lookups per gather are reported by the
TCP_TCC_READ_REQ counter.
"""
import argparse
import ctypes
import json

from . import _lib as L


def run(window, lines, iters, reps):
    ms = ctypes.c_float()
    cyc = ctypes.c_double()
    L.check(L.lib().pd_gather_bench(0, window, lines, iters, reps, ctypes.byref(ms), ctypes.byref(cyc)))
    return ms.value, cyc.value


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--iters", type=int, default=2048)
    ap.add_argument("--reps", type=int, default=3)
    a = ap.parse_args()
    rows = []
    for window in (16 << 10, 2 << 20, 64 << 20):
        for lines in (1, 4, 16, 32, 64):
            ms, cyc = run(window, lines, a.iters, a.reps)
            r = {"window_bytes": window, "lines_per_load": lines, "ms": ms, "cu_cycles_per_load": cyc,
                 "cu_cycles_per_line": cyc / lines}
            rows.append(r)
            print(json.dumps(r), flush=True)
    print("\n| window | lines per load | CU cycles per load | per line |\n|---|---|---|---|")
    for r in rows:
        print(f"| {r['window_bytes'] >> 10} KiB | {r['lines_per_load']} | {r['cu_cycles_per_load']:.1f} | "
              f"{r['cu_cycles_per_line']:.2f} |")


if __name__ == "__main__":
    main()
