"""Autotuner of the histogram launch.

    python -m padenti_amd.autotune --config c2 [--config c3] [--out padenti_amd/tuning.json]

For a workload it sweeps the launch tuning of the histogram kernel

    waves per block    4, 8, 16          (samples of a work item are strided
                                          over the waves; the block shares one
                                          64-candidate LDS histogram tile)
    blocks per CU      4, 8, 16, 32, 64  (the chunk size: a level's samples
                                          are cut into work items so that the
                                          launch holds about this many items
                                          per CU; it also decides which nodes
                                          are searched inside the kernel)

over level shapes of the real training set (the samples cut into 1, 16, 256
and 4096 equal nodes: the root, shallow, middle and deep levels of a tree),
times each (hist + split device time, pd_trainer_time_level) and keeps the
tuning with the smallest total.  Entries are keyed by (feature hook,
threshold bucket, classes); HIPTreeTrainer applies the matching entry before
every train() call.
"""
import argparse
import ctypes
import json
import os
import sys
import time

from . import _lib as L

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE = os.path.join(HERE, "tuning.json")
WAVES = (4, 8, 16)
BLOCKS_PER_CU = (4, 8, 16, 32, 64)
SHAPES = (1, 16, 256, 4096)
THR_BUCKETS = (1, 8, 10, 16, 20, 24, 32, 48, 64)


def thr_bucket(nT):
    for b in THR_BUCKETS:
        if nT <= b:
            return b
    return 64


def key(kind, nT, nC):
    return f"{int(kind)}:{thr_bucket(nT)}:{int(nC)}"


def load_table(path=TABLE):
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return {e["key"]: e for e in json.load(f).get("entries", [])}


def set_tuning(trainer, waves, blocks_per_cu):
    L.check(L.lib().pd_trainer_set_tuning(trainer._h, ctypes.byref(L.Tuning(waves, blocks_per_cu))))


def sweep(trainer, params, shapes=SHAPES, reps=2, log=None):
    """Device ms of every (waves, blocks_per_cu) on every shape."""
    out = {}
    ms = ctypes.c_double()
    for w in WAVES:
        for b in BLOCKS_PER_CU:
            set_tuning(trainer, w, b)
            row = {}
            for n in shapes:
                L.check(L.lib().pd_trainer_time_level(trainer._h, params.c(), 0, n, reps, ctypes.byref(ms)))
                row[n] = ms.value
            out[(w, b)] = row
            if log:
                log(f"[autotune] waves {w:2d} blocks/CU {b:2d}: " +
                    " ".join(f"{n}:{row[n]:.2f}ms" for n in shapes))
    return out


def tune_workload(cfg_name, n_images=None, reps=2, log=None):
    import padenti_amd as pa
    from padenti_amd import synth
    c = synth.CONFIGS[cfg_name]
    ts = synth.training_set(cfg_name, n_images=n_images)
    p = synth.params_for(cfg_name)
    tr = pa.HIPTreeTrainer(c["kind"], c["n_classes"], tuned=False)
    tr.set_training_set(ts)
    S = sum(im.samples.size for im in ts.images)
    shapes = tuple(n for n in SHAPES if n * 64 <= S)
    t0 = time.perf_counter()
    res = sweep(tr, p, shapes, reps, log)
    total = {k: sum(v.values()) for k, v in res.items()}
    best = min(total, key=total.get)
    default = total[(8, 16)]
    tr.close()
    return {
        "key": key(c["kind"], p.n_thresholds, c["n_classes"]),
        "workload": f"{cfg_name}: {len(ts.images)} images {c['w']}x{c['h']}, {S} samples, "
                    f"{p.n_features} x {p.n_thresholds}, {c['n_classes']} classes",
        "waves": best[0], "blocks_per_cu": best[1],
        "total_ms": total[best], "default_total_ms": default, "gain_vs_default": default / total[best],
        "shapes": list(shapes),
        "sweep": [{"waves": k[0], "blocks_per_cu": k[1], "ms": [v[n] for n in shapes]} for k, v in res.items()],
        "sweep_s": time.perf_counter() - t0,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--config", action="append", default=None)
    ap.add_argument("--images", type=int, default=None)
    ap.add_argument("--reps", type=int, default=2)
    ap.add_argument("--out", default=TABLE)
    a = ap.parse_args()
    log = lambda m: print(m, file=sys.stderr, flush=True)   # noqa: E731
    table = load_table(a.out)
    for cfg in a.config or ["c2"]:
        e = tune_workload(cfg, a.images, a.reps, log)
        table[e["key"]] = e
        log(f"[autotune] {cfg}: waves {e['waves']} blocks/CU {e['blocks_per_cu']} "
            f"({e['gain_vs_default']:.3f}x vs the default 8 / 16)")
    with open(a.out, "w") as f:
        json.dump({"entries": sorted(table.values(), key=lambda e: e["key"])}, f, indent=1)
    print(json.dumps({k: (v["waves"], v["blocks_per_cu"], round(v["gain_vs_default"], 3)) for k, v in table.items()}))


if __name__ == "__main__":
    main()
