"""Benchmark of the random-forest training hot path on one MI355X.

A step trains one whole tree of the mid-size configuration (500 synthetic
320x240 uint16 depth images, 1M sampled pixels, 8 classes, 2000 candidate
features x 20 thresholds per node, depth 16; tree ids cycle 0..2) in the
level-batched form, with the training set already resident in HBM.

value = sample*feature*threshold evaluations per second over the timed steps
(evals = sum over levels of active samples * nF * nT).

roofline: the dominant kernel is the node histogram.  Its algorithmic bytes
per launch are
    B = S*(4+4+1+b_px) + S*nF*2*b_px + nodes*nF*(FD+nT)*4 + nodes*nC*nT*nF*4
(S active samples and `nodes` frontier nodes of the launch, b_px bytes per
pixel); achieved = sum B / sum kernel time (HIP events on the trainer's
stream), peak = 8 TB/s HBM.

lds: the same kernel's LDS-atomic work, one 32-bit LDS add per (active
sample, candidate feature): U = sum S*nF; achieved = U / kernel time against
the chip's sustained LDS-atomic rate measured in the same run by the
library's probe kernel (pd_lds_atomic_bench).

extras (after the timed steps, skipped with --no-extras): batched
prediction Mpix/s and the learned split proposal's speedup / gain loss /
accuracy, the configuration's other headline figures.

cpu_baseline: the CPU oracle (an op-for-op restatement of the reference
histogram) timed on this host's cores on a bounded sample of the same
workload: the root node of the same data set with a subset of the candidate
features, sized to ~15 s.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def hist_bytes(S, nodes, nF, nT, nC, fd, b_px):
    return S * (4 + 4 + 1 + b_px) + S * nF * 2 * b_px + nodes * nF * (fd + nT) * 4 + nodes * nC * nT * nF * 4


def cpu_baseline(ts, params, cfg, target_s):
    """Oracle root-node histogram on a subset of candidates, all host cores."""
    import numpy as np

    from oracle import oracle as orc
    o = orc.Tset(cfg["kind"], ts)
    S = o.total
    sample_node = np.zeros(S, dtype=np.int32)
    from padenti_amd import TreeTrainerParameters
    nT = params.n_thresholds

    def run(nf):
        p = TreeTrainerParameters(nf, nT, params.feat_low_bounds, params.feat_up_bounds, params.thr_low_bound,
                                  params.thr_up_bound, params.per_leaf_samples_thr)
        t = time.perf_counter()
        orc.hist_nodes(o, p, 0, [0], sample_node)
        return time.perf_counter() - t

    threads = orc.num_threads()
    nf = max(threads, 8)
    dt = run(nf)
    nf2 = int(max(nf, min(params.n_features, nf * target_s / max(dt, 1e-3))))
    nf2 = max(threads, (nf2 // threads) * threads)
    dt2 = run(nf2)
    return {
        "value": S * nf2 * nT / dt2,
        "unit": "evals/s",
        "cores": threads,
        "kind": "port",
        "sample": f"oracle histogram of the root node: {S} samples x {nf2} of {params.n_features} candidates "
                  f"x {nT} thresholds ({dt2:.1f} s)",
    }


def lds_peak():
    """Chip-wide LDS atomic adds/s of the probe kernel (best of 3)."""
    import ctypes

    import padenti_amd as pa
    ms, rate = ctypes.c_float(), ctypes.c_double()
    best = 0.0
    for _ in range(3):
        pa.check(pa.lib().pd_lds_atomic_bench(0, 4096, 3, ctypes.byref(ms), ctypes.byref(rate)))
        best = max(best, rate.value)
    return best


def traffic_per_launch(cfg_name):
    """HBM bytes per histogram launch measured with rocprofv3 --pmc FETCH_SIZE /
    WRITE_SIZE on this workload (scaled by the copy-kernel calibration),
    recorded in profiles/k_hist_<cfg>_traffic.json; None if not measured."""
    path = os.path.join(ROOT, "profiles", f"k_hist_{cfg_name}_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("traffic_bytes_per_launch")


def extras():
    """The other headline figures of the configuration, measured after the
    timed steps (each in its own guarded section, so that a failure there is
    reported and cannot hide the training line):
      predict:  256 mixed RGB-D frames (76.2 Mpx) through 5 depth-20 trees in
                one hipGraph launch (padenti_amd/predict_bench.py);
      proposal: learned split proposal vs exhaustive evaluation of a
                20000-candidate pool on the same data, plain (network top
                2000) and refined (candidates ranked on a subsample of each
                node) (padenti_amd/proposal_bench.py);
      ingest:   uniform and class-weighted sampling of 5000 640x480 uint16
                images resident in HBM (padenti_amd/ingest_bench.py), with
                the CPU oracle (oracle/ingest.c, images in parallel over the
                host cores) timed on the same images as its baseline."""
    res = {}
    try:
        from padenti_amd import predict_bench
        r = predict_bench.run(predict_bench.parser().parse_args(["--reps", "5"]))
        res["predict"] = {k: r[k] for k in ("value", "unit", "frames", "pixels", "ms_graph_launch",
                                            "frame_latency_ms", "roofline")}
        print(f"[bench] predict {r['value']:.0f} Mpix/s", file=sys.stderr, flush=True)
    except Exception as e:   # reported, never fatal
        res["predict"] = {"error": repr(e)}
    try:
        from padenti_amd import proposal_bench
        r = proposal_bench.run(proposal_bench.parser().parse_args([]))
        res["proposal"] = {k: r[k] for k in ("value", "unit", "exhaustive_tree_s", "proposal_tree_s", "pool", "keep",
                                             "mean_gain_loss", "mean_gain_loss_screened",
                                             "mean_gain_loss_no_network", "heldout_pixel_accuracy",
                                             "accuracy_loss_pt", "records", "network_train_s", "refined")}
        print(f"[bench] proposal {r['value']:.2f}x over exhaustive, gain loss {r['mean_gain_loss']:.3f}; refined "
              f"{r['refined']['speedup']:.2f}x, gain loss {r['refined']['mean_gain_loss']:.4f}",
              file=sys.stderr, flush=True)
    except Exception as e:
        res["proposal"] = {"error": repr(e)}
    try:
        res["ingest"] = ingest_extra()
    except Exception as e:
        res["ingest"] = {"error": repr(e)}
    return res


def ingest_extra(n_images=5000):
    from padenti_amd import ingest_bench
    from oracle import ingest as oi
    from oracle import oracle as orc
    a = ingest_bench.parser().parse_args(["--images", str(n_images)])
    datas, labs = ingest_bench.make_dataset(a.images, a.w, a.h, a.classes, a.seed)
    r = ingest_bench.run(a, datas, labs)
    cpu = {}
    for kind, name in ((0, "uniform"), (1, "class_weighted")):
        t = time.perf_counter()
        oi.sample_images(kind, datas, labs, a.samples, a.classes if kind else 0, seed=7)
        dt = time.perf_counter() - t
        cpu[name] = {"value": len(datas) / dt, "unit": "images/s", "cores": orc.num_threads(), "kind": "port",
                     "sample": f"all {len(datas)} images, oracle/ingest.c (the reference's per-image loop), "
                               "images in parallel"}
        print(f"[bench] ingest {name}: GPU {r[name]['images_per_s']:.0f} images/s, CPU oracle "
              f"{len(datas) / dt:.0f} images/s ({orc.num_threads()} threads)", file=sys.stderr, flush=True)
    r["cpu_baseline"] = cpu
    return r


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c2")
    ap.add_argument("--mode", default="batched", choices=["batched", "per_node"])
    ap.add_argument("--images", type=int, default=None, help="scale the data set down (debug)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--levels-json", default=None, help="write per-level records to this file")
    ap.add_argument("--per-level", action="store_true", help="print per-level timings and evals/s")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the batched-prediction and split-proposal measurements after the timed steps")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (trees or candidate features "
              "would shard across GPUs); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import padenti_amd as pa
    from padenti_amd import synth
    if pa.device_count() < 1:
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)
    cfg = synth.CONFIGS[args.config]
    t0 = time.perf_counter()
    ts = synth.training_set(args.config, n_images=args.images)
    params = synth.params_for(args.config)
    gen_s = time.perf_counter() - t0
    print(f"[bench] generated {len(ts.images)} images in {gen_s:.1f} s", file=sys.stderr, flush=True)
    mode = pa.PD_MODE_LEVEL_BATCHED if args.mode == "batched" else pa.PD_MODE_PER_NODE
    tr = pa.HIPTreeTrainer(cfg["kind"], cfg["n_classes"])
    tr.set_training_set(ts)
    depth = cfg["depth"]
    fd = pa.FEATURES[cfg["kind"]][3]
    b_px = 2 if cfg["kind"] != pa.PD_FEAT_DEPTH_F32 else 4
    if cfg["kind"] == pa.PD_FEAT_RGBD_U16X4:
        b_px = 8

    def step(i):
        tree = pa.Tree(i % cfg["trees"], depth, cfg["kind"], cfg["n_classes"])
        t = time.perf_counter()
        tr.train(tree, ts, params, 1, depth, mode=mode)
        pa.lib().pd_dev_sync(0)
        return time.perf_counter() - t, tr.last_level_stats, tree

    for i in range(args.warmup):
        dt, _, _ = step(i)
        print(f"[bench] warmup tree {i}: {dt:.3f} s", file=sys.stderr, flush=True)
    pa.lib().pd_dev_sync(0)
    t_all = time.perf_counter()
    evals = 0.0
    hist_s = 0.0
    hist_b = 0.0
    lds_u = 0.0
    launches = 0
    levels = []
    for i in range(args.steps):
        dt, st, tree = step(i)
        e = sum(s["evals"] for s in st)
        evals += e
        for s in st:
            hist_s += s["hist_ms"] / 1e3
            launches += s["hist_launches"]
            hist_b += hist_bytes(s["active_samples"], s["frontier_nodes"], params.n_features, params.n_thresholds,
                                 cfg["n_classes"], fd, b_px)
            lds_u += s["active_samples"] * params.n_features
        levels.append(st)
        if args.per_level:
            for s in st:
                ms = s["hist_ms"] + s["split_ms"] + s["route_ms"]
                print(f"[bench]   level {s['level']:2d}: nodes {s['frontier_nodes']:7d} samples {s['active_samples']:9d} "
                      f"slices {s['slices']:3d} hist {s['hist_ms']:8.2f} ms split {s['split_ms']:7.2f} ms "
                      f"route {s['route_ms']:6.2f} ms  {s['evals'] / max(ms, 1e-6) / 1e9:7.2f} Teval/s",
                      file=sys.stderr, flush=True)
        print(f"[bench] tree {i % cfg['trees']}: {dt:.3f} s, {e / dt / 1e12:.2f} Teval/s, "
              f"splits {int((tree.left_children >= 0).sum())}", file=sys.stderr, flush=True)
    pa.lib().pd_dev_sync(0)
    total = time.perf_counter() - t_all
    value = evals / total
    achieved = hist_b / hist_s / 1e9 if hist_s > 0 else 0.0
    lds_p = lds_peak()
    lds_a = lds_u / hist_s if hist_s > 0 else 0.0
    if args.levels_json:
        with open(args.levels_json, "w") as f:
            json.dump(levels, f, indent=1)
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(ts, params, cfg, args.cpu_seconds)
        print(f"[bench] cpu baseline {cpu['value'] / 1e9:.3f} Geval/s on {cpu['cores']} threads",
              file=sys.stderr, flush=True)
    out = {
        "metric": "sample*feature*threshold evals/s (plain level-batched tree training)",
        "value": value,
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": total / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"{args.config}: train one depth-{depth} tree per step, {len(ts.images)} images "
                        f"{cfg['w']}x{cfg['h']}, {sum(im.samples.size for im in ts.images)} samples, "
                        f"{params.n_features} features x {params.n_thresholds} thresholds, "
                        f"{cfg['n_classes']} classes, {args.mode}",
            "tree_train_s": total / args.steps,
            "hist_kernel_launches": launches,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic_per_launch(args.config),
        },
        "lds": {"achieved": lds_a, "peak": lds_p, "unit": "atomic adds/s", "frac": lds_a / lds_p if lds_p else None},
        "cpu_baseline": cpu,
    }
    tr.close()
    if not args.no_extras and args.config == "c2" and args.images is None:
        out["extras"] = extras()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
