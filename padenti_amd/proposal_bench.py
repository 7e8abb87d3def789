"""Learned split-proposal benchmark (mid-size depth workload).

    python -m padenti_amd.proposal_bench [--seeds 1 2 3 4] [--per-node 350] [--epochs 2]

1. Records: plain level-batched trainings with the mid-size settings (500
   synthetic 320x240 uint16 capsule images, 2000 samples each, 2000 features
   x 20 thresholds, depth 16, trees 0..2) on data seeds 1..4 emit
   `per_node` (network input, measured gain) records per trained node.
2. The network is trained on the GPU (BF16 MFMA forward/backward, Adam).
3. On seed-0 data the same tree is trained
     exhaustive: the whole pool of K = 20000 candidates per node (nF = K,
                 nT = 1), the judge;
     proposal:   the network scores the pool and the best 2000 are
                 histogrammed exactly (nodes below --min-samples samples
                 evaluate the whole pool: the fallback);
     no network: the first 2000 candidates (plain trainer with nF = 2000);
   timed end to end (HIP-synchronised wall clock of pd_trainer_train), and
   the proposal is run once more with the judge on to get the gain lost per
   level against the exhaustive pool; the judge also runs for the screened
   selection (keep taken in sampler order from the best --screen scores) and
   for constant scores (the first 2000 candidates: no network).
   refined:    the network's best --refine-screen candidates of a node
               (default: the whole pool) are ranked by their exact gain on
               every --refine-stride-th sample (at least --refine-min) and the
               best 2000 histogrammed; with a screen below the pool the same
               refinement of the whole pool runs beside it.
4. Held-out pixel accuracy of each tree on 20 held-out images.
Prints per-level lines on stderr and one JSON line.
"""
import argparse
import json
import sys
import time

import numpy as np

import padenti_amd as pa
from padenti_amd import proposal as pp
from padenti_amd import synth


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def accuracy(tree, kind, nC, held):
    clf = pa.HIPClassifier(kind, nC)
    clf << tree
    ok = tot = 0
    for px, lab in held:
        h, w = lab.shape
        post = clf.predict_posterior(px, w, h)
        pred = np.argmax(post.reshape(nC, -1), axis=0)
        m = lab.reshape(-1) > 0
        ok += int((pred[m] == lab.reshape(-1)[m].astype(np.int64) - 1).sum())
        tot += int(m.sum())
    clf.close()
    return ok / tot


def timed_train(tr, ts, p, depth, mode, tree_id=0):
    tree = pa.Tree(tree_id, depth, synth.CONFIGS["c2"]["kind"], synth.CONFIGS["c2"]["n_classes"])
    pa.lib().pd_dev_sync(0)
    t = time.perf_counter()
    tr.train(tree, ts, p, 1, depth, mode=mode)
    pa.lib().pd_dev_sync(0)
    return tree, time.perf_counter() - t, tr.last_level_stats


def main():
    print(json.dumps(run(parser().parse_args())), flush=True)


def parser():
    ap = argparse.ArgumentParser()
    ap.add_argument("--seeds", type=int, nargs="+", default=[1, 2, 3, 4])
    ap.add_argument("--images", type=int, default=500)
    ap.add_argument("--trees", type=int, default=3)
    ap.add_argument("--per-node", type=int, default=350)
    ap.add_argument("--epochs", type=int, default=2)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--pool", type=int, default=20000)
    ap.add_argument("--keep", type=int, default=2000)
    ap.add_argument("--min-samples", type=int, default=256)
    ap.add_argument("--screen", type=int, default=10000,
                    help="also judge the screened selection: keep taken in sampler order from the best `screen`")
    ap.add_argument("--depth", type=int, default=16)
    ap.add_argument("--refine-screen", type=int, default=0,
                    help="refinement: the network's best `refine-screen` candidates of a node (0: the whole "
                         "pool) ...")
    ap.add_argument("--refine-stride", type=int, default=16,
                    help="... are ranked by their exact gain on every refine-stride-th sample ...")
    ap.add_argument("--refine-min", type=int, default=1024,
                    help="... (at least refine-min samples) and the best `keep` histogrammed")
    return ap


def run(a):
    """The whole benchmark; returns the result record."""
    c = synth.CONFIGS["c2"]
    kind, nC = c["kind"], c["n_classes"]
    tr = pa.HIPTreeTrainer(kind, nC)

    # 1. records
    cap = len(a.seeds) * a.trees * 20000 * a.per_node
    buf = pp.RecordBuffer(cap)
    tr.set_record_sink(buf, per_node=a.per_node)
    prec = synth.params_for("c2")
    t0 = time.perf_counter()
    for seed in a.seeds:
        ts = synth.training_set("c2", seed=seed, n_images=a.images)
        for tid in range(a.trees):
            tr.train(pa.Tree(tid, a.depth, kind, nC), ts, prec, 1, a.depth)
        log(f"[proposal] seed {seed}: {tr.records_written()} records so far")
    n_rec = tr.records_written()
    tr.set_record_sink(None, 0)
    rec_s = time.perf_counter() - t0

    # 2. network training
    net = pp.SplitProposalNet(seed=7)
    ep_loss, train_ms = net.train_records(buf, batch=a.batch, epochs=a.epochs, opt=pp.adam(lr=a.lr))
    log(f"[proposal] {n_rec} records in {rec_s:.1f} s; network: epoch losses {ep_loss.tolist()}, "
        f"{train_ms / 1e3:.2f} s on the GPU")
    trained = net.params
    buf.close()

    # 3. seed-0 trees
    ts = synth.training_set("c2", seed=0, n_images=a.images)
    tr.set_training_set(ts)
    held = synth.heldout("c2", seed=0)
    pool = synth.params_for("c2", n_features=a.pool, n_thresholds=1)
    small = synth.params_for("c2", n_features=a.keep, n_thresholds=1)
    timed_train(tr, ts, small, a.depth, pa.PD_MODE_LEVEL_BATCHED)   # warm-up
    t_exh_tree, t_exh, st_exh = timed_train(tr, ts, pool, a.depth, pa.PD_MODE_LEVEL_BATCHED)
    tr.set_proposal(net, keep=a.keep, min_samples=a.min_samples, judge=False)
    t_prop_tree, t_prop, st_prop = timed_train(tr, ts, pool, a.depth, pa.PD_MODE_PROPOSAL)
    tr.set_proposal(net, keep=a.keep, min_samples=a.min_samples, judge=True)
    judged_tree, _, st_judge = timed_train(tr, ts, pool, a.depth, pa.PD_MODE_PROPOSAL)
    assert np.array_equal(judged_tree.left_children, t_prop_tree.left_children)
    tr.set_proposal(net, keep=a.keep, min_samples=a.min_samples, judge=True, screen=a.screen)
    screened_tree, t_screen, st_screen = timed_train(tr, ts, pool, a.depth, pa.PD_MODE_PROPOSAL)
    # refinement: the network's best candidates ranked exactly on a subsample;
    # beside it the same refinement of the whole pool (no use of the network)
    # the same ranked nodes as the plain proposal (a ranked node below
    # --refine-min samples is ranked on all of them)
    rscreen, rmin = a.refine_screen or a.pool, a.min_samples
    tr.set_proposal_refine(a.refine_stride, a.refine_min)
    tr.set_proposal(net, keep=a.keep, min_samples=rmin, judge=False, screen=rscreen)
    ref_tree, t_ref, st_ref = timed_train(tr, ts, pool, a.depth, pa.PD_MODE_PROPOSAL)
    tr.set_proposal(net, keep=a.keep, min_samples=rmin, judge=True, screen=rscreen)
    _, _, st_ref_j = timed_train(tr, ts, pool, a.depth, pa.PD_MODE_PROPOSAL)
    if rscreen < a.pool:
        tr.set_proposal(net, keep=a.keep, min_samples=rmin, judge=False, screen=a.pool)
        refall_tree, t_refall, _ = timed_train(tr, ts, pool, a.depth, pa.PD_MODE_PROPOSAL)
        tr.set_proposal(net, keep=a.keep, min_samples=rmin, judge=True, screen=a.pool)
        _, _, st_refall_j = timed_train(tr, ts, pool, a.depth, pa.PD_MODE_PROPOSAL)
    else:
        refall_tree, t_refall, st_refall_j = ref_tree, t_ref, st_ref_j
    tr.set_proposal_refine(0, 0)
    t_small_tree, t_small, _ = timed_train(tr, ts, small, a.depth, pa.PD_MODE_LEVEL_BATCHED)
    net.params = np.zeros_like(trained)   # constant scores: the first `keep` candidates
    tr.set_proposal(net, keep=a.keep, min_samples=a.min_samples, judge=True)
    _, _, st_const = timed_train(tr, ts, pool, a.depth, pa.PD_MODE_PROPOSAL)

    def gl(st):
        n = sum(s["judged_nodes"] for s in st)
        return sum(s["gain_loss_rel"] for s in st) / max(n, 1)

    levels = []
    for sp, sj, sc, se, sr, srj in zip(st_prop, st_judge, st_const, st_exh, st_ref, st_ref_j):
        n = sj["judged_nodes"]
        row = {"level": sp["level"], "proposal_nodes": sp["proposal_nodes"], "fallback_nodes": sp["fallback_nodes"],
               "gain_loss": sj["gain_loss_rel"] / n if n else None,
               "gain_loss_no_network": sc["gain_loss_rel"] / sc["judged_nodes"] if sc["judged_nodes"] else None,
               "gain_loss_refined": srj["gain_loss_rel"] / srj["judged_nodes"] if srj["judged_nodes"] else None,
               "refined_ms": sr["hist_ms"] + sr["split_ms"] + sr["mlp_ms"] + sr["select_ms"] + sr["refine_ms"],
               "refine_ms": sr["refine_ms"],
               "exhaustive_ms": se["hist_ms"] + se["split_ms"],
               "proposal_ms": sp["hist_ms"] + sp["split_ms"] + sp["mlp_ms"] + sp["select_ms"],
               "mlp_ms": sp["mlp_ms"], "select_ms": sp["select_ms"]}
        levels.append(row)
        log(f"[proposal] level {row['level']:2d}: ranked {row['proposal_nodes']:5d} fallback {row['fallback_nodes']:5d} "
            f"gain loss {row['gain_loss'] if row['gain_loss'] is not None else float('nan'):.4f} "
            f"(no network {row['gain_loss_no_network'] if row['gain_loss_no_network'] is not None else float('nan'):.4f}) "
            f"exhaustive {row['exhaustive_ms']:8.2f} ms proposal {row['proposal_ms']:8.2f} ms "
            f"(mlp {row['mlp_ms']:.2f}, select {row['select_ms']:.2f}); refined: gain loss "
            f"{row['gain_loss_refined'] if row['gain_loss_refined'] is not None else float('nan'):.4f} "
            f"{row['refined_ms']:8.2f} ms (subsample ranking {row['refine_ms']:.2f})")
    acc = {"exhaustive": accuracy(t_exh_tree, kind, nC, held), "proposal": accuracy(t_prop_tree, kind, nC, held),
           "screened": accuracy(screened_tree, kind, nC, held),
           "refined": accuracy(ref_tree, kind, nC, held), "refined_whole_pool": accuracy(refall_tree, kind, nC, held),
           "first_keep_no_network": accuracy(t_small_tree, kind, nC, held)}
    rows_scored = sum(s["proposal_nodes"] for s in st_prop) * a.pool
    mlp_s = sum(s["mlp_ms"] for s in st_prop) / 1e3
    flop = rows_scored * 2 * (22 * 256 + 2 * 256 * 256 + 256)
    out = {
        "metric": "split-proposal tree training speedup over exhaustive pool evaluation (same data)",
        "value": t_exh / t_prop,
        "unit": "x",
        "exhaustive_tree_s": t_exh,
        "proposal_tree_s": t_prop,
        "first_keep_tree_s": t_small,
        "pool": a.pool, "keep": a.keep, "min_samples": a.min_samples,
        "mean_gain_loss": gl(st_judge),
        "mean_gain_loss_screened": gl(st_screen), "screen": a.screen,
        "mean_gain_loss_no_network": gl(st_const),
        "refined": {"screen": rscreen, "stride": a.refine_stride, "min_samples": a.refine_min, "min_ranked": rmin,
                    "tree_s": t_ref, "speedup": t_exh / t_ref, "mean_gain_loss": gl(st_ref_j),
                    "accuracy_loss_pt": 100 * (acc["exhaustive"] - acc["refined"]),
                    "whole_pool": {"tree_s": t_refall, "speedup": t_exh / t_refall,
                                   "mean_gain_loss": gl(st_refall_j),
                                   "accuracy_loss_pt": 100 * (acc["exhaustive"] - acc["refined_whole_pool"])}},
        "heldout_pixel_accuracy": acc,
        "accuracy_loss_pt": 100 * (acc["exhaustive"] - acc["proposal"]),
        "records": n_rec, "record_s": rec_s,
        "network_epoch_loss": ep_loss.tolist(), "network_train_s": train_ms / 1e3,
        "scoring": {"rows": rows_scored, "s": mlp_s,
                    "tflops": flop / mlp_s / 1e12 if mlp_s else None, "peak_bf16_tflops": 2500.0},
        "levels": levels,
    }
    net.close()
    tr.close()
    return out


if __name__ == "__main__":
    main()
