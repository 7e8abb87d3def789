"""Writes the golden fixtures of tests/golden (run from the repo root).

md5rand_vectors.json: md5Rand(seed) for fixed seeds, computed with hashlib
(MD5 of the 16 little-endian seed bytes) and, where oracle/_ref is built,
checked against the reference's own md5Rand from src/padenti/prng.hpp.
c1_param_table.json: candidate parameters / thresholds of tree 0, nodes
0-6, first 8 candidates, with the C1 bounds (numpy float32 closed form).
synth_digests.json: SHA-256 of generated images (generator determinism).
"""
import hashlib
import json
import os
import struct
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from tests import helpers as hp  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def md5_vectors():
    rng = np.random.default_rng(20240)
    seeds = [[0, 0, 0, 0], [0, 0, 0, 1], [0, 1, 0, 0], [0, 0, 99, 0], [1, 2, 3, 4], [0xffffffff] * 4]
    seeds += rng.integers(0, 2 ** 32, size=(26, 4), dtype=np.uint64).tolist()
    vec = [{"seed": [int(x) for x in s], "out": hp.md5rand(s)} for s in seeds]
    try:
        from oracle import oracle as orc
        ref = orc.ref_md5rand()
    except Exception:
        ref = None
    if ref is not None:
        for v in vec:
            assert ref(v["seed"]).tolist() == v["out"]
    return {"source": "hashlib md5 of struct.pack('<4I', seed)", "checked_against_reference_prng": ref is not None,
            "vectors": vec}


def param_table():
    lo, hi = [-60.0] * 4, [60.0] * 4
    rows = []
    for node in range(7):
        for f in range(8):
            p = hp.params(4, 0, node, f, lo, hi)
            t = hp.thresholds(10, 0, node, f, -0.5, 0.5)
            rows.append({"node": node, "f": f, "params_hex": p.tobytes().hex(), "thr_hex": t.tobytes().hex()})
    return {"tree": 0, "feat_bounds": [-60.0, 60.0], "thr_bounds": [-0.5, 0.5], "n_thresholds": 10, "rows": rows}


def synth_digests():
    from padenti_amd import synth
    out = {}
    for cfg, idx in [("c1", 0), ("c1", 5), ("c2", 0), ("c3", 1), ("c4", 3)]:
        c = synth.CONFIGS[cfg]
        px, lab = synth.gen_image(c["scene"], 0, idx, c["w"], c["h"], c["n_classes"])
        smp = synth.gen_samples(0, idx, lab, c["samples"])
        out[f"{cfg}:{idx}"] = {
            "pixels": hashlib.sha256(px.tobytes()).hexdigest(),
            "labels": hashlib.sha256(lab.tobytes()).hexdigest(),
            "samples": hashlib.sha256(smp.tobytes()).hexdigest(),
        }
    return out


if __name__ == "__main__":
    with open(os.path.join(OUT, "md5rand_vectors.json"), "w") as f:
        json.dump(md5_vectors(), f, indent=1)
    with open(os.path.join(OUT, "c1_param_table.json"), "w") as f:
        json.dump(param_table(), f, indent=1)
    with open(os.path.join(OUT, "synth_digests.json"), "w") as f:
        json.dump(synth_digests(), f, indent=1)
    print("golden fixtures written")
