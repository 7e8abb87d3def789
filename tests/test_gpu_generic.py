"""Every FeatType and images of more than four channels on the GPU: the
ChannelDiff hook (padenti_amd/examples/channel_diff.hpp) instantiated ten
times as plugins trains trees bit-identical to the oracle's generic kinds
and predicts identical leaves and posteriors.  Both threshold buckets of the
plugins' light kernel set are used (nT = 5 -> 16 slots, nT = 20 -> 64)."""
import os

import numpy as np
import pytest

import padenti_amd as pa
from oracle import oracle as orc

from .test_oracle_generic import COMBOS, bounds_for, generic_tset, random_image

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(pa.__file__))
NAMES = ["cd_u8x10_f64", "cd_i16x5_i32", "cd_u16x3_u8", "cd_f32x6_i64", "cd_u8x1_i8",
         "cd_u16x2_u32", "cd_i8x7_u16", "cd_u16x16_u64", "cd_f64x4_f32", "cd_i32x3_i16"]


def plugin(i):
    it, nch, ft = COMBOS[i]
    lib = os.path.join(HERE, "libpd_channel_diff_%s.so" % ("a" if i < 5 else "b"))
    return pa.Plugin(lib, NAMES[i], it, nch, ft, 4)


@pytest.mark.parametrize("i", range(10))
def test_generic_hook_trains_and_predicts_like_the_oracle(gpu, i):
    it, nch, ft = COMBOS[i]
    print(f"[generic] {NAMES[i]}", flush=True)
    pl = plugin(i)
    k = orc.generic_kind(it, nch, ft)
    rng = np.random.default_rng(100 + i)
    nC = 4 if i % 2 == 0 else 6
    ts = generic_tset(it, nch, nC, rng, n_images=5)
    lo, hi = bounds_for(ft)
    nT = 5 if i < 5 else 20
    p = pa.TreeTrainerParameters(48, nT, lo, hi, -40 if np.dtype(ft).kind != "u" else 0, 40, 2)
    tree = pa.Tree(1, 7, pl, nC)
    tr = pa.HIPTreeTrainer(None, nC, plugin=pl)
    tr.train(tree, ts, p, 1, 7)
    out = orc.train_tree(orc.Tset(k, ts), p, ts.priors, 1, 7)
    assert np.array_equal(tree.left_children, out["left_children"])
    assert tree.features.tobytes() == out["features"].tobytes()
    assert tree.thresholds.tobytes() == out["thresholds"].tobytes()
    assert tree.posteriors.tobytes() == out["posteriors"].tobytes()
    assert (tree.left_children >= 0).sum() >= 3   # it actually split
    clf = pa.HIPClassifier(None, nC, plugin=pl)
    clf << tree
    img = random_image(it, nch, 21, 27, rng)
    mask = (np.arange(21 * 27) % 5 != 0).astype(np.uint8).reshape(21, 27)
    assert np.array_equal(clf.predict_leaves(0, img, 27, 21, mask), orc.predict_leaves(k, img, mask, tree))
    assert clf.predict_posterior(img, 27, 21).tobytes() == orc.predict_posterior(k, img, None, [tree], nC).tobytes()
    tr.close()
    clf.close()
