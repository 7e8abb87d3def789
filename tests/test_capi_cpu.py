"""C-ABI library checks that need no GPU: it loads, exports every symbol
the header declares, reports no device here and fails cleanly."""
import ctypes
import hashlib
import json
import os
import re

import numpy as np
import pytest

import padenti_amd
from padenti_amd import _lib as L
from padenti_amd import synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "padenti_hip.h")).read()
    return sorted(set(re.findall(r"\b(pd_[a-z0-9_]+)\s*\(", src)))


def test_header_symbols_exported():
    lib = ctypes.CDLL(L.LIB_PATH)
    names = declared_symbols()
    assert len(names) >= 25
    for n in names:
        assert hasattr(lib, n), n


def test_version_and_no_device_here():
    assert L.lib().pd_version() >= 100
    n = padenti_amd.device_count()
    assert n >= 0


@pytest.mark.skipif(padenti_amd.device_count() > 0, reason="checks the no-device error path")
def test_create_without_device_fails_cleanly():
    with pytest.raises(padenti_amd.PdError) as e:
        padenti_amd.HIPTreeTrainer(padenti_amd.PD_FEAT_DEPTH_F32, 4)
    assert e.value.code == -2 and "device" in str(e.value)


def test_bad_arguments_rejected_before_device_use():
    h = ctypes.c_void_p()
    assert L.lib().pd_trainer_create(ctypes.byref(h), 0, 0, 0) == -1          # n_classes 0
    assert b"n_classes" in L.lib().pd_last_error()
    assert L.lib().pd_forest_create(ctypes.byref(h), 0, 0, 65) == -1          # too many classes
    assert L.lib().pd_trainer_create(None, 0, 0, 4) == -1


def test_synth_generator_is_deterministic():
    with open(os.path.join(ROOT, "tests", "golden", "synth_digests.json")) as f:
        gold = json.load(f)
    for key, want in gold.items():
        cfg, idx = key.split(":")
        c = synth.CONFIGS[cfg]
        px, lab = synth.gen_image(c["scene"], 0, int(idx), c["w"], c["h"], c["n_classes"])
        smp = synth.gen_samples(0, int(idx), lab, c["samples"])
        assert hashlib.sha256(px.tobytes()).hexdigest() == want["pixels"], key
        assert hashlib.sha256(lab.tobytes()).hexdigest() == want["labels"], key
        assert hashlib.sha256(smp.tobytes()).hexdigest() == want["samples"], key


def test_synth_shapes_and_classes():
    c = synth.CONFIGS["c2"]
    px, lab = synth.gen_image(c["scene"], 0, 1, 320, 240, 8)
    assert px.dtype == np.uint16 and px.shape == (240, 320)
    assert 600 <= int(px.min()) and int(px.max()) <= 4100
    assert set(np.unique(lab)) <= set(range(1, 9)) and len(np.unique(lab)) >= 4
    px, lab = synth.gen_image(synth.CONFIGS["c4"]["scene"], 0, 1, 64, 48, 21)
    assert px.shape == (48, 64, 4) and int(px[..., :3].max()) <= 255


def test_c1_param_table_fixture():
    from tests import helpers as hp
    with open(os.path.join(ROOT, "tests", "golden", "c1_param_table.json")) as f:
        gold = json.load(f)
    for r in gold["rows"]:
        p = hp.params(4, 0, r["node"], r["f"], [-60.0] * 4, [60.0] * 4)
        t = hp.thresholds(10, 0, r["node"], r["f"], -0.5, 0.5)
        assert p.tobytes().hex() == r["params_hex"] and t.tobytes().hex() == r["thr_hex"]
    from oracle import oracle as orc
    for r in gold["rows"]:
        assert orc.feature_params(0, 0, r["node"], r["f"], [-60.0] * 4, [60.0] * 4).tobytes().hex() == r["params_hex"]
        assert orc.thresholds(0, 0, r["node"], r["f"], 10, -0.5, 0.5).tobytes().hex() == r["thr_hex"]
