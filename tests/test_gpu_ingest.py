"""Dataset ingestion on the GPU against the CPU oracle (oracle/ingest.c):
device CRC-32 (vs zlib too), both samplers bit-exact on ragged batches and
edge cases, the device-resident entry point, RGB -> label mapping, and a
training set built from a directory of PNGs that trains the oracle's tree."""
import ctypes as C
import zlib

import numpy as np
import pytest

import padenti_amd as pa
from padenti_amd import dataset as ds
from padenti_amd import synth
from padenti_amd._lib import check, lib, ptr
from oracle import ingest as oi
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


class DevArray:
    def __init__(self, nbytes):
        self.p = C.c_void_p()
        check(lib().pd_dev_alloc(0, max(int(nbytes), 16), C.byref(self.p)))

    def put(self, a, off=0):
        a = np.ascontiguousarray(a)
        check(lib().pd_memcpy_h2d(self.p.value + off, ptr(a), a.nbytes))

    def get(self, n, dtype, off=0):
        out = np.empty(n, dtype=dtype)
        check(lib().pd_memcpy_d2h(ptr(out), self.p.value + off, out.nbytes))
        return out

    def free(self):
        check(lib().pd_dev_free(self.p))


def test_crc32_ragged_ranges(gpu):
    rng = np.random.default_rng(1)
    lengths = [0, 1, 3, 15, 63, 64, 65, 127, 4031, 4095, 4096, 4097, 8191, 8192, 12345, 100003, 614400,
               (1 << 20) + 7]
    offs, o = [], 5   # odd start: every alignment case
    for n in lengths:
        offs.append(o)
        o += n + int(rng.integers(0, 70))
    buf = rng.integers(0, 256, o + 64, dtype=np.uint8)
    d = DevArray(buf.size)
    d.put(buf)
    crc = DevArray(4 * len(lengths))
    ms = C.c_float()
    offs_a, len_a = np.array(offs, np.uint64), np.array(lengths, np.uint64)   # alive across the call
    assert int((offs_a + len_a).max()) <= buf.size
    check(lib().pd_crc32_images_dev(0, len(lengths), d.p, buf.size, ptr(offs_a), ptr(len_a), crc.p, C.byref(ms)))
    got = crc.get(len(lengths), np.uint32)
    for i, (s, n) in enumerate(zip(offs, lengths)):
        assert got[i] == zlib.crc32(buf[s:s + n].tobytes()), (i, s, n)
        assert got[i] == oi.crc32(buf[s:s + n])
    d.free()
    crc.free()


def _batch(seed, n=24):
    """Ragged images: sizes from 1x1 to 640x480, label densities from none to all."""
    rng = np.random.default_rng(seed)
    shapes = [(1, 1), (1, 7), (5, 1), (3, 3), (17, 33), (31, 32), (64, 64), (240, 320), (480, 640)]
    datas, labs = [], []
    for i in range(n):
        h, w = shapes[i % len(shapes)] if i < len(shapes) else (int(rng.integers(1, 130)), int(rng.integers(1, 170)))
        kind = i % 4
        if kind == 0:
            d = rng.integers(0, 65536, (h, w), dtype=np.uint16)
        elif kind == 1:
            d = rng.random((h, w), dtype=np.float32)
        elif kind == 2:
            d = rng.integers(0, 256, (h, w, 3), dtype=np.uint8)
        else:
            d = rng.integers(0, 65536, (h, w, 4), dtype=np.uint16)
        dens = [0.0, 0.02, 0.5, 1.0][(i // 4) % 4]
        lab = rng.integers(1, 9, (h, w)).astype(np.uint8)
        lab[rng.random((h, w)) >= dens] = 0
        if i == 5:   # exactly one labelled pixel
            lab[:] = 0
            lab[h // 2, w // 2] = 3
        datas.append(d)
        labs.append(lab)
    return datas, labs


@pytest.mark.parametrize("n_samples,seed", [(1, 0), (7, 3), (2000, 0), (2048, 0xFFFFFFFF)])
def test_uniform_sampler_matches_oracle(gpu, n_samples, seed):
    datas, labs = _batch(11 + n_samples)
    smp = ds.UniformImageSampler(n_samples, seed=seed)
    got, crc = smp.sample_batch(datas, labs, crc=True)
    want, wcrc = oi.sample_images(0, datas, labs, n_samples, seed=seed)
    assert np.array_equal(crc, wcrc)
    for i, (g, w) in enumerate(zip(got, want)):
        assert np.array_equal(g, w), i
        if (labs[i] > 0).any():
            assert g.size == n_samples


@pytest.mark.parametrize("n_classes,weights,n_samples", [
    (8, None, 2000), (8, "random", 300), (8, "zeros", 64), (1, [1.0], 5), (8, "big", 50), (255, None, 3000),
    (8, "big", 2000)])   # the last: 48000 slots per image
def test_weighted_sampler_matches_oracle(gpu, n_classes, weights, n_samples):
    datas, labs = _batch(7 + n_classes + n_samples)
    rng = np.random.default_rng(n_samples)
    if n_classes == 1:
        labs = [np.minimum(l, 1) for l in labs]
    if n_classes == 255:
        labs = [np.where(l > 0, rng.integers(1, 256, l.shape), 0).astype(np.uint8) for l in labs]
    if weights == "random":
        weights = rng.random(n_classes).astype(np.float32)
    elif weights == "zeros":
        weights = np.array([0.5, 0, 0.25, 0, 0.1, 0, 0.7, 0.3], np.float32)
    elif weights == "big":
        weights = np.full(n_classes, 3.0, np.float32)   # buckets above n_samples
    smp = ds.ClassWeightedImageSampler(n_samples, n_classes, weights, seed=5)
    got = smp.sample_batch(datas, labs)
    want, _ = oi.sample_images(1, datas, labs, n_samples, n_classes, weights, seed=5)
    for i, (g, w) in enumerate(zip(got, want)):
        assert np.array_equal(g, w), i


def test_weighted_sampler_rejects_label_above_classes(gpu):
    lab = np.array([[1, 2, 5]], np.uint8)
    with pytest.raises(pa.PdError, match="exceeds n_classes"):
        ds.ClassWeightedImageSampler(4, 3).sample(np.zeros((1, 3), np.uint16), lab)


def test_device_entry_point_and_stats(gpu):
    datas, labs = _batch(99, n=12)
    raws = [np.ascontiguousarray(d).view(np.uint8).reshape(-1) for d in datas]
    dbytes = np.array([r.size for r in raws], np.uint64)
    doff = np.concatenate([[3], 3 + np.cumsum(dbytes)[:-1]]).astype(np.uint64)   # unaligned image bytes
    lsz = np.array([l.size for l in labs], np.uint64)
    loff = np.concatenate([[0], np.cumsum((lsz + 15) // 16 * 16)[:-1]]).astype(np.uint64)
    w = np.array([l.shape[1] for l in labs], np.uint32)
    h = np.array([l.shape[0] for l in labs], np.uint32)
    dd, dl = DevArray(int(doff[-1] + dbytes[-1])), DevArray(int(loff[-1] + lsz[-1]))
    for r, o in zip(raws, doff):
        dd.put(r, int(o))
    for l, o in zip(labs, loff):
        dl.put(l, int(o))
    dsize, lsize = int(doff[-1] + dbytes[-1]), int(loff[-1] + lsz[-1])
    for smp in (ds.UniformImageSampler(300, seed=2), ds.ClassWeightedImageSampler(300, 8, seed=2)):
        cap = smp.capacity()
        out, nout, crc = DevArray(4 * cap * len(labs)), DevArray(4 * len(labs)), DevArray(4 * len(labs))
        st = smp.sample_dev(len(labs), dd.p, dsize, doff, dbytes, dl.p, lsize, loff, w, h, out.p, nout.p, crc.p)
        assert st["crc_bytes"] == int(dbytes.sum()) and st["label_bytes"] == int(lsz.sum())
        assert st["total_ms"] > 0
        n_out = nout.get(len(labs), np.uint32)
        allv = out.get(cap * len(labs), np.uint32)
        want = smp.sample_batch(datas, labs)
        for i in range(len(labs)):
            assert np.array_equal(allv[i * cap:i * cap + n_out[i]], want[i]), i
        assert np.array_equal(crc.get(len(labs), np.uint32), [zlib.crc32(r.tobytes()) for r in raws])
        for a in (out, nout, crc):
            a.free()
    bad = loff.copy()
    bad[1] += 1
    with pytest.raises(pa.PdError, match="16-byte aligned"):
        ds.UniformImageSampler(10).sample_dev(len(labs), dd.p, dsize, doff, dbytes, dl.p, lsize, bad, w, h, dd.p,
                                              dd.p)
    # ranges are checked against the buffer sizes before any launch
    with pytest.raises(pa.PdError, match="outside its buffer"):
        ds.UniformImageSampler(10).sample_dev(len(labs), dd.p, dsize - 1, doff, dbytes, dl.p, lsize, loff, w, h,
                                              dd.p, dd.p)
    with pytest.raises(pa.PdError, match="outside its buffer"):
        ds.UniformImageSampler(10).sample_dev(len(labs), dd.p, dsize, doff, dbytes, dl.p, lsize - 16, loff, w, h,
                                              dd.p, dd.p)
    dd.free()
    dl.free()


@pytest.mark.parametrize("ch", [3, 4])
def test_rgb_to_labels_matches_oracle(gpu, ch):
    rng = np.random.default_rng(ch)
    m = rng.integers(0, 256, (40, 3)).astype(np.uint8)
    m[7] = m[3]   # repeated colour: the first class wins
    pick = rng.integers(0, 40, (97, 131))
    px = m[pick]
    stray = rng.random((97, 131)) < 0.2
    px[stray] = rng.integers(0, 256, (int(stray.sum()), 3))
    if ch == 4:
        px = np.concatenate([px, rng.integers(0, 256, (97, 131, 1)).astype(np.uint8)], -1)
    got = ds.rgb_to_labels(px, m)
    assert np.array_equal(got, oi.rgb_to_labels(px, m))
    assert not np.any(got == 8)   # class 7 duplicates class 3


def test_training_set_from_directory_trains_oracle_tree(gpu, tmp_path):
    ts0 = synth.training_set("c2", seed=4, n_images=6, samples=64)
    colours = np.array([[255, 0, 0], [0, 0, 255], [0, 255, 0], [255, 255, 0], [255, 0, 255], [0, 255, 255],
                        [128, 64, 32], [250, 250, 250]], np.uint8)
    for i, im in enumerate(ts0.images):
        depth = im.data.reshape(im.height, im.width)
        lab = im.labels.reshape(im.height, im.width)
        rgb = np.where(lab[..., None] > 0, colours[np.maximum(lab, 1) - 1], 0).astype(np.uint8)
        ds.write_png(str(tmp_path / f"f{i:03d}_depth.png"), depth)
        ds.write_png(str(tmp_path / f"f{i:03d}_labels.png"), rgb[..., ::-1])   # files hold B, G, R like cv::imwrite
    (tmp_path / "f999_depth.png").write_bytes(b"")   # no labels file: skipped
    ts = ds.training_set_from_directory(str(tmp_path), "_depth.png", "_labels.png", 8,
                                        ds.PNGImageLoader(np.uint16, 1), ds.RGBLabelsLoader(colours, 8),
                                        ds.UniformImageSampler(500, seed=9))
    assert len(ts.images) == 6
    acc = np.zeros(8)
    for im, im0 in zip(ts.images, ts0.images):
        assert np.array_equal(im.data.reshape(-1), im0.data.reshape(-1))
        assert np.array_equal(im.labels, im0.labels)
        want = oi.sample_uniform(im0.data.reshape(im0.height, im0.width), im0.labels.reshape(im0.height, im0.width),
                                 500, seed=9)
        assert np.array_equal(im.samples, want)
        acc += im.priors.astype(np.float64)
    assert ts.priors.tobytes() == (acc / 6).astype(np.float32).tobytes()
    p = synth.params_for("c2", n_features=128, n_thresholds=8)
    tree = pa.Tree(0, 8, pa.PD_FEAT_DEPTH_U16, 8)
    tr = pa.HIPTreeTrainer(pa.PD_FEAT_DEPTH_U16, 8)
    tr.train(tree, ts, p, 1, 8)
    out = orc.train_tree(orc.Tset(pa.PD_FEAT_DEPTH_U16, ts), p, ts.priors, 0, 8)
    assert np.array_equal(tree.left_children, out["left_children"])
    assert tree.features.tobytes() == out["features"].tobytes()
    assert tree.posteriors.tobytes() == out["posteriors"].tobytes()
    tr.close()
