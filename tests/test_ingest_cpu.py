"""Host parts of the ingestion path, on the CPU: the PNG codec against an
independent encoder / decoder (tests/pngenc.py), the directory pairing, the
sampler capacity and the loaders' type checks.  Nothing here launches a
kernel."""
import os

import numpy as np
import pytest

from padenti_amd import dataset as ds
from padenti_amd._lib import PdError

from . import pngenc


def _roundtrip(tmp_path, name, png_bytes):
    p = tmp_path / name
    p.write_bytes(png_bytes)
    return ds.read_png(str(p))


@pytest.mark.parametrize("interlace", [False, True])
@pytest.mark.parametrize("depth", [1, 2, 4, 8, 16])
def test_png_gray(tmp_path, depth, interlace):
    rng = np.random.default_rng(depth)
    h, w = 13, 11   # odd sizes: every Adam7 pass and partial bytes
    s = rng.integers(0, 2 ** depth, (h, w))
    got = _roundtrip(tmp_path, "g.png", pngenc.encode(s, 0, depth, interlace, first_filter=depth))
    scale = {1: 255, 2: 85, 4: 17, 8: 1, 16: 1}[depth]   # libpng's low-depth gray expansion
    assert got.dtype == (np.uint16 if depth == 16 else np.uint8) and got.shape == (h, w)
    assert np.array_equal(got.astype(np.int64), s * scale)


@pytest.mark.parametrize("interlace", [False, True])
@pytest.mark.parametrize("depth", [8, 16])
@pytest.mark.parametrize("ct", [2, 4, 6])
def test_png_colour(tmp_path, ct, depth, interlace):
    rng = np.random.default_rng(ct * 100 + depth)
    h, w = 9, 17
    ch = pngenc.CHANNELS[ct]
    s = rng.integers(0, 2 ** depth, (h, w, ch))
    got = _roundtrip(tmp_path, "c.png", pngenc.encode(s, ct, depth, interlace, first_filter=ct))
    if ct == 2:
        exp = s[..., ::-1]                                   # R, G, B -> B, G, R
    elif ct == 4:
        exp = np.stack([s[..., 0]] * 3 + [s[..., 1]], -1)    # gray + alpha -> B, G, R, A
    else:
        exp = s[..., [2, 1, 0, 3]]                          # R, G, B, A -> B, G, R, A
    assert got.shape == exp.shape and np.array_equal(got.astype(np.int64), exp)


@pytest.mark.parametrize("depth", [1, 2, 4, 8])
def test_png_palette(tmp_path, depth):
    rng = np.random.default_rng(depth + 50)
    n_pal = min(2 ** depth, 40)
    pal = rng.integers(0, 256, (n_pal, 3))
    idx = rng.integers(0, n_pal, (7, 23))
    got = _roundtrip(tmp_path, "p.png", pngenc.encode(idx, 3, depth, palette=pal))
    assert np.array_equal(got, pal[idx][..., ::-1].astype(np.uint8))
    trns = rng.integers(0, 256, n_pal // 2 + 1)               # shorter than the palette: the rest opaque
    got = _roundtrip(tmp_path, "pt.png", pngenc.encode(idx, 3, depth, palette=pal, trns=trns.tolist()))
    alpha = np.where(idx < trns.size, trns[np.minimum(idx, trns.size - 1)], 255)
    assert np.array_equal(got, np.concatenate([pal[idx][..., ::-1], alpha[..., None]], -1).astype(np.uint8))


def test_png_rgb_trns_key(tmp_path):
    s = np.array([[[1, 2, 3], [4, 5, 6]], [[1, 2, 3], [0, 0, 0]]])
    got = _roundtrip(tmp_path, "k.png", pngenc.encode(s, 2, 8, trns=[0, 1, 0, 2, 0, 3]))
    assert got[..., 3].tolist() == [[0, 255], [0, 255]]
    assert np.array_equal(got[..., :3], s[..., ::-1].astype(np.uint8))


@pytest.mark.parametrize("ch,dt", [(1, np.uint8), (1, np.uint16), (3, np.uint8), (3, np.uint16), (4, np.uint8),
                                   (4, np.uint16)])
def test_png_write_roundtrip(tmp_path, ch, dt):
    rng = np.random.default_rng(ch)
    shape = (19, 7) if ch == 1 else (19, 7, ch)
    a = rng.integers(0, np.iinfo(dt).max + 1, shape).astype(dt)
    p = str(tmp_path / "w.png")
    ds.write_png(p, a)
    assert np.array_equal(ds.read_png(p), a)
    file_samples, depth = pngenc.decode_simple(open(p, "rb").read())   # independent reader: file order R, G, B, A
    exp = a[..., None] if ch == 1 else (a[..., ::-1] if ch == 3 else a[..., [2, 1, 0, 3]])
    assert depth == 8 * np.dtype(dt).itemsize and np.array_equal(file_samples, exp)


def test_png_rejects_corruption(tmp_path):
    good = pngenc.encode(np.zeros((4, 4), np.uint8), 0, 8)
    bad = bytearray(good)
    bad[40] ^= 0xFF   # inside a chunk: CRC mismatch
    p = tmp_path / "bad.png"
    p.write_bytes(bytes(bad))
    with pytest.raises(PdError):
        ds.read_png(str(p))
    p.write_bytes(b"not a png at all")
    with pytest.raises(PdError):
        ds.read_png(str(p))
    with pytest.raises(PdError):
        ds.read_png(str(tmp_path / "missing.png"))


def test_list_pairs_by_suffix(tmp_path):
    for n in ["b_depth.png", "b_labels.png", "a_depth.png", "a_labels.png", "c_depth.png", "_depth.png",
              "_labels.png", "x.png", "d_depth.png.bak"]:
        (tmp_path / n).write_bytes(b"")
    (tmp_path / "e_depth.png").mkdir()
    (tmp_path / "e_labels.png").write_bytes(b"")
    pairs = ds.list_pairs(str(tmp_path), "_depth.png", "_labels.png")
    names = [(os.path.basename(a), os.path.basename(b)) for a, b in pairs]
    # c has no labels file, d does not end with the suffix, e is a directory
    assert names == [("_depth.png", "_labels.png"), ("a_depth.png", "a_labels.png"), ("b_depth.png", "b_labels.png")]
    # the reference's own example: data "_depth.png", labels ".png"
    (tmp_path / "a.png").write_bytes(b"")
    names = [os.path.basename(a) for a, _ in ds.list_pairs(str(tmp_path), "_depth.png", ".png")]
    assert names == ["a_depth.png"]
    with pytest.raises(PdError):
        ds.list_pairs(str(tmp_path / "nope"), "_depth.png", ".png")


def test_sampler_capacity_host():
    assert ds.UniformImageSampler(2048).capacity() == 2048
    assert ds.ClassWeightedImageSampler(100, 3).capacity() == 102           # 3 x ceilf(33.333336f)
    assert ds.ClassWeightedImageSampler(10, 1, [0.3]).capacity() == 3       # float, not double, product
    assert ds.ClassWeightedImageSampler(2000, 2, [0.25, 0.0]).capacity() == 500
    with pytest.raises(PdError):
        ds.ClassWeightedImageSampler(10, 2, [0.5, -1.0]).capacity()
    with pytest.raises(ValueError):
        ds.ClassWeightedImageSampler(10, 2, [0.5])


def test_loader_type_checks(tmp_path):
    p8 = str(tmp_path / "g8.png")
    ds.write_png(p8, np.zeros((3, 3), np.uint8))
    p16 = str(tmp_path / "g16.png")
    ds.write_png(p16, np.zeros((3, 3), np.uint16))
    assert ds.PNGImageLoader(np.uint16, 1).load(p16).dtype == np.uint16
    with pytest.raises(ValueError, match="Image type different"):
        ds.PNGImageLoader(np.uint16, 1).load(p8)
    with pytest.raises(ValueError, match="CV_8UC3 or CV_8UC4"):
        ds.RGBLabelsLoader([[255, 0, 0]]).load(p8)


def test_device_entry_points_check_ranges_before_launch():
    """A byte range past the caller's buffer is an argument error, raised on
    the host before any device work (no GPU needed to reach it)."""
    import ctypes as C
    from padenti_amd._lib import lib, ptr
    off = np.array([0, 100], np.uint64)
    n = np.array([100, 29], np.uint64)
    ms = C.c_float()
    fake = C.c_void_p(0x10000)   # never dereferenced: the check fails first
    rc = lib().pd_crc32_images_dev(0, 2, fake, 128, ptr(off), ptr(n), fake, C.byref(ms))
    assert rc == -1 and b"outside its buffer" in lib().pd_last_error()
    smp = ds.UniformImageSampler(4)
    w = np.array([10, 1], np.uint32)
    h = np.array([10, 29], np.uint32)
    with pytest.raises(PdError, match="image data range 1 lies outside"):
        smp.sample_dev(2, fake, 128, off, n, fake, 1000, np.array([0, 112], np.uint64), w, h, fake, fake)
    with pytest.raises(PdError, match="label image range 1 lies outside"):
        smp.sample_dev(2, fake, 129, off, n, fake, 140, np.array([0, 112], np.uint64), w, h, fake, fake)
