"""Dataset ingestion: image loaders, the two pixel samplers and the
directory-built training set, mirroring the reference's

  ImageLoader / CVImageLoader        src/padenti/image_loader.hpp, cv_image_loader_impl.hpp:36-72
  CVRGBLabelsLoader (+ _rgb2label)    src/padenti/cv_image_loader_impl.hpp:78-133, 235-267
  UniformImageSampler                 src/padenti/uniform_image_sampler_impl.hpp:27-63
  ClassWeightedImageSampler           src/padenti/classweighted_image_sampler_impl.hpp:26-123
  TrainingSet(dir, suffixes, ...)     src/padenti/training_set_impl.hpp:43-133

Images are read by the library's PNG codec (host) in OpenCV's channel order
(B, G, R[, A]); label mapping, CRC-32 and sampling run on the GPU and give
the reference's sample lists (MT19937 seeded with crc32(image) ^ seed).
"""
import ctypes as C
import os

import numpy as np

from . import TrainingSet, TrainingSetImage
from ._lib import (PD_SAMPLER_CLASS_WEIGHTED, PD_SAMPLER_UNIFORM, IngestStats, PdError, PngInfo, Sampler, check,
                   lib, ptr)

__all__ = ["read_png", "write_png", "list_pairs", "PNGImageLoader", "RGBLabelsLoader", "UniformImageSampler",
           "ClassWeightedImageSampler", "training_set_from_directory", "rgb_to_labels"]


def read_png(path):
    """cv::imread(path, -1) for PNG: (h, w) or (h, w, c) uint8 / uint16, B, G, R(, A)."""
    info = PngInfo()
    p = os.fsencode(path)
    check(lib().pd_png_read(p, C.byref(info), None, 0))
    dt = np.uint8 if info.bit_depth == 8 else np.uint16
    shape = (info.height, info.width) if info.channels == 1 else (info.height, info.width, info.channels)
    out = np.empty(shape, dtype=dt)
    check(lib().pd_png_read(p, C.byref(info), ptr(out), out.nbytes))
    return out


def write_png(path, img):
    """cv::imwrite for PNG: (h, w) or (h, w, 3|4) uint8 / uint16 in B, G, R(, A) order."""
    a = np.ascontiguousarray(img)
    if a.dtype not in (np.uint8, np.uint16):
        raise ValueError("PNG pixels must be uint8 or uint16")
    ch = 1 if a.ndim == 2 else a.shape[2]
    check(lib().pd_png_write(os.fsencode(path), a.shape[1], a.shape[0], ch, 8 * a.itemsize, ptr(a)))


def list_pairs(ts_path, data_suffix, labels_suffix):
    """(data path, labels path) pairs of a training-set directory, by file name."""
    n = C.c_uint32()
    need = C.c_uint64()
    args = (os.fsencode(ts_path), os.fsencode(data_suffix), os.fsencode(labels_suffix))
    check(lib().pd_list_pairs(*args, None, 0, C.byref(n), C.byref(need)))
    buf = C.create_string_buffer(max(need.value, 1))
    check(lib().pd_list_pairs(*args, buf, need.value, C.byref(n), C.byref(need)))
    parts = buf.raw[:need.value].split(b"\0")[:-1]
    return [(os.fsdecode(parts[2 * i]), os.fsdecode(parts[2 * i + 1])) for i in range(n.value)]


class PNGImageLoader:
    """CVImageLoader<type, nChannels>: the file must decode to exactly this
    element type and channel count ("Image type different from expected one")."""

    def __init__(self, dtype, channels):
        self.dtype = np.dtype(dtype)
        self.channels = int(channels)

    def load(self, path):
        img = read_png(path)
        ch = 1 if img.ndim == 2 else img.shape[2]
        if img.dtype != self.dtype or ch != self.channels:
            raise ValueError("Image type different from expected one")
        return img


def rgb_to_labels(rgb, rgb2label, device=0):
    """_rgb2label on the GPU: (h, w, 3|4) R, G, B(, A) uint8 -> (h, w) labels."""
    a = np.ascontiguousarray(rgb, dtype=np.uint8)
    m = np.ascontiguousarray(rgb2label, dtype=np.uint8).reshape(-1, 3)
    out = np.empty(a.shape[:2], dtype=np.uint8)
    check(lib().pd_rgb_to_labels(device, out.size, a.shape[2], ptr(a), ptr(m), m.shape[0], ptr(out)))
    return out


class RGBLabelsLoader:
    """CVRGBLabelsLoader(rgb2labelMap, nClasses): an 8-bit 3- or 4-channel
    PNG whose colours map to labels 1..n_classes (0 when unmatched)."""

    def __init__(self, rgb2label, n_classes=None, device=0):
        m = np.ascontiguousarray(rgb2label, dtype=np.uint8).reshape(-1, 3)
        self.map = m[:n_classes] if n_classes is not None else m
        self.device = device

    def load(self, path):
        img = read_png(path)
        if img.dtype != np.uint8 or img.ndim != 3 or img.shape[2] not in (3, 4):
            raise ValueError("Labels image type must be CV_8UC3 or CV_8UC4")
        rgb = img[..., [2, 1, 0]]   # cv::cvtColor BGR(A) -> RGB(A), alpha unused
        return rgb_to_labels(rgb, self.map, self.device)


class _Sampler:
    def __init__(self, kind, n_samples, seed, n_classes=0, weights=None, device=0):
        self.kind = kind
        self.n_samples = int(n_samples)
        self.seed = int(seed) & 0xFFFFFFFF
        self.n_classes = int(n_classes)
        self.weights = None if weights is None else np.ascontiguousarray(weights, dtype=np.float32)
        if self.weights is not None and self.weights.size != self.n_classes:
            raise ValueError("one weight per class")
        self.device = device
        self.last_stats = None

    def _desc(self):
        s = Sampler()
        s.kind, s.n_samples, s.seed, s.n_classes = self.kind, self.n_samples, self.seed, self.n_classes
        s.weights = None if self.weights is None else self.weights.ctypes.data
        return s

    def capacity(self):
        c = C.c_uint32()
        check(lib().pd_sampler_capacity(C.byref(self._desc()), C.byref(c)))
        return c.value

    def sample(self, data, labels):
        """sample(data, labels, w, h, samples): sorted pixel indices y * w + x."""
        return self.sample_batch([data], [labels])[0]

    def sample_batch(self, datas, labels_list, crc=False):
        """All images in one device call; list of sample arrays (and the CRCs)."""
        n = len(datas)
        if n == 0:
            return ([], np.zeros(0, np.uint32)) if crc else []
        raws = [np.ascontiguousarray(d).view(np.uint8).reshape(-1) for d in datas]
        labs = [np.ascontiguousarray(l, dtype=np.uint8) for l in labels_list]
        data = np.concatenate(raws)
        dbytes = np.array([r.size for r in raws], dtype=np.uint64)
        doff = np.concatenate([[0], np.cumsum(dbytes)[:-1]]).astype(np.uint64)
        lab = np.concatenate([l.reshape(-1) for l in labs])
        lsz = np.array([l.size for l in labs], dtype=np.uint64)
        loff = np.concatenate([[0], np.cumsum(lsz)[:-1]]).astype(np.uint64)
        w = np.array([l.shape[1] for l in labs], dtype=np.uint32)
        h = np.array([l.shape[0] for l in labs], dtype=np.uint32)
        cap = self.capacity()
        out = np.empty(max(n * cap, 1), dtype=np.uint32)
        n_out = np.empty(n, dtype=np.uint32)
        crcs = np.empty(n, dtype=np.uint32)
        desc = self._desc()
        check(lib().pd_sample_images(self.device, C.byref(desc), n, ptr(data), ptr(doff), ptr(dbytes), ptr(lab),
                                     ptr(loff), ptr(w), ptr(h), ptr(out), ptr(n_out), ptr(crcs)))
        res = [out[i * cap:i * cap + n_out[i]].copy() for i in range(n)]
        return (res, crcs) if crc else res

    def sample_dev(self, n, d_data, data_size, data_off, data_bytes, d_labels, labels_size, lab_off, widths, heights,
                   d_samples, d_n_out, d_crc=None):
        """Device-resident form (pd_sample_images_dev): d_* are device
        pointers, data_size / labels_size the bytes of those buffers, the
        rest host arrays (uint64 offsets and byte counts, uint32 sizes).
        Returns the stage timings."""
        st = IngestStats()
        desc = self._desc()
        arrs = [np.ascontiguousarray(a, dtype=t) for a, t in ((data_off, np.uint64), (data_bytes, np.uint64),
                                                             (lab_off, np.uint64), (widths, np.uint32),
                                                             (heights, np.uint32))]
        for a in arrs:
            if a.size != n:
                raise ValueError("one offset / size per image")
        check(lib().pd_sample_images_dev(self.device, C.byref(desc), n, d_data, int(data_size), ptr(arrs[0]),
                                         ptr(arrs[1]), d_labels, int(labels_size), ptr(arrs[2]), ptr(arrs[3]),
                                         ptr(arrs[4]), d_samples, d_n_out, d_crc, C.byref(st)))
        self.last_stats = {f: getattr(st, f) for f, _ in IngestStats._fields_}
        return self.last_stats


class UniformImageSampler(_Sampler):
    """UniformImageSampler(nSamples, seed=0): n_samples draws with replacement
    among the labelled pixels, sorted."""

    def __init__(self, n_samples, seed=0, device=0):
        super().__init__(PD_SAMPLER_UNIFORM, n_samples, seed, 0, None, device)


class ClassWeightedImageSampler(_Sampler):
    """ClassWeightedImageSampler(nSamples, nClasses[, weights], seed=0): per
    class a reservoir of ceilf(w_c * n_samples) pixels (w_c = 1/nClasses by
    default)."""

    def __init__(self, n_samples, n_classes, weights=None, seed=0, device=0):
        super().__init__(PD_SAMPLER_CLASS_WEIGHTED, n_samples, seed, n_classes, weights, device)


def training_set_from_directory(ts_path, data_suffix, labels_suffix, n_classes, data_loader, labels_loader,
                                sampler):
    """TrainingSet(tsPath, dataSuffix, labelsSuffix, nClasses, dataLoader,
    labelsLoader, sampler): pairs by suffix, skips pairs whose sizes differ,
    samples every image (one device call) and sets the priors to the FP64
    mean of the per-image priors cast to float (training_set_impl.hpp:113-129).
    Pairs are taken in file-name order."""
    datas, labs = [], []
    for dpath, lpath in list_pairs(ts_path, data_suffix, labels_suffix):
        d = data_loader.load(dpath)
        l = labels_loader.load(lpath)
        if d.shape[:2] != l.shape[:2]:
            continue   # the reference logs a warning and skips the pair
        datas.append(d)
        labs.append(l)
    ts = TrainingSet(n_classes)
    samples = sampler.sample_batch(datas, labs) if datas else []
    for d, l, s in zip(datas, labs, samples):
        ts.images.append(TrainingSetImage(d, l.shape[1], l.shape[0], l, n_classes, s))
    if ts.images:
        acc = np.zeros(n_classes, dtype=np.float64)
        for im in ts.images:
            acc += im.priors.astype(np.float64)
        ts.priors = (acc / len(ts.images)).astype(np.float32)
    return ts
