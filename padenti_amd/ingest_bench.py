"""Dataset-ingestion benchmark: sampling a large training set resident in HBM.

    python -m padenti_amd.ingest_bench [--images 5000] [--w 640] [--h 480] [--reps 5]

Workload: synthetic 640x480 uint16 depth images (the blob scene, 16 classes)
and their label images, already in device memory; every image is hashed
(CRC-32 of its bytes), then sampled by
  uniform:        UniformImageSampler(2000, seed) - 2000 draws per image;
  class-weighted: ClassWeightedImageSampler(2000, 16, seed) - a reservoir of
                  ceilf(2000 / 16) = 125 pixels per class and image,
and the samples are sorted per image (pd_sample_images_dev, one call per
sampler).  Reports images/s, the per-stage times (HIP events) and the HBM
roofline of the CRC kernel (algorithmic bytes = the image bytes) and of the
whole call (image + label bytes).  The host form (pd_sample_images: host
buffers in, host sample lists out, PCIe included) is timed once beside it.
"""
import argparse
import ctypes as C
import json
import sys
import time

import numpy as np

from . import _lib as L
from . import dataset as ds
from . import synth

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parser():
    ap = argparse.ArgumentParser()
    ap.add_argument("--images", type=int, default=5000)
    ap.add_argument("--w", type=int, default=640)
    ap.add_argument("--h", type=int, default=480)
    ap.add_argument("--classes", type=int, default=16)
    ap.add_argument("--samples", type=int, default=2000)
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--seed", type=int, default=0)
    return ap


def make_dataset(n, w, h, n_classes, seed):
    """(depth images [n][h][w] uint16, label images [n][h][w] uint8)."""
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(max_workers=16) as ex:
        parts = list(ex.map(lambda i: synth.gen_image(L.PD_SYN_BLOBS_U16, seed, i, w, h, n_classes), range(n)))
    return [p[0] for p in parts], [p[1] for p in parts]


class _Dev:
    def __init__(self, nbytes):
        self.p = C.c_void_p()
        L.check(L.lib().pd_dev_alloc(0, max(int(nbytes), 16), C.byref(self.p)))

    def free(self):
        L.check(L.lib().pd_dev_free(self.p))


def run(a, datas=None, labs=None):
    if datas is None:
        t0 = time.perf_counter()
        datas, labs = make_dataset(a.images, a.w, a.h, a.classes, a.seed)
        log(f"[ingest] {len(datas)} images {a.w}x{a.h} generated in {time.perf_counter() - t0:.1f} s")
    n = len(datas)
    npx = a.w * a.h
    img_bytes = npx * 2
    dd, dl = _Dev(n * img_bytes), _Dev(n * npx)
    # one copy each (images and label images packed back to back; 640x480 keeps 16-byte alignment);
    # the packed host arrays stay referenced until the copies return
    packed = np.ascontiguousarray(np.stack(datas))
    L.check(L.lib().pd_memcpy_h2d(dd.p, L.ptr(packed), n * img_bytes))
    packed = np.ascontiguousarray(np.stack(labs))
    L.check(L.lib().pd_memcpy_h2d(dl.p, L.ptr(packed), n * npx))
    del packed
    doff = (np.arange(n, dtype=np.uint64) * img_bytes).astype(np.uint64)
    dbytes = np.full(n, img_bytes, np.uint64)
    loff = (np.arange(n, dtype=np.uint64) * npx).astype(np.uint64)
    w = np.full(n, a.w, np.uint32)
    h = np.full(n, a.h, np.uint32)
    out = {"workload": f"{n} x {a.w}x{a.h} uint16 depth images + uint8 labels, {a.classes} classes, in HBM",
           "images": n, "image_bytes": n * img_bytes, "label_bytes": n * npx}
    for name, smp in (("uniform", ds.UniformImageSampler(a.samples, seed=7)),
                      ("class_weighted", ds.ClassWeightedImageSampler(a.samples, a.classes, seed=7))):
        cap = smp.capacity()
        ds_out, dn, dc = _Dev(4 * n * cap), _Dev(4 * n), _Dev(4 * n)
        args = (n, dd.p, n * img_bytes, doff, dbytes, dl.p, n * npx, loff, w, h, ds_out.p, dn.p, dc.p)
        smp.sample_dev(*args)   # warm-up
        runs = []
        for _ in range(a.reps):
            t = time.perf_counter()
            st = smp.sample_dev(*args)
            st["wall_ms"] = (time.perf_counter() - t) * 1e3
            runs.append(st)
        med = {k: float(np.median([r[k] for r in runs])) for k in ("crc_ms", "index_ms", "sample_ms", "sort_ms",
                                                                  "total_ms", "wall_ms")}
        n_out = np.empty(n, np.uint32)
        L.check(L.lib().pd_memcpy_d2h(L.ptr(n_out), dn.p, 4 * n))
        crc_gbs = n * img_bytes / (med["crc_ms"] * 1e6)
        all_gbs = (n * img_bytes + n * npx) / (med["total_ms"] * 1e6)
        out[name] = {"images_per_s": n / (med["total_ms"] / 1e3), "samples": int(n_out.sum()), "ms": med,
                     "crc_roofline": {"bound": "hbm", "achieved": crc_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                      "frac": crc_gbs / HBM_PEAK_GBS},
                     "call_gbs": all_gbs}
        log(f"[ingest] {name}: {out[name]['images_per_s']:.0f} images/s (total {med['total_ms']:.2f} ms: crc "
            f"{med['crc_ms']:.2f} index {med['index_ms']:.2f} sample {med['sample_ms']:.2f} sort "
            f"{med['sort_ms']:.2f}); CRC {crc_gbs:.0f} GB/s")
        for x in (ds_out, dn, dc):
            x.free()
    dd.free()
    dl.free()
    # host form once (PCIe both ways, staging included)
    smp = ds.UniformImageSampler(a.samples, seed=7)
    k = min(n, 1000)
    t = time.perf_counter()
    smp.sample_batch(datas[:k], labs[:k])
    host_s = time.perf_counter() - t
    out["host_path_images_per_s"] = k / host_s
    out["host_path_images"] = k
    log(f"[ingest] host path (PCIe, {k} images): {k / host_s:.0f} images/s")
    return out


def main():
    print(json.dumps(run(parser().parse_args())), flush=True)


if __name__ == "__main__":
    main()
