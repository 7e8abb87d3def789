"""numpy restatements used to pin the oracle (independent of oracle.c)."""
import hashlib
import math
import struct

import numpy as np


def md5rand(seed):
    d = hashlib.md5(struct.pack("<4I", *[int(s) & 0xffffffff for s in seed])).digest()
    return list(struct.unpack("<4I", d))


def sample_f32(w, lo, hi):
    u = np.float32(np.uint32(w)) * np.float32(2.0 ** -32)
    span = np.float32(hi) - np.float32(lo)
    return np.float32(np.float32(lo) + np.float32(u * span))


def sample_s16(w, lo, hi):
    u = np.float32(np.uint32(w)) * np.float32(2.0 ** -32)
    span = int(hi) - int(lo)
    q = np.float32(u * np.float32(span))
    qi = int(math.trunc(float(q)))
    return np.int16(int(lo) + qi)


def params(fd, tree, node, f, lo, hi, short=False):
    seed = [tree, node, f, 0]
    out = []
    for j in range(fd):
        if j % 4 == 0:
            seed = md5rand(seed)
        out.append(sample_s16(seed[j % 4], lo[j], hi[j]) if short else sample_f32(seed[j % 4], lo[j], hi[j]))
    return np.array(out, dtype=np.int16 if short else np.float32)


def thresholds(nT, tree, node, f, lo, hi, short=False):
    seed = [tree, node, f, 1]
    out = []
    for t in range(nT):
        if t % 4 == 0:
            seed = md5rand(seed)
        out.append(sample_s16(seed[t % 4], lo, hi) if short else sample_f32(seed[t % 4], lo, hi))
    return np.array(out, dtype=np.int16 if short else np.float32)


def round_half_away(q):
    q = float(q)
    return math.copysign(math.floor(abs(q) + 0.5), q)


def _probe(img, x, y, d, px, py, bg, ch=0):
    qx = np.float32(np.float32(px) / np.float32(d))
    qy = np.float32(np.float32(py) / np.float32(d))
    rx, ry = round_half_away(qx), round_half_away(qy)
    if not (abs(rx) < 1048576.0 and abs(ry) < 1048576.0):
        return np.float32(bg)
    ox, oy = x + int(rx), y + int(ry)
    h, w = img.shape[0], img.shape[1]
    if ox < 0 or ox >= w or oy < 0 or oy >= h:
        return np.float32(bg)
    v = img[oy, ox] if img.ndim == 2 else img[oy, ox, ch]
    return np.float32(v)


def feature(kind, img, x, y, p):
    """Responses of the four hooks (see include/padenti/pd_features.hpp)."""
    if kind in (0, 1):
        bg = 10.0 if kind == 0 else 10000.0
        d = np.float32(img[y, x])
        if not d > 0:
            return np.float32(0.0)
        return np.float32(_probe(img, x, y, d, p[0], p[1], bg) - _probe(img, x, y, d, p[2], p[3], bg))
    if kind == 2:
        d = np.float32(img[y, x, 3])

        def chan(c):
            c = np.float32(c)
            if not c > 0:
                return 0
            if c >= 3:
                return 3
            return int(c)
        if not d > 0:
            return np.float32(0.0)
        a = _probe(img, x, y, d, p[0], p[1], 10000.0, chan(p[4]))
        b = _probe(img, x, y, d, p[2], p[3], 10000.0, chan(p[5]))
        return np.float32(a - b)
    if kind == 3:
        depth = np.int16(np.uint16(img[y, x]).astype(np.int16))
        if depth == 0:
            return np.int16(0)
        fd = np.float32(depth)
        qx = np.float32(np.float32(np.float32(p[0]) * np.float32(500.0)) / fd)
        qy = np.float32(np.float32(np.float32(p[1]) * np.float32(500.0)) / fd)
        rx, ry = round_half_away(qx), round_half_away(qy)
        h, w = img.shape
        if not (abs(rx) < 1048576.0 and abs(ry) < 1048576.0):
            d2 = 10000
        else:
            ox, oy = x + int(rx), y + int(ry)
            d2 = 10000 if (ox < 0 or ox >= w or oy < 0 or oy >= h) else int(np.uint16(img[oy, ox]).astype(np.int16))
        d2 = d2 if d2 else 10000
        return np.int16(np.int32(int(depth) - d2).astype(np.int16))
    raise ValueError(kind)


def entropy64(counts):
    counts = np.asarray(counts, dtype=np.float64)
    n = counts.sum()
    if n <= 0:
        return 0.0
    p = counts / n
    p = p[p > 1e-6]
    return float(-(p * np.log2(p)).sum())


def gain64(left, tot):
    left = np.asarray(left, dtype=np.float64)
    tot = np.asarray(tot, dtype=np.float64)
    right = tot - left
    n = tot.sum()
    return entropy64(tot) - (left.sum() / n * entropy64(left) + right.sum() / n * entropy64(right))
