"""Independent PNG encoder and a minimal decoder for the codec tests (plain
numpy + zlib, written from ISO/IEC 15948, not from the library's codec).

encode() takes samples exactly as stored in the file (R, G, B order, the
file's bit depth, gray values before any scaling) and exercises every
scanline filter (chosen per row) and, on request, Adam7 interlacing.
"""
import struct
import zlib

import numpy as np

ADAM7 = [(0, 0, 8, 8), (4, 0, 8, 8), (0, 4, 4, 8), (2, 0, 4, 4), (0, 2, 2, 4), (1, 0, 2, 2), (0, 1, 1, 2)]
CHANNELS = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}


def _chunk(t, data):
    return struct.pack(">I", len(data)) + t + data + struct.pack(">I", zlib.crc32(t + data) & 0xFFFFFFFF)


def _pack_row(samples, depth):
    """(w, ch) samples -> row bytes at bit depth `depth`."""
    flat = samples.reshape(-1).astype(np.uint32)
    if depth == 16:
        return flat.astype(">u2").tobytes()
    if depth == 8:
        return flat.astype(np.uint8).tobytes()
    per = 8 // depth
    nbytes = (flat.size * depth + 7) // 8
    out = bytearray(nbytes)
    for i, v in enumerate(flat):
        out[i // per] |= int(v) << (8 - depth * (i % per + 1))
    return bytes(out)


def _paeth(a, b, c):
    p = a + b - c
    pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
    if pa <= pb and pa <= pc:
        return a
    return b if pb <= pc else c


def _filter(rows, bpp, first_filter=0):
    out = bytearray()
    prev = bytes(len(rows[0])) if rows else b""
    for y, row in enumerate(rows):
        ft = (first_filter + y) % 5
        r = bytearray(len(row))
        for i in range(len(row)):
            a = row[i - bpp] if i >= bpp else 0
            b = prev[i] if y > 0 else 0
            c = prev[i - bpp] if (y > 0 and i >= bpp) else 0
            pred = [0, a, b, (a + b) // 2, _paeth(a, b, c)][ft]
            r[i] = (row[i] - pred) & 0xFF
        out.append(ft)
        out += r
        prev = row
    return bytes(out)


def encode(samples, color_type, depth, interlace=False, palette=None, trns=None, first_filter=0):
    """samples: (h, w, channels of the colour type) integer array."""
    s = np.asarray(samples)
    if s.ndim == 2:
        s = s[:, :, None]
    h, w, ch = s.shape
    assert ch == CHANNELS[color_type]
    bpp = max(1, ch * depth // 8)
    passes = ADAM7 if interlace else [(0, 0, 1, 1)]
    raw = b""
    for k, (x0, y0, dx, dy) in enumerate(passes):
        sub = s[y0::dy, x0::dx]
        if sub.shape[0] == 0 or sub.shape[1] == 0:
            continue
        rows = [_pack_row(sub[y], depth) for y in range(sub.shape[0])]
        raw += _filter(rows, bpp, first_filter + k)
    out = b"\x89PNG\r\n\x1a\n"
    out += _chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, depth, color_type, 0, 0, 1 if interlace else 0))
    out += _chunk(b"tEXt", b"Comment\0ancillary chunk")
    if palette is not None:
        out += _chunk(b"PLTE", np.asarray(palette, dtype=np.uint8).reshape(-1).tobytes())
    if trns is not None:
        out += _chunk(b"tRNS", bytes(trns))
    z = zlib.compress(raw, 9)
    mid = len(z) // 2   # two IDAT chunks
    out += _chunk(b"IDAT", z[:mid]) + _chunk(b"IDAT", z[mid:])
    out += _chunk(b"IEND", b"")
    return out


def decode_simple(data):
    """Decoder for what the library writes: non-interlaced, filter 0 rows.
    Returns (h, w, ch) samples in file order (R, G, B, A) and the depth."""
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    o, idat = 8, b""
    while o < len(data):
        n, = struct.unpack(">I", data[o:o + 4])
        t = data[o + 4:o + 8]
        body = data[o + 8:o + 8 + n]
        assert zlib.crc32(t + body) & 0xFFFFFFFF == struct.unpack(">I", data[o + 8 + n:o + 12 + n])[0]
        if t == b"IHDR":
            w, h, depth, ct, _, _, il = struct.unpack(">IIBBBBB", body)
        elif t == b"IDAT":
            idat += body
        o += 12 + n
    assert il == 0
    ch = CHANNELS[ct]
    raw = zlib.decompress(idat)
    rb = w * ch * depth // 8
    rows = []
    for y in range(h):
        r = raw[y * (rb + 1):(y + 1) * (rb + 1)]
        assert r[0] == 0
        rows.append(np.frombuffer(r[1:], dtype=">u2" if depth == 16 else np.uint8))
    return np.stack(rows).reshape(h, w, ch).astype(np.uint16 if depth == 16 else np.uint8), depth
