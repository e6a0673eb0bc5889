"""Minimal PNG decoder (non-interlaced, bit depths 1-16, colour types 0/2/3/4/6)
for reading the reference's own test fixtures without an image library."""
import struct
import zlib

import numpy as np


def _unfilter(raw, h, stride, bpp):
    out = np.zeros((h, stride), np.uint8)
    prev = np.zeros(stride, np.int32)
    pos = 0
    for y in range(h):
        ft = raw[pos]
        line = np.frombuffer(raw, np.uint8, stride, pos + 1).astype(np.int32)
        pos += 1 + stride
        cur = np.zeros(stride, np.int32)
        if ft == 0:
            cur = line
        elif ft == 2:
            cur = (line + prev) & 255
        else:
            for i in range(stride):
                a = cur[i - bpp] if i >= bpp else 0
                b = prev[i]
                c = prev[i - bpp] if i >= bpp else 0
                if ft == 1:
                    p = a
                elif ft == 3:
                    p = (a + b) >> 1
                else:
                    pa, pb, pc = abs(b - c), abs(a - c), abs(a + b - 2 * c)
                    p = a if (pa <= pb and pa <= pc) else (b if pb <= pc else c)
                cur[i] = (line[i] + p) & 255
        out[y] = cur
        prev = cur
    return out


def read_png(path):
    """Returns (samples HxWxC, bit_depth, colour_type); palette images are
    expanded to RGB(A) 8-bit."""
    data = open(path, "rb").read()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, idat, plte, trns = 8, b"", None, None
    while pos < len(data):
        n, = struct.unpack(">I", data[pos:pos + 4])
        typ = data[pos + 4:pos + 8]
        body = data[pos + 8:pos + 8 + n]
        pos += 12 + n
        if typ == b"IHDR":
            w, h, depth, ctype, _, _, interlace = struct.unpack(">IIBBBBB", body)
            assert interlace == 0
        elif typ == b"PLTE":
            plte = np.frombuffer(body, np.uint8).reshape(-1, 3)
        elif typ == b"tRNS":
            trns = np.frombuffer(body, np.uint8)
        elif typ == b"IDAT":
            idat += body
    ch = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}[ctype]
    bits = ch * depth
    stride = (w * bits + 7) // 8
    bpp = max(1, bits // 8)
    rows = _unfilter(zlib.decompress(idat), h, stride, bpp)
    if depth == 16:
        s = rows.reshape(h, -1).view(">u2").astype(np.uint16).reshape(h, w, ch)
    elif depth == 8:
        s = rows.reshape(h, w, ch)
    else:
        bitsarr = np.unpackbits(rows, axis=1)[:, : w * depth].reshape(h, w, depth)
        s = np.zeros((h, w), np.uint16)
        for k in range(depth):
            s = (s << 1) | bitsarr[..., k]
        s = s.reshape(h, w, 1)
    if ctype == 3:
        idx = s[..., 0].astype(np.int64)
        rgb = plte[idx]
        if trns is not None:
            a = np.full(len(plte), 255, np.uint8)
            a[: len(trns)] = trns
            return np.concatenate([rgb, a[idx][..., None]], -1), 8, ctype
        return rgb, 8, ctype
    return s, depth, ctype


def to_q16_rgbaf(samples, depth, ctype):
    """What IM's decoder hands the engine: Q16 floats, grey expanded, alpha
    65535 when absent."""
    s = samples.astype(np.float32)
    if depth == 8:
        s = s * np.float32(257.0)
    elif depth < 8:
        s = s * np.float32(65535.0 / ((1 << depth) - 1))
    h, w, c = s.shape
    out = np.empty((h, w, 4), np.float32)
    if c in (1, 2):
        out[..., 0] = out[..., 1] = out[..., 2] = s[..., 0]
        out[..., 3] = s[..., 1] if c == 2 else 65535.0
    else:
        out[..., :3] = s[..., :3]
        out[..., 3] = s[..., 3] if c == 4 else 65535.0
    return out, c in (2, 4)
