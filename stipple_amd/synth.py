"""Seeded synthetic inputs for the five measured workloads.

Every generator returns 8- or 16-bit integer samples (the wire format a
caller hands over) and never depends on anything outside numpy.  8-bit
samples widen to Q16 by x257 and 16-bit samples are used as-is, exactly as an
8-bit or 16-bit PNG decodes into a Q16 float image (widen_q16).
"""
from __future__ import annotations

import math

import numpy as np


def widen_q16(img: np.ndarray) -> np.ndarray:
    """uint8/uint16 (H, W, 3|4) -> float32 (H, W, 4) Q16; alpha 65535 if absent."""
    h, w, c = img.shape
    out = np.empty((h, w, 4), np.float32)
    scale = 257.0 if img.dtype == np.uint8 else 1.0
    out[..., :c] = img.astype(np.float32) * np.float32(scale)
    if c == 3:
        out[..., 3] = 65535.0
    return out


def is_gray(img: np.ndarray) -> bool:
    """The grayscale test rd_cli applies before ColorTree::Build (r == g == b)."""
    return bool(np.all(img[..., 0] == img[..., 1]) and np.all(img[..., 1] == img[..., 2]))


# -- C1: one 640x480 RGB8 ramp/sine image -----------------------------------
def ramp_image(w=640, h=480, sigma=8.0, seed=1) -> np.ndarray:
    rng = np.random.default_rng(seed)
    x = np.arange(w, dtype=np.float64)[None, :]
    y = np.arange(h, dtype=np.float64)[:, None]
    r = np.round(255.0 * x / max(w - 1, 1)) + 0 * y
    g = np.round(255.0 * np.sin(math.pi * y / max(h - 1, 1))) + 0 * x
    b = np.floor((r + g) / 2.0)
    img = np.stack([r, g, b], axis=-1)
    img = img + rng.normal(0.0, sigma, img.shape)
    return np.clip(np.round(img), 0, 255).astype(np.uint8)


# -- C2: a batch of uneven RGBA8 stills --------------------------------------
def sinusoid_image(w, h, rng, channels=4, sigma=6.0, rects=True) -> np.ndarray:
    """RGB = sum of 3 random-phase 2-D sinusoids per channel + N(0, sigma);
    alpha from 1-8 random opaque/transparent rectangles (when 4 channels).
    sin(ax + by + c) is evaluated separably as sin(ax+c)cos(by) + cos(ax+c)sin(by)."""
    x = np.arange(w, dtype=np.float64)
    y = np.arange(h, dtype=np.float64)
    planes = np.empty((h, w, 3), np.float32)
    for c in range(3):
        acc = np.zeros((h, w), np.float32)
        for _ in range(3):
            fx, fy = rng.uniform(-0.03, 0.03, 2)
            ph = rng.uniform(0, 2 * math.pi)
            ax = 2 * math.pi * fx * x + ph
            by = 2 * math.pi * fy * y
            acc += np.outer(np.cos(by), np.sin(ax)).astype(np.float32)
            acc += np.outer(np.sin(by), np.cos(ax)).astype(np.float32)
        planes[..., c] = 127.5 + 127.5 * acc / 3.0
    planes += rng.standard_normal((h, w, 3), dtype=np.float32) * np.float32(sigma)
    rgb = np.clip(np.rint(planes), 0, 255).astype(np.uint8)
    if channels == 3:
        return rgb
    alpha = np.full((h, w), 255, np.uint8)
    if rects:
        for _ in range(int(rng.integers(1, 9))):
            x0, x1 = np.sort(rng.integers(0, w + 1, 2))
            y0, y1 = np.sort(rng.integers(0, h + 1, 2))
            alpha[y0:y1, x0:x1] = 255 if rng.random() < 0.5 else 0
    return np.concatenate([rgb, alpha[..., None]], -1)


def still_batch(n=64, seed=2, max_side=2048):
    """[(img uint8 HxWx4, colours)] with w, h uniform in [1, max_side],
    including 1x1, 3x5 and 17x13; colours log-uniform on {2..256}."""
    rng = np.random.default_rng(seed)
    sizes = [(1, 1), (3, 5), (17, 13)]
    while len(sizes) < n:
        sizes.append((int(rng.integers(1, max_side + 1)), int(rng.integers(1, max_side + 1))))
    out = []
    for (w, h) in sizes[:n]:
        colours = int(round(math.exp(rng.uniform(math.log(2), math.log(256)))))
        colours = min(max(colours, 2), 256)
        out.append((sinusoid_image(w, h, rng), colours))
    return out


# -- C3: plasma video -------------------------------------------------------
def plasma_params(seed=3):
    rng = np.random.default_rng(seed)
    return {
        "freq": rng.uniform(0.002, 0.02, (3, 4, 2)),
        "phase": rng.uniform(0, 2 * math.pi, (3, 4)),
        "drift": 0.05,
    }


def plasma_frame(t, w=1920, h=1080, params=None, sigma=4.0, seed=3) -> np.ndarray:
    """Per channel 127.5 + 127.5 * mean of 4 sinusoids sin(2 pi (fx x + fy y) + phase
    + drift t), plus N(0, sigma); sin(a + b) is evaluated separably as
    sin(a) cos(b) + cos(a) sin(b)."""
    p = params or plasma_params(seed)
    rng = np.random.default_rng(seed * 100003 + t)
    x = np.arange(w, dtype=np.float64)
    y = np.arange(h, dtype=np.float64)
    planes = np.empty((h, w, 3), np.float32)
    for c in range(3):
        acc = np.zeros((h, w), np.float32)
        for k in range(4):
            fx, fy = p["freq"][c, k]
            a = 2 * math.pi * fx * x + p["phase"][c, k] + p["drift"] * t
            b = 2 * math.pi * fy * y
            acc += np.outer(np.cos(b), np.sin(a)).astype(np.float32)
            acc += np.outer(np.sin(b), np.cos(a)).astype(np.float32)
        planes[..., c] = 127.5 + 127.5 * acc / 4.0
    planes += rng.standard_normal((h, w, 3), dtype=np.float32) * np.float32(sigma)
    return np.clip(np.rint(planes), 0, 255).astype(np.uint8)


# -- C4: RGB16 radial gradients with salt and pepper -------------------------
def radial_frame16(t, w=3840, h=2160, seed=4, salt=0.05) -> np.ndarray:
    rng = np.random.default_rng(seed * 7919 + t)
    base = np.random.default_rng(seed)
    centres = base.uniform(0, 1, (3, 2)) * [w, h]
    x = np.arange(w, dtype=np.float64)[None, :]
    y = np.arange(h, dtype=np.float64)[:, None]
    diag = math.hypot(w, h)
    planes = []
    for c in range(3):
        cx, cy = centres[c] + 3.0 * t
        r = np.sqrt((x - cx) ** 2 + (y - cy) ** 2) / diag
        planes.append(65535.0 * np.clip(1.0 - r, 0, 1))
    img = np.stack(planes, -1)
    mask = rng.random((h, w)) < salt
    vals = np.where(rng.random((h, w)) < 0.5, 0.0, 65535.0)
    img[mask] = vals[mask][:, None]
    return np.clip(np.round(img), 0, 65535).astype(np.uint16)


def rgb8_to_yuv420(img: np.ndarray) -> np.ndarray:
    """An RGB8 frame (even sides) as planar 8-bit 4:2:0 bytes, the way an
    encoder would hand it over: limited-range BT.601 luma per pixel, chroma of
    each 2x2 block's mean colour.  Synthetic input only (the engines read the
    result; nothing here is on their path)."""
    h, w = img.shape[:2]
    rgb = img[..., :3].astype(np.int64)
    r, g, b = rgb[..., 0], rgb[..., 1], rgb[..., 2]
    y = ((66 * r + 129 * g + 25 * b + 128) >> 8) + 16
    m = rgb.reshape(h // 2, 2, w // 2, 2, 3).sum(axis=(1, 3)) // 4
    r2, g2, b2 = m[..., 0], m[..., 1], m[..., 2]
    u = ((-38 * r2 - 74 * g2 + 112 * b2 + 128) >> 8) + 128
    v = ((112 * r2 - 94 * g2 - 18 * b2 + 128) >> 8) + 128
    planes = [np.clip(p, 0, 255).astype(np.uint8).reshape(-1) for p in (y, u, v)]
    return np.concatenate(planes)


# -- serving mix: plasma + noise generated on the device ----------------------
def plasma_device(w, h, t=0, seed=5, sigma=4.0, device="cuda"):
    """An RGB8 plasma frame (the formula of plasma_frame: per channel 127.5 +
    127.5 * mean of 4 sinusoids with seeded frequencies in [0.002, 0.02]
    cycles/px, phases drifting 0.05 rad per frame t, plus N(0, sigma)), made
    with torch on `device` and returned there as h*w*3 contiguous uint8 --
    for input pools too large to generate on the host at serving rates."""
    import torch
    p = plasma_params(seed)
    g = torch.Generator(device=device)
    g.manual_seed(seed * 100003 + t)
    x = torch.arange(w, dtype=torch.float32, device=device)
    y = torch.arange(h, dtype=torch.float32, device=device)
    planes = torch.empty((h, w, 3), dtype=torch.float32, device=device)
    for c in range(3):
        acc = torch.zeros((h, w), dtype=torch.float32, device=device)
        for k in range(4):
            fx, fy = (float(v) for v in p["freq"][c, k])
            ph = float(p["phase"][c, k]) + p["drift"] * t
            acc += torch.sin(2 * math.pi * fy * y[:, None] + 2 * math.pi * fx * x[None, :] + ph)
        planes[..., c] = 127.5 + 127.5 * acc / 4.0
    planes += torch.randn((h, w, 3), generator=g, device=device) * sigma
    return planes.round_().clamp_(0, 255).to(torch.uint8).contiguous()


def radial_device(t, w=3840, h=2160, seed=4, salt=0.05, device="cuda"):
    """An RGB16 frame of radial_frame16's formula (three radial gradients from
    seeded centres drifting 3 px per frame t, then `salt` of the pixels set to
    0 or 65535 per pixel), made with torch on `device` as h*w*3 contiguous
    uint16 -- for frame counts too large to generate on the host."""
    import torch
    base = np.random.default_rng(seed)
    centres = base.uniform(0, 1, (3, 2)) * [w, h]
    g = torch.Generator(device=device)
    g.manual_seed(seed * 7919 + t)
    x = torch.arange(w, dtype=torch.float64, device=device)[None, :]
    y = torch.arange(h, dtype=torch.float64, device=device)[:, None]
    diag = math.hypot(w, h)
    img = torch.empty((h, w, 3), dtype=torch.float64, device=device)
    for c in range(3):
        cx, cy = (float(v) for v in centres[c] + 3.0 * t)
        r = torch.sqrt((x - cx) ** 2 + (y - cy) ** 2) / diag
        img[..., c] = 65535.0 * torch.clamp(1.0 - r, 0, 1)
    mask = torch.rand((h, w), generator=g, device=device) < salt
    vals = torch.where(torch.rand((h, w), generator=g, device=device) < 0.5, 0.0, 65535.0)
    img = torch.where(mask[..., None], vals[..., None].to(torch.float64), img)
    return img.round_().clamp_(0, 65535).to(torch.int32).to(torch.uint16).contiguous()
