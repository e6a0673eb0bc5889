"""Shared helpers for the parity tests: the oracle is the checker, the HIP
library is the thing tested."""
import numpy as np

from stipple_amd import synth


def fmt_of(img):
    import stipple_amd as sa
    c = img.shape[-1]
    if img.dtype == np.uint8:
        return sa.PIX_RGB8 if c == 3 else sa.PIX_RGBA8
    if img.dtype == np.uint16:
        return sa.PIX_RGB16 if c == 3 else sa.PIX_RGBA16
    return sa.PIX_RGBAF32


def to_device(img):
    import torch
    a = np.ascontiguousarray(img)
    return torch.from_numpy(a.reshape(-1).view(np.uint8).copy()).cuda()


def from_device(t, shape, dtype):
    return t.cpu().numpy().view(dtype).reshape(shape)


def oracle_tree(oracle, img, colors, assoc=None):
    """Octree from the oracle (as the reference's callers build ColorTree)."""
    f = synth.widen_q16(img) if img.dtype != np.float32 else img
    if assoc is None:
        assoc = img.shape[-1] == 4 and img.dtype != np.float32
    return oracle.tree_build(f, colors, assoc, synth.is_gray(img)), f


def out_alloc(shape_hw, fmt):
    import torch
    import stipple_amd as sa
    h, w = shape_hw
    nbytes = {sa.PIX_RGBAF32: 16, sa.PIX_RGB8: 3, sa.PIX_RGBA8: 4, sa.PIX_RGB16: 6,
              sa.PIX_RGBA16: 8, sa.PIX_INDEX8: 1, sa.PIX_INDEX16: 2}[fmt]
    return torch.zeros(h * w * nbytes, dtype=torch.uint8, device="cuda")


def scale_char_np(q):
    """ScaleQuantumToChar on a float32 array (host, for expected 8-bit outputs)."""
    q = np.asarray(q, np.float32)
    s = (q / np.float32(257.0)).astype(np.float32)
    r = np.where(np.isnan(q) | (q <= 0), 0, np.where(s >= 255, 255, (s + np.float32(0.5)).astype(np.uint32)))
    return r.astype(np.uint8)
