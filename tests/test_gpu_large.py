"""Parity at the workloads' own sizes and formats, through the C-ABI.

* Jacobi in the 4K regime: RGB16 input, a 256-colour octree palette, 8 sweeps,
  taps 96, fp32 -- on a 960x540 crop and on one whole 3840x2160 frame, compared
  with the oracle; the stated tolerance (at most 0.5% of pixels with a
  differing palette index, mean per-channel Q16 error stated) is computed and
  asserted, and exact agreement is asserted too.  The whole frame also runs as
  a captured frame chain, which must equal the eager call, and into INDEX8,
  where every index must be below the palette size.
* The sequential walk on a 2048x2048 RGBA8 still with 256 colours, associated
  alpha and the default memo (the largest still of the batched workload), and
  on the degenerate 2048x1 and 1x2048 strips.
* The block walk at B=64 on one 1920x1080 plasma frame (curve level 11), from
  RGB8 and from planar 4:2:0.

Each case prints its name and shape before touching the device.
"""
import numpy as np
import pytest

from stipple_amd import synth
from tests import helpers as H

pytestmark = pytest.mark.gpu
_KEEP = []
TOL_INDEX_FRACTION = 0.005


def say(*a):
    print(*a, flush=True)


def _index_stats(oracle, got, exp, q_exp, w, h):
    """Fraction of pixels whose chosen entry differs and the mean per-channel Q16
    error, over the pixels that own a curve position."""
    differ = np.any(got != exp, axis=-1)
    err = np.abs(got[..., :3].astype(np.float64) - exp[..., :3].astype(np.float64)).mean()
    return float(differ.mean()), float(err)


def _jacobi(gpu, oracle, img16, colours, iterations=8, taps=96, dst_name="RGBAF32",
            prefill=None):
    import torch
    h, w = img16.shape[:2]
    f = synth.widen_q16(img16)
    tree = oracle.tree_build(f, colours, False, False)
    pal = gpu.Palette.from_tree(tree.data())
    src = H.to_device(img16)
    dst_fmt = getattr(gpu, "PIX_" + dst_name)
    dst = H.out_alloc((h, w), dst_fmt)
    if prefill is not None:
        dst.view(torch.float32).fill_(prefill)
    spec = gpu.image_spec(src, dst, w, h, gpu.PIX_RGB16, dst_fmt, pal)
    _KEEP.append((pal, src, dst))
    gpu.walk_approx([spec], iterations=iterations, taps=taps, fp64=False, use_tree=True)
    torch.cuda.synchronize()
    return tree, f, pal, src, dst, spec


def test_jacobi_4k_regime_crop(gpu, oracle):
    """960x540 crop of the first 4K radial frame, RGB16, 256 colours, 8 sweeps, taps 96."""
    crop = np.ascontiguousarray(synth.radial_frame16(0)[:540, :960])
    say("jacobi crop 960x540 RGB16 256 colours 8 sweeps taps 96")
    tree, f, pal, src, dst, _ = _jacobi(gpu, oracle, crop, 256)
    got = H.from_device(dst, (540, 960, 4), np.float32)
    exp, q = oracle.jacobi(tree, f, 8, 96, want_q=True)
    frac, err = _index_stats(oracle, got, exp, q, 960, 540)
    say(f"  differing indices {frac:.6f}, mean per-channel Q16 error {err:.4f}")
    assert frac <= TOL_INDEX_FRACTION
    assert np.array_equal(got, exp)


def test_jacobi_full_palette_index_limit(gpu, oracle):
    """Uniform RGB16 noise quantised to exactly 256 colours: index 255 is used and
    survives the one-byte index buffer (the reference's limit, not an alias)."""
    rng = np.random.default_rng(17)
    img = rng.integers(0, 65536, (256, 384, 3), dtype=np.uint16)
    say("jacobi noise 384x256 RGB16 256 colours, INDEX8 out")
    tree, f, pal, src, dst, _ = _jacobi(gpu, oracle, img, 256, dst_name="INDEX8")
    assert tree.data().colormap.shape[0] == 256
    got = H.from_device(dst, (256 * 384,), np.uint8)
    _, q = oracle.jacobi(tree, f, 8, 96, want_q=True)
    order, owner = oracle.block_order(384, 256)
    want = np.zeros(256 * 384, np.uint8)
    own = owner >= 0
    want[own] = q[owner[own]]
    assert got.max() == 255
    assert np.array_equal(got, want)


def test_jacobi_4k_frame(gpu, oracle):
    """One whole 3840x2160 RGB16 frame: every pixel written (the output starts as
    NaN), equal to the oracle, and a captured chain replay equals the eager call."""
    import torch
    frame = synth.radial_frame16(0)
    h, w = frame.shape[:2]
    say("jacobi 3840x2160 RGB16 256 colours 8 sweeps taps 96")
    tree, f, pal, src, dst, spec = _jacobi(gpu, oracle, frame, 256, prefill=float("nan"))
    got = H.from_device(dst, (h, w, 4), np.float32).copy()
    assert not np.isnan(got).any()
    cm = tree.data().colormap.astype(np.float32)
    exp, q = oracle.jacobi(tree, f, 8, 96, want_q=True)
    frac, err = _index_stats(oracle, got, exp, q, w, h)
    say(f"  differing indices {frac:.6f}, mean per-channel Q16 error {err:.4f}")
    assert frac <= TOL_INDEX_FRACTION
    assert np.array_equal(got, exp)
    assert int(q.max()) < cm.shape[0]
    say("  frame chain replay")
    dst2 = H.out_alloc((h, w), gpu.PIX_RGBAF32)
    dst2.view(torch.float32).fill_(float("nan"))
    _KEEP.append(dst2)
    chain = gpu.FrameChain.approx([gpu.image_spec(src, dst2, w, h, gpu.PIX_RGB16,
                                                  gpu.PIX_RGBAF32, pal)],
                                  iterations=8, taps=96)
    for _ in range(2):
        chain.run()
        torch.cuda.synchronize()
        assert torch.equal(dst2, dst)
    chain.close()


@pytest.mark.parametrize("w,h", [(2048, 2048), (2048, 1), (1, 2048)])
def test_seq_walk_large_rgba(gpu, oracle, w, h):
    """The sequential walk at the largest still size, RGBA8 with transparent
    rectangles, 256 colours, associated alpha, the default memo (2^24 keys)."""
    import torch
    rng = np.random.default_rng(w + 5 * h)
    img = synth.sinusoid_image(w, h, rng, channels=4)
    say("seq walk", w, h, "RGBA8 256 colours assoc")
    tree, f = H.oracle_tree(oracle, img, 256)
    assert tree.data().assoc
    pal = gpu.Palette.from_tree(tree.data())
    src = H.to_device(img)
    dst = H.out_alloc((h, w), gpu.PIX_RGBAF32)
    _KEEP.append((pal, src, dst))
    gpu.walk_seq_batch([gpu.image_spec(src, dst, w, h, gpu.PIX_RGBA8, gpu.PIX_RGBAF32, pal)])
    torch.cuda.synchronize()
    got = H.from_device(dst, (h, w, 4), np.float32)
    exp = oracle.walk_seq(tree, f)
    assert np.array_equal(got.view(np.uint32), exp.view(np.uint32))


def test_blocks_1080p_b64(gpu, oracle):
    """One 1920x1080 plasma frame (level 11), B=64, 16 colours, RGB8 in and out."""
    import torch
    frame = synth.plasma_frame(0)
    say("blocks 1920x1080 B=64 RGB8")
    tree, f = H.oracle_tree(oracle, frame, 16, assoc=False)
    pal = gpu.Palette.from_tree(tree.data())
    src = H.to_device(frame)
    dst = H.out_alloc((1080, 1920), gpu.PIX_RGB8)
    dstf = H.out_alloc((1080, 1920), gpu.PIX_RGBAF32)
    _KEEP.append((pal, src, dst, dstf))
    gpu.walk_blocks([gpu.image_spec(src, dst, 1920, 1080, gpu.PIX_RGB8, gpu.PIX_RGB8, pal),
                     gpu.image_spec(src, dstf, 1920, 1080, gpu.PIX_RGB8, gpu.PIX_RGBAF32, pal)],
                    block=64)
    torch.cuda.synchronize()
    exp = oracle.walk_blocks(tree, f, 64)[0]
    gotf = H.from_device(dstf, (1080, 1920, 4), np.float32)
    assert np.array_equal(gotf, exp)
    got = H.from_device(dst, (1080, 1920, 3), np.uint8)
    assert np.array_equal(got, H.scale_char_np(exp[..., :3]))


def test_blocks_1080p_yuv420(gpu, oracle):
    """The same frame as planar 4:2:0 in, planar 4:4:4 out, B=64."""
    import torch
    frame = synth.plasma_frame(1)
    planes = synth.rgb8_to_yuv420(frame)
    say("blocks 1920x1080 B=64 yuv420 -> yuv444")
    f = oracle.yuv_to_rgbaf(planes, 1920, 1080, 8)[0]
    tree = oracle.tree_build(f, 16, False)
    pal = gpu.Palette.from_tree(tree.data())
    src = H.to_device(planes)
    dst = torch.zeros(3 * 1920 * 1080, dtype=torch.uint8, device="cuda")
    _KEEP.append((pal, src, dst))
    gpu.walk_blocks([gpu.image_spec(src, dst, 1920, 1080, 8, 7, pal)], block=64)
    torch.cuda.synchronize()
    exp = oracle.walk_blocks_yuv444(tree, f, 64)[0]
    assert np.array_equal(dst.cpu().numpy(), np.ascontiguousarray(exp).reshape(-1).view(np.uint8))
