"""Parity of the HIP engines with the oracle, through the C-ABI.

Bit-exact for every integer, index and float output.  Each case prints its
name and shape before touching the device.
"""
import numpy as np
import pytest

from stipple_amd import synth
from tests import helpers as H

pytestmark = pytest.mark.gpu

_KEEP = []  # device buffers referenced by pending specs stay alive until the module ends

GEOMS = [(1, 1), (3, 5), (17, 13), (2, 1), (1, 2), (16, 16), (64, 64), (100, 37), (256, 192),
         (33, 200), (640, 480), (512, 480), (1000, 999), (2048, 7), (5, 2048)]


def say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("w,h", GEOMS)
def test_curve_order(gpu, oracle, w, h):
    say("curve", w, h)
    got = gpu.curve_order(w, h).cpu().numpy()
    assert np.array_equal(got, oracle.seq_order(w, h))


def _seq_case(gpu, oracle, img, colors, use_cache=True, dst_fmt=None, diffusion=1.0):
    tree, f = H.oracle_tree(oracle, img, colors)
    pal = gpu.Palette.from_tree(tree.data())
    h, w = img.shape[:2]
    src = H.to_device(img)
    fmt = H.fmt_of(img)
    dst_fmt = gpu.PIX_RGBAF32 if dst_fmt is None else dst_fmt
    dst = H.out_alloc((h, w), dst_fmt)
    spec = gpu.image_spec(src, dst, w, h, fmt, dst_fmt, pal)
    _KEEP.append((src, pal, tree))
    return tree, f, pal, spec, dst


def _expect_seq(oracle, tree, f, use_cache, diffusion=1.0):
    return oracle.walk_seq(tree, f, diffusion=diffusion, use_cache=use_cache)


@pytest.mark.parametrize("use_cache", [True, False])
@pytest.mark.parametrize("assoc", [False, True])
def test_seq_small_batch(gpu, oracle, use_cache, assoc):
    rng = np.random.default_rng(11 + assoc)
    sizes = [(1, 1), (3, 5), (17, 13), (64, 48), (96, 64), (128, 128), (200, 150), (31, 257)]
    colours = [2, 4, 16, 16, 64, 256, 16, 8]
    cases, specs = [], []
    for (w, h), col in zip(sizes, colours):
        img = synth.sinusoid_image(w, h, rng, channels=4 if assoc else 3)
        say("seq case", w, h, col, "assoc", assoc, "cache", use_cache)
        tree, f, pal, spec, dst = _seq_case(gpu, oracle, img, col, use_cache)
        cases.append((tree, f, pal, dst, (h, w)))
        specs.append(spec)
    gpu.walk_seq_batch(specs, gpu.SeqOptions(use_cache=use_cache))
    import torch
    torch.cuda.synchronize()
    for tree, f, pal, dst, hw in cases:
        got = H.from_device(dst, (hw[0], hw[1], 4), np.float32)
        exp = _expect_seq(oracle, tree, f, use_cache)
        assert np.array_equal(got.view(np.uint32), exp.view(np.uint32)), f"mismatch at {hw}"


@pytest.mark.parametrize("dst_name", ["RGBA8", "RGB8", "RGBA16", "INDEX16"])
def test_seq_output_formats(gpu, oracle, dst_name):
    rng = np.random.default_rng(5)
    img = synth.sinusoid_image(120, 90, rng, channels=4)
    dst_fmt = getattr(gpu, "PIX_" + dst_name)
    say("seq format", dst_name)
    tree, f, pal, spec, dst = _seq_case(gpu, oracle, img, 16, True, dst_fmt)
    gpu.walk_seq_batch([spec])
    import torch
    torch.cuda.synchronize()
    exp, idx = oracle.walk_seq(tree, f, want_index=True)
    h, w = img.shape[:2]
    if dst_name == "INDEX16":
        got = H.from_device(dst, (h * w,), np.uint16)
        order = oracle.seq_order(w, h)
        want = np.zeros(h * w, np.uint16)
        want[order] = idx
        assert np.array_equal(got, want)
    elif dst_name == "RGBA16":
        got = H.from_device(dst, (h, w, 4), np.uint16)
        assert np.array_equal(got, exp.astype(np.uint16))
    else:
        c = 4 if dst_name == "RGBA8" else 3
        got = H.from_device(dst, (h, w, c), np.uint8)
        assert np.array_equal(got, H.scale_char_np(exp[..., :c]))


def test_seq_ramp_640x480(gpu, oracle):
    img = synth.ramp_image()
    say("seq ramp 640x480 16 colours")
    tree, f, pal, spec, dst = _seq_case(gpu, oracle, img, 16)
    gpu.walk_seq_batch([spec])
    got = H.from_device(dst, (480, 640, 4), np.float32)
    exp = oracle.walk_seq(tree, f)
    assert np.array_equal(got, exp)


def test_seq_small_memo_spills(gpu, oracle):
    """A tiny LDS memo forces evictions into the global table; first-touch
    results must not change."""
    rng = np.random.default_rng(9)
    img = synth.sinusoid_image(300, 200, rng, channels=3, sigma=30.0)
    say("seq spill 300x200 memo_bits=8")
    tree, f, pal, spec, dst = _seq_case(gpu, oracle, img, 64)
    gpu.walk_seq_batch([spec], gpu.SeqOptions(memo_bits=8))
    got = H.from_device(dst, (200, 300, 4), np.float32)
    assert np.array_equal(got, oracle.walk_seq(tree, f))


@pytest.mark.parametrize("w,h,ch,colors,slice_visits", [
    (320, 240, 3, 4096, 0), (256, 200, 4, 2048, 0), (320, 240, 3, 4096, 48 * 31),
    (384, 384, 3, 65536, 0), (384, 384, 3, 65536, 48 * 211)])
def test_seq_large_palette(gpu, oracle, w, h, ch, colors, slice_visits):
    """Palettes whose search image is above the walk's 96 KiB LDS share (more
    than about 1400 colours) are read from HBM, and a full 65536-colour palette
    keeps index+1 in 32-bit memo entries; every output is the oracle's, in one
    launch and in resumable slices."""
    import torch
    img = np.random.default_rng(w + colors).integers(0, 256, (h, w, ch), dtype=np.uint8)
    say("seq large palette", w, h, ch, colors, "slices", slice_visits)
    tree, f, pal, spec, dst = _seq_case(gpu, oracle, img, colors)
    assert pal.info()["count"] > 1400
    if colors == 65536:
        assert pal.info()["count"] == 65536
    if slice_visits:
        job = gpu.WalkJob()
        job.add([spec])
        steps = 0
        while job.active:
            job.step(slice_visits)
            steps += 1
            assert steps < 100000
    else:
        gpu.walk_seq_batch([spec])
    torch.cuda.synchronize()
    got = H.from_device(dst, (h, w, 4), np.float32)
    assert np.array_equal(got, oracle.walk_seq(tree, f))


def test_seq_diffusion(gpu, oracle):
    rng = np.random.default_rng(4)
    img = synth.sinusoid_image(80, 60, rng, channels=3)
    say("seq diffusion 0.5")
    tree, f, pal, spec, dst = _seq_case(gpu, oracle, img, 16)
    gpu.walk_seq_batch([spec], gpu.SeqOptions(diffusion=0.5))
    got = H.from_device(dst, (60, 80, 4), np.float32)
    assert np.array_equal(got, oracle.walk_seq(tree, f, diffusion=0.5))


def _blocks_case(gpu, oracle, imgs, colors, block, dst_fmt):
    import torch
    trees, specs, dsts, fs = [], [], [], []
    for img in imgs:
        tree, f = H.oracle_tree(oracle, img, colors)
        pal = gpu.Palette.from_tree(tree.data())
        h, w = img.shape[:2]
        src = H.to_device(img)
        dst = H.out_alloc((h, w), dst_fmt)
        specs.append(gpu.image_spec(src, dst, w, h, H.fmt_of(img), dst_fmt, pal))
        trees.append((tree, pal, src))
        dsts.append(dst)
        fs.append(f)
    gpu.walk_blocks(specs, block=block)
    torch.cuda.synchronize()
    return trees, dsts, fs


@pytest.mark.parametrize("w,h", [(1, 1), (3, 5), (17, 13), (64, 64), (200, 150), (256, 192),
                                 (97, 301)])
@pytest.mark.parametrize("block", [16, 64, 100])
def test_blocks_float(gpu, oracle, w, h, block):
    rng = np.random.default_rng(w * 7 + h)
    imgs = [synth.sinusoid_image(w, h, rng, channels=3) for _ in range(2)]
    say("blocks", w, h, "B", block)
    trees, dsts, fs = _blocks_case(gpu, oracle, imgs, 16, block, gpu.PIX_RGBAF32)
    for (tree, _, _), dst, f in zip(trees, dsts, fs):
        got = H.from_device(dst, (h, w, 4), np.float32)
        exp = oracle.walk_blocks(tree, f, block)[0]
        assert np.array_equal(got, exp)


@pytest.mark.parametrize("assoc", [False, True])
def test_blocks_u16(gpu, oracle, assoc):
    rng = np.random.default_rng(21)
    w, h = 128, 96
    imgs = [synth.sinusoid_image(w, h, rng, channels=4 if assoc else 3) for _ in range(3)]
    say("blocks u16 assoc", assoc)
    trees, dsts, fs = _blocks_case(gpu, oracle, imgs, 64 if assoc else 16, 64, gpu.PIX_RGBA16)
    for (tree, _, _), dst, f in zip(trees, dsts, fs):
        got = H.from_device(dst, (h, w, 4), np.uint16)
        exp = oracle.walk_blocks_u16(tree, f, 64)[0]
        assert np.array_equal(got, exp)


def _mixed_frames(rng, w, h, kinds):
    out = []
    for kind in kinds:
        base = synth.sinusoid_image(w, h, rng, channels=3 if kind in ("rgb8", "rgb16", "f32") else 4)
        if kind in ("rgb8", "rgba8"):
            out.append(base)
        elif kind in ("rgb16", "rgba16"):
            noise = rng.integers(0, 257, base.shape)
            out.append(np.clip(base.astype(np.int64) * 257 + noise - 128, 0, 65535).astype(np.uint16))
        else:
            f = synth.widen_q16(base) + rng.uniform(-300, 300, (h, w, 4)).astype(np.float32)
            f[..., 3] = 65535.0
            out.append(np.clip(f, 0, 65535).astype(np.float32))
    return out


def _expect_blocks_out(oracle, tree, f, block, dst_name, w, h):
    exp = oracle.walk_blocks(tree, f, block)[0]
    if dst_name == "RGBAF32":
        return exp
    if dst_name == "RGBA16":
        return exp.astype(np.uint16)
    if dst_name in ("RGB8", "RGBA8"):
        return H.scale_char_np(exp[..., :3 if dst_name == "RGB8" else 4])
    # INDEX8: the last visit of each pixel (position 0, the origin, is visited
    # again at the end); the pixel the curve never reaches stays 0.
    _, idx = oracle.walk_blocks(tree, f, block, want_index=True)
    order = oracle.block_order(w, h)[0]
    want = np.zeros(h * w, np.uint8)
    want[order[1:]] = idx[0][1:]
    return want.reshape(h, w)


@pytest.mark.parametrize("kinds,block", [
    (("rgb8", "rgba8", "rgb16", "rgba16", "f32"), 64),
    (("rgb8", "rgb16", "rgba8"), 40),
    (("rgba8", "rgb8"), 100),
])
def test_blocks_mixed_sources_and_outputs(gpu, oracle, kinds, block):
    """Frames of every source class in one batch (staged as the widest class),
    each written in another output format, against the oracle; block lengths
    that are not a multiple of the 16-position staging chunk."""
    import torch
    w, h = 150, 110
    rng = np.random.default_rng(len(kinds) * 100 + block)
    imgs = _mixed_frames(rng, w, h, kinds)
    outs = ["RGBA8", "RGB8", "INDEX8", "RGBA16", "RGBAF32"]
    specs, cases = [], []
    for k, img in enumerate(imgs):
        tree, f = H.oracle_tree(oracle, img, 8 + 24 * k, assoc=False)
        pal = gpu.Palette.from_tree(tree.data())
        dst_name = outs[k % len(outs)]
        dst_fmt = getattr(gpu, "PIX_" + dst_name)
        src = H.to_device(img)
        dst = H.out_alloc((h, w), dst_fmt)
        specs.append(gpu.image_spec(src, dst, w, h, H.fmt_of(img), dst_fmt, pal))
        cases.append((tree, f, dst_name, dst, pal, src))
    say("blocks mixed", kinds, "B", block)
    gpu.walk_blocks(specs, block=block)
    torch.cuda.synchronize()
    for tree, f, dst_name, dst, _, _ in cases:
        say("  frame out", dst_name)
        exp = _expect_blocks_out(oracle, tree, f, block, dst_name, w, h)
        got = H.from_device(dst, exp.shape, exp.dtype)
        assert np.array_equal(got.view(np.uint8), exp.view(np.uint8)), dst_name


def test_blocks_refusals(gpu, oracle):
    rng = np.random.default_rng(1)
    img = synth.sinusoid_image(32, 32, rng, channels=3)
    tree, f = H.oracle_tree(oracle, img, 16)
    pal = gpu.Palette.from_tree(tree.data())
    src = H.to_device(img)
    dst = H.out_alloc((32, 32), gpu.PIX_RGBAF32)
    spec = gpu.image_spec(src, dst, 32, 32, gpu.PIX_RGB8, gpu.PIX_RGBAF32, pal)
    with pytest.raises(gpu.EngineError) as ei:
        gpu.walk_blocks([spec], block=15)
    assert ei.value.code == 2
    with pytest.raises(gpu.EngineError):
        gpu.walk_blocks([], block=64)
