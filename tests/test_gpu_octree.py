"""ColorTree::Build on the device against the oracle: same colormap, same node
numbering, the same doubles in every node (the root's quantize_error exactly
when it was summed serially, otherwise within the bound that makes it
irrelevant)."""
import os

import numpy as np
import pytest

from stipple_amd import synth
from tests import helpers as H
from tests.pngread import read_png, to_q16_rgbaf

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
_KEEP = []


def say(*a):
    print(*a, flush=True)


def _compare(gpu, oracle, img_dev, fmt, w, h, f, colours, assoc, gray):
    say("octree", w, h, "colours", colours, "assoc", assoc, "gray", gray)
    pal = gpu.Palette.build(img_dev, w, h, fmt, colours, assoc, gray, keep_tree=True)
    _KEEP.append(pal)
    t = oracle.tree_build(f, colours, assoc, gray)
    d = t.data()
    info = pal.info()
    assert info["count"] == t.count
    assert info["depth"] == t.depth
    assert info["nodes"] == t.node_count
    assert np.array_equal(pal.colormap(), d.colormap)
    g = pal.tree()
    assert np.array_equal(g["child"], d.child)
    assert np.array_equal(g["parent"], d.parent)
    assert np.array_equal(g["level"], d.level)
    assert np.array_equal(g["id"], d.ids)
    assert np.array_equal(g["number_unique"], d.number_unique)
    assert np.array_equal(g["color_number"], d.color_number)
    assert np.array_equal(g["total"].view(np.uint64), d.total.view(np.uint64))
    assert np.array_equal(g["qe"][1:].view(np.uint64), d.qe[1:].view(np.uint64))
    if info["exact_root"]:
        assert g["qe"][0] == d.qe[0]
    else:
        assert abs(g["qe"][0] - d.qe[0]) <= 1e-9 * abs(d.qe[0])
    return pal, t


@pytest.mark.parametrize("w,h,colours,ch", [
    (1, 1, 16, 3), (3, 5, 4, 3), (17, 13, 16, 3), (64, 48, 16, 3), (96, 64, 256, 4),
    (200, 150, 16, 3), (128, 128, 2, 3), (64, 64, 64, 4), (257, 3, 8, 3), (37, 91, 32, 4),
    (640, 480, 16, 3), (500, 300, 256, 3), (300, 200, 100, 4), (1000, 20, 7, 3)])
def test_octree_synthetic(gpu, oracle, w, h, colours, ch):
    rng = np.random.default_rng(w * 31 + h * 7 + colours)
    img = synth.sinusoid_image(w, h, rng, channels=ch)
    src = H.to_device(img)
    _KEEP.append(src)
    _compare(gpu, oracle, src, H.fmt_of(img), w, h, synth.widen_q16(img), colours, ch == 4,
             synth.is_gray(img))


def test_octree_few_colours_no_break(gpu, oracle):
    """Fewer distinct colours than the palette: loop 1 covers the image."""
    rng = np.random.default_rng(2)
    lut = rng.integers(0, 256, (5, 3)).astype(np.uint8)
    img = lut[rng.integers(0, 5, (90, 70))]
    src = H.to_device(img)
    _KEEP.append(src)
    _compare(gpu, oracle, src, gpu.PIX_RGB8, 70, 90, synth.widen_q16(img), 16, False, False)


def test_octree_gray(gpu, oracle):
    rng = np.random.default_rng(3)
    g = np.clip(np.round(np.linspace(0, 255, 120)[None, :] + rng.normal(0, 9, (80, 120))), 0, 255)
    img = np.repeat(g.astype(np.uint8)[..., None], 3, axis=-1)
    src = H.to_device(img)
    _KEEP.append(src)
    _compare(gpu, oracle, src, gpu.PIX_RGB8, 120, 80, synth.widen_q16(img), 16, False, True)


@pytest.mark.parametrize("name,colours", [("in_gray_alpha", 2), ("in_alpha", 64), ("in_plasma", 16),
                                          ("t_17x13", 16), ("t_3x5", 4), ("t_1x1", 16)])
def test_octree_reference_fixtures(gpu, oracle, name, colours):
    s, d, c = read_png(os.path.join(HERE, "golden", name + ".png"))
    f, has_alpha = to_q16_rgbaf(s, d, c)
    src = H.to_device(f)
    _KEEP.append(src)
    h, w = f.shape[:2]
    _compare(gpu, oracle, src, gpu.PIX_RGBAF32, w, h, f, colours, has_alpha, synth.is_gray(f))


def test_octree_plasma_1080p(gpu, oracle):
    img = synth.plasma_frame(0)
    src = H.to_device(img)
    _KEEP.append(src)
    _compare(gpu, oracle, src, gpu.PIX_RGB8, 1920, 1080, synth.widen_q16(img), 16, False, False)


def test_device_palette_drives_the_walk(gpu, oracle):
    """classify -> reduce -> walk entirely on the device equals the oracle."""
    import torch
    rng = np.random.default_rng(8)
    img = synth.sinusoid_image(150, 110, rng, channels=4)
    src = H.to_device(img)
    pal = gpu.Palette.build(src, 150, 110, gpu.PIX_RGBA8, 32, True)
    dst = H.out_alloc((110, 150), gpu.PIX_RGBAF32)
    gpu.walk_seq_batch([gpu.image_spec(src, dst, 150, 110, gpu.PIX_RGBA8, gpu.PIX_RGBAF32, pal)])
    torch.cuda.synchronize()
    f = synth.widen_q16(img)
    t = oracle.tree_build(f, 32, True)
    got = H.from_device(dst, (110, 150, 4), np.float32)
    assert np.array_equal(got, oracle.walk_seq(t, f))
