"""sa_walk_blocks_noalloc (the block walk over the library's fixed device
scratch) against sa_walk_blocks, through the C-ABI: byte-identical outputs on
two calls in a row, the second with a smaller input (the scratch keeps the
first call's contents), and a clear refusal beyond the scratch's limits.  Each
case prints its name before touching the device."""
import numpy as np
import pytest

from stipple_amd import synth
from tests import helpers as H

pytestmark = pytest.mark.gpu
_KEEP = []


def say(*a):
    print(*a, flush=True)


def _frames(gpu, oracle, rng, w, h, outs, colours):
    """One frame per output format, each with its own palette; returns the specs
    for both engines (separate destinations) and the destination pairs."""
    import torch
    specs_a, specs_b, pairs = [], [], []
    for k, out in enumerate(outs):
        img = synth.sinusoid_image(w, h, rng, channels=3)
        tree, f = H.oracle_tree(oracle, img, colours[k % len(colours)], assoc=False)
        pal = gpu.Palette.from_tree(tree.data())
        src = H.to_device(img)
        fmt = getattr(gpu, "PIX_" + out)
        da = H.out_alloc((h, w), fmt)
        db = H.out_alloc((h, w), fmt)
        db.fill_(0xA5)                        # anything the walk fails to write shows
        _KEEP.append((pal, src, da, db))
        specs_a.append(gpu.image_spec(src, da, w, h, gpu.PIX_RGB8, fmt, pal))
        specs_b.append(gpu.image_spec(src, db, w, h, gpu.PIX_RGB8, fmt, pal))
        pairs.append((da, db))
    torch.cuda.synchronize()
    return specs_a, specs_b, pairs


def test_noalloc_blocks_equal_the_allocating_call_twice(gpu, oracle):
    import torch
    rng = np.random.default_rng(31)
    outs = ["RGB8", "RGBAF32", "RGBA16", "INDEX8", "RGB8", "RGBA8"]
    say("noalloc call 1: 6 frames 640x480, B=64")
    a1, b1, p1 = _frames(gpu, oracle, rng, 640, 480, outs, [16, 64, 256, 5])
    gpu.walk_blocks(a1, block=64)
    gpu.walk_blocks_noalloc(b1, block=64)
    torch.cuda.synchronize()
    for da, db in p1:
        assert torch.equal(da, db)
    say("noalloc call 2: 3 frames 200x150 (smaller), B=40, diffusion 0.7")
    a2, b2, p2 = _frames(gpu, oracle, rng, 200, 150, outs[:3], [8, 32, 200])
    gpu.walk_blocks(a2, block=40, diffusion=0.7)
    gpu.walk_blocks_noalloc(b2, block=40, diffusion=0.7)
    torch.cuda.synchronize()
    for da, db in p2:
        assert torch.equal(da, db)
    say("noalloc call 3: 1x1 and a power-of-two width (the unvisited pixel)")
    for (w, h) in [(1, 1), (64, 40)]:
        a3, b3, p3 = _frames(gpu, oracle, rng, w, h, ["RGBAF32"], [4])
        gpu.walk_blocks(a3, block=16)
        gpu.walk_blocks_noalloc(b3, block=16)
        torch.cuda.synchronize()
        for da, db in p3:
            assert torch.equal(da, db)


def test_noalloc_blocks_refusals(gpu, oracle):
    """Beyond the fixed scratch (a side above 4096, too many pixels) and the
    reference's own refusals: an error before any launch."""
    import torch
    rng = np.random.default_rng(5)
    img = synth.sinusoid_image(32, 32, rng, channels=3)
    tree, f = H.oracle_tree(oracle, img, 16, assoc=False)
    pal = gpu.Palette.from_tree(tree.data())
    _KEEP.append(pal)
    for (w, h) in [(4097, 2), (4096, 4096)]:
        say("noalloc refusal", w, h)
        src = torch.zeros(w * h * 3, dtype=torch.uint8, device="cuda")
        dst = torch.zeros(w * h * 3, dtype=torch.uint8, device="cuda")
        spec = gpu.image_spec(src, dst, w, h, gpu.PIX_RGB8, gpu.PIX_RGB8, pal)
        with pytest.raises(gpu.EngineError) as ei:
            gpu.walk_blocks_noalloc([spec], block=64)
        assert ei.value.code == gpu.SA_EUNSUPPORTED
        assert "scratch" in str(ei.value)
    src = torch.zeros(32 * 32 * 3, dtype=torch.uint8, device="cuda")
    dst = torch.zeros(32 * 32 * 3, dtype=torch.uint8, device="cuda")
    spec = gpu.image_spec(src, dst, 32, 32, gpu.PIX_RGB8, gpu.PIX_RGB8, pal)
    say("noalloc refusal B=8")
    with pytest.raises(gpu.EngineError) as ei:
        gpu.walk_blocks_noalloc([spec], block=8)
    assert ei.value.code == 2
    torch.cuda.synchronize()
