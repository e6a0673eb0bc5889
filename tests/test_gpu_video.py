"""Video palette sampling on the GPU, through the C-ABI.

The device montage (sa_video_montage) against the oracle's restatement, byte
for byte; the montage palette (sa_video_palette) against ColorTree::Build on
the oracle's montage -- every node -- and against the reference's own
ColorTree::Build (oracle/_ref) on the same montage; a palette .txt imported
with its tree built on the device against the same steps on the CPU.  Each
case prints its name first.
"""
import numpy as np
import pytest

from oracle.oracle import Ref
from stipple_amd import synth
from tests import helpers as H

pytestmark = pytest.mark.gpu
_KEEP = []


def say(*a):
    print(*a, flush=True)


def _fmt(gpu, f):
    if f.dtype == np.uint8:
        return gpu.PIX_RGB8 if f.shape[-1] == 3 else gpu.PIX_RGBA8
    return gpu.PIX_RGB16 if f.shape[-1] == 3 else gpu.PIX_RGBA16


def _plasma_frames(n, w, h, ch=3, dtype=np.uint8, seed=3):
    out = []
    for t in range(n):
        f = synth.plasma_frame(t * 7, w, h, seed=seed)
        if ch == 4:
            a = np.full((h, w, 1), 255, np.uint8)
            a[h // 3:h // 2, w // 4:w // 2] = 0
            f = np.concatenate([f, a], -1)
        if dtype == np.uint16:
            f = f.astype(np.uint16) * 257 + (np.arange(w, dtype=np.uint16)[None, :, None] % 7)
        out.append(np.ascontiguousarray(f))
    return out


@pytest.mark.parametrize("n,want,tile,cols,ch,dtype", [
    (1, 1, 0, 0, 3, np.uint8), (5, 5, 0, 0, 3, np.uint8), (3, 7, 0, 0, 4, np.uint8),
    (6, 6, 48, 0, 3, np.uint8), (4, 9, 100, 0, 4, np.uint16), (3, 3, 0, 3, 3, np.uint16),
    (7, 8, 33, 0, 3, np.uint8)])
def test_montage_equals_the_oracle(gpu, oracle, n, want, tile, cols, ch, dtype):
    import torch
    w, h = 160, 90
    fr = _plasma_frames(n, w, h, ch, dtype)
    dev = [H.to_device(f) for f in fr]
    _KEEP.append(dev)
    say("montage", n, want, tile, cols, ch, np.dtype(dtype).name)
    m, mw, mh, gray = gpu.video_montage(dev, w, h, _fmt(gpu, fr[0]), want=want, tile=tile,
                                        cols=cols)
    torch.cuda.synchronize()
    exp = oracle.video_montage(fr, want=want, tile=tile, cols=cols)
    assert (mh, mw) == exp.shape[:2]
    got = H.from_device(m, (mh, mw, 4), np.uint16)
    assert np.array_equal(got.astype(np.float32), exp)
    assert gray == bool(np.all((exp[..., 0] == exp[..., 1]) & (exp[..., 1] == exp[..., 2])))


def test_montage_gray_flag(gpu, oracle):
    import torch
    g = np.repeat(np.arange(64, dtype=np.uint8).reshape(8, 8, 1), 3, axis=-1)
    dev = [H.to_device(g), H.to_device(g[::-1].copy())]
    _KEEP.append(dev)
    say("montage gray flag")
    _, _, _, gray = gpu.video_montage(dev, 8, 8, gpu.PIX_RGB8, want=2)
    assert gray
    c = g.copy()
    c[5, 6, 1] += 1
    dev2 = [H.to_device(g), H.to_device(c)]
    _KEEP.append(dev2)
    _, _, _, gray2 = gpu.video_montage(dev2, 8, 8, gpu.PIX_RGB8, want=2)
    torch.cuda.synchronize()
    assert not gray2


@pytest.mark.parametrize("n,want,tile,colours,ch", [(8, 8, 0, 16, 3), (5, 9, 120, 16, 4),
                                                    (12, 12, 64, 256, 3), (3, 4, 0, 7, 3)])
def test_video_palette_equals_colortree_build(gpu, oracle, n, want, tile, colours, ch):
    """Every node of the device tree equals the oracle's ColorTree::Build on the
    montage, and its colormap equals the reference's own ColorTree::Build."""
    from tests.test_gpu_octree import _check_built
    w, h = 320, 180
    fr = _plasma_frames(n, w, h, ch)
    dev = [H.to_device(f) for f in fr]
    _KEEP.append(dev)
    say("video palette", n, want, tile, colours, ch)
    pal = gpu.Palette.from_video(dev, w, h, _fmt(gpu, fr[0]), colours, want=want, tile=tile,
                                 keep_tree=True)
    _KEEP.append(pal)
    m = oracle.video_montage(fr, want=want, tile=tile)
    gray = bool(np.all((m[..., 0] == m[..., 1]) & (m[..., 1] == m[..., 2])))
    _check_built(oracle, pal, m, colours, False, gray)
    if Ref.available():
        ref = Ref()
        rt = ref.tree_build(m, colours, False, gray)
        assert np.array_equal(pal.colormap(), rt.data().colormap)


def test_palette_txt_import_builds_the_tree(gpu, oracle, tmp_path):
    """ImReadPalette's text path: entries read, narrowed and sorted, a count x 1
    strip's tree built at `count` colours, its colormap adopted."""
    rng = np.random.default_rng(8)
    n = 23
    ent = np.zeros((n, 4))
    ent[:, :3] = rng.integers(0, 256, (n, 3)) * 257.0 + rng.uniform(0, 1, (n, 3))
    ent[:, 3] = 65535.0
    p = tmp_path / "pal.txt"
    gpu.write_palette_txt(ent, str(p))
    say("palette txt import")
    pal = gpu.Palette.import_txt(str(p), expect_colors=n)
    _KEEP.append(pal)
    got_ent, _ = gpu.read_palette_txt(str(p))
    strip = got_ent.astype(np.float32).reshape(1, n, 4)
    t = oracle.tree_build(strip, n, False, False)
    assert t.count == n
    assert np.array_equal(pal.colormap(), t.data().colormap)
    # Duplicate colours: the tree holds fewer leaves than rows -- refused.
    ent[5, :3] = ent[9, :3]
    q = tmp_path / "dup.txt"
    gpu.write_palette_txt(ent, str(q))
    with pytest.raises(gpu.EngineError) as ei:
        gpu.Palette.import_txt(str(q))
    assert "reduced 23 colours to" in str(ei.value)
