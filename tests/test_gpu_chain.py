"""Frame chains (sa_chain_*): an engine's batch captured into a hipGraph and
replayed.  Every replay must equal the eager engine call on the buffers'
current contents (itself pinned to the oracle elsewhere; one oracle check is
repeated here), with eager work and host synchronisation between replays."""
import numpy as np
import pytest

from stipple_amd import synth
from tests import helpers as H

pytestmark = pytest.mark.gpu
_KEEP = []


def say(*a):
    print(*a, flush=True)


def _frames(gpu, oracle, w, h, n, seed, colours=16, channels=3, dst_fmt=None):
    rng = np.random.default_rng(seed)
    dst_fmt = gpu.PIX_RGBAF32 if dst_fmt is None else dst_fmt
    out = []
    for _ in range(n):
        img = synth.sinusoid_image(w, h, rng, channels=channels)
        tree, f = H.oracle_tree(oracle, img, colours)
        pal = gpu.Palette.from_tree(tree.data())
        src = H.to_device(img)
        chain_dst = H.out_alloc((h, w), dst_fmt)
        eager_dst = H.out_alloc((h, w), dst_fmt)
        _KEEP.append((pal, src, chain_dst, eager_dst, tree))
        out.append({"img": img, "tree": tree, "pal": pal, "src": src, "cd": chain_dst,
                    "ed": eager_dst, "fmt": H.fmt_of(img)})
    return out


def _specs(gpu, fr, w, h, key, dst_fmt):
    return [gpu.image_spec(f["src"], f[key], w, h, f["fmt"], dst_fmt, f["pal"]) for f in fr]


def _refill(fr, rng, channels):
    import torch
    for f in fr:
        h, w = f["img"].shape[:2]
        img = synth.sinusoid_image(w, h, rng, channels=channels)
        f["img"] = img
        f["src"].copy_(torch.from_numpy(np.ascontiguousarray(img).reshape(-1).view(np.uint8).copy()).cuda())


@pytest.mark.parametrize("w,h,block", [(200, 150, 64), (128, 96, 16), (64, 64, 32)])
def test_chain_blocks_replays_equal_eager(gpu, oracle, w, h, block):
    import torch
    say("chain blocks", w, h, block)
    fr = _frames(gpu, oracle, w, h, 3, w + h)
    fmt = gpu.PIX_RGBAF32
    chain = gpu.FrameChain.blocks(_specs(gpu, fr, w, h, "cd", fmt), block=block)
    rng = np.random.default_rng(5)
    for rep in range(3):
        if rep:
            _refill(fr, rng, 3)
        chain.run()
        gpu.walk_blocks(_specs(gpu, fr, w, h, "ed", fmt), block=block)
        torch.cuda.synchronize()
        for k, f in enumerate(fr):
            got = H.from_device(f["cd"], (h, w, 4), np.float32)
            exp = H.from_device(f["ed"], (h, w, 4), np.float32)
            assert np.array_equal(got, exp), (rep, k)
        if rep == 0:  # the first replay against the oracle too (same inputs as its tree)
            f0 = fr[0]
            exp0 = oracle.walk_blocks(f0["tree"], synth.widen_q16(f0["img"]), block)[0]
            assert np.array_equal(H.from_device(f0["cd"], (h, w, 4), np.float32), exp0)
    chain.close()


@pytest.mark.parametrize("w,h,iterations,taps,fp64", [(200, 150, 8, 96, False),
                                                      (64, 64, 3, 16, True),
                                                      (97, 61, 0, 96, False)])
def test_chain_approx_replays_equal_eager(gpu, oracle, w, h, iterations, taps, fp64):
    import torch
    say("chain approx", w, h, iterations, taps, fp64)
    fr = _frames(gpu, oracle, w, h, 2, 3 * w + h, colours=32)
    fmt = gpu.PIX_RGBAF32
    chain = gpu.FrameChain.approx(_specs(gpu, fr, w, h, "cd", fmt), iterations=iterations,
                                  taps=taps, fp64=fp64)
    rng = np.random.default_rng(6)
    for rep in range(3):
        if rep:
            _refill(fr, rng, 3)
        chain.run()
        gpu.walk_approx(_specs(gpu, fr, w, h, "ed", fmt), iterations=iterations, taps=taps,
                        fp64=fp64)
        torch.cuda.synchronize()
        for k, f in enumerate(fr):
            got = H.from_device(f["cd"], (h, w, 4), np.float32)
            exp = H.from_device(f["ed"], (h, w, 4), np.float32)
            assert np.array_equal(got, exp), (rep, k)
    chain.close()


def test_chain_outputs_rgb8_and_streams(gpu, oracle):
    """RGB8 out, replayed on a side stream with the host synchronising in between."""
    import torch
    say("chain rgb8 side stream")
    w, h = 150, 100
    fr = _frames(gpu, oracle, w, h, 2, 9, dst_fmt=gpu.PIX_RGB8)
    st = torch.cuda.Stream()
    chain = gpu.FrameChain.blocks(_specs(gpu, fr, w, h, "cd", gpu.PIX_RGB8), block=64, stream=st)
    for rep in range(2):
        chain.run(stream=st)
        st.synchronize()
        gpu.walk_blocks(_specs(gpu, fr, w, h, "ed", gpu.PIX_RGB8), block=64)
        torch.cuda.synchronize()
        for f in fr:
            assert torch.equal(f["cd"], f["ed"])
    chain.close()


def test_chain_refusals(gpu, oracle):
    fr = _frames(gpu, oracle, 32, 32, 1, 1)
    specs = _specs(gpu, fr, 32, 32, "cd", gpu.PIX_RGBAF32)
    with pytest.raises(gpu.EngineError) as ei:
        gpu.FrameChain.blocks(specs, block=15)
    assert ei.value.code == gpu.SA_EREFUSED
    with pytest.raises(gpu.EngineError):
        gpu.FrameChain.approx(specs, taps=5000)
