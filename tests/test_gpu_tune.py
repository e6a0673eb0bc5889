"""The autotuner on the device: every value of a knob gives byte-identical
outputs (tune() itself raises otherwise), each equals the oracle's, and the
fastest is left set."""
import numpy as np
import pytest

from stipple_amd import synth, tune
from tests import helpers as H

pytestmark = pytest.mark.gpu
_KEEP = []


def say(*a):
    print(*a, flush=True)


def _batch(gpu, oracle, w, h, n, seed, colours=16, dst_fmt=None):
    rng = np.random.default_rng(seed)
    dst_fmt = gpu.PIX_RGBAF32 if dst_fmt is None else dst_fmt
    specs, dsts, cases = [], [], []
    for _ in range(n):
        img = synth.sinusoid_image(w, h, rng, channels=3)
        tree, f = H.oracle_tree(oracle, img, colours)
        pal = gpu.Palette.from_tree(tree.data())
        src = H.to_device(img)
        dst = H.out_alloc((h, w), dst_fmt)
        _KEEP.append((pal, src, dst, tree))
        specs.append(gpu.image_spec(src, dst, w, h, gpu.PIX_RGB8, dst_fmt, pal))
        dsts.append(dst)
        cases.append((tree, f))
    return specs, dsts, cases


def test_tune_blocks_threads(gpu, oracle):
    say("tune blocks.threads")
    w, h = 256, 192
    specs, dsts, cases = _batch(gpu, oracle, w, h, 3, 1)
    before = gpu.tune_get("blocks.threads")
    try:
        res = tune.tune("blocks.threads", lambda: gpu.walk_blocks(specs, block=64),
                        outputs=lambda: dsts)
        say(res)
        assert res["best"] in tune.KNOBS["blocks.threads"] and len(res["ms"]) == 4
        gpu.walk_blocks(specs, block=64)
        import torch
        torch.cuda.synchronize()
        for (tree, f), dst in zip(cases, dsts):
            exp = oracle.walk_blocks(tree, f, 64)[0]
            assert np.array_equal(H.from_device(dst, (h, w, 4), np.float32), exp)
    finally:
        gpu.tune_set("blocks.threads", before)


def test_tune_approx_per(gpu, oracle):
    say("tune approx.per")
    w, h = 200, 150
    specs, dsts, cases = _batch(gpu, oracle, w, h, 2, 2, colours=32)
    before = gpu.tune_get("approx.per")
    try:
        res = tune.tune("approx.per", lambda: gpu.walk_approx(specs, iterations=4, taps=96),
                        outputs=lambda: dsts)
        say(res)
        assert res["best"] in tune.KNOBS["approx.per"]
        gpu.walk_approx(specs, iterations=4, taps=96)
        import torch
        torch.cuda.synchronize()
        for (tree, f), dst in zip(cases, dsts):
            exp = oracle.jacobi(tree, f, 4, 96, False, True)
            assert np.array_equal(H.from_device(dst, (h, w, 4), np.float32), exp)
    finally:
        gpu.tune_set("approx.per", before)


def test_tune_seq_memo_bits(gpu, oracle):
    say("tune seq.memo_bits")
    w, h = 120, 90
    specs, dsts, cases = _batch(gpu, oracle, w, h, 3, 3, colours=64)
    before = gpu.tune_get("seq.memo_bits")
    try:
        res = tune.tune("seq.memo_bits", lambda: gpu.walk_seq_batch(specs),
                        outputs=lambda: dsts, reps=1)
        say(res)
        assert res["best"] in tune.KNOBS["seq.memo_bits"]
        tree, f = cases[0]
        assert np.array_equal(H.from_device(dsts[0], (h, w, 4), np.float32),
                              oracle.walk_seq(tree, f))
    finally:
        gpu.tune_set("seq.memo_bits", before)
