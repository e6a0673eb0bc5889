"""Resumable sequential walks (sa_walk_job_*): continuous batching.

Each image must be bit-exact with one uninterrupted RiemersmaWalkCpu walk
(the oracle) whatever it shared the device with, however its walk was cut
into slices and whenever it joined: the same result alone and in the mix.
"""
import numpy as np
import pytest

from stipple_amd import synth
from tests import helpers as H

pytestmark = pytest.mark.gpu

_KEEP = []


def say(*a):
    print(*a, flush=True)


def _case(gpu, oracle, img, colors, dst_fmt=None):
    tree, f = H.oracle_tree(oracle, img, colors)
    pal = gpu.Palette.from_tree(tree.data())
    h, w = img.shape[:2]
    src = H.to_device(img)
    dst_fmt = gpu.PIX_RGBAF32 if dst_fmt is None else dst_fmt
    dst = H.out_alloc((h, w), dst_fmt)
    spec = gpu.image_spec(src, dst, w, h, H.fmt_of(img), dst_fmt, pal)
    _KEEP.append((src, pal, tree))
    return {"tree": tree, "f": f, "dst": dst, "hw": (h, w), "spec": spec}


def _check(oracle, c, use_cache=True):
    h, w = c["hw"]
    got = H.from_device(c["dst"], (h, w, 4), np.float32)
    exp = oracle.walk_seq(c["tree"], c["f"], use_cache=use_cache)
    assert np.array_equal(got.view(np.uint32), exp.view(np.uint32)), f"mismatch at {w}x{h}"


def _images(seed, assoc_mix=True):
    rng = np.random.default_rng(seed)
    sizes = [(1, 1), (3, 5), (17, 13), (64, 48), (200, 150), (31, 257), (333, 111)]
    colours = [2, 4, 16, 9, 64, 256, 24]
    out = []
    for k, ((w, h), col) in enumerate(zip(sizes, colours)):
        ch = 4 if (assoc_mix and k % 2 == 0) else 3
        out.append((synth.sinusoid_image(w, h, rng, channels=ch), col))
    return out


@pytest.mark.parametrize("slice_visits", [48, 1000, 48 * 97, 0])
def test_job_slices_equal_one_walk(gpu, oracle, slice_visits):
    import torch
    say("job slices", slice_visits)
    cases = [_case(gpu, oracle, img, col) for img, col in _images(3)]
    job = gpu.WalkJob()
    ids = job.add([c["spec"] for c in cases])
    assert ids == list(range(len(cases)))
    done = []
    steps = 0
    while job.active:
        done += job.step(slice_visits)
        steps += 1
        assert steps < 100000
    torch.cuda.synchronize()
    assert sorted(done) == ids
    for c in cases:
        _check(oracle, c)
    job.close()


def test_job_joins_mid_flight(gpu, oracle):
    """Images that join while others are half-walked; small memo so the
    slices carry evicted and dirty lines across launches."""
    import torch
    say("job joins mid-flight")
    cases = [_case(gpu, oracle, img, col) for img, col in _images(4)]
    job = gpu.WalkJob(gpu.SeqOptions(memo_bits=4))
    order = [4, 6, 0, 5, 1, 3, 2]
    job.add([cases[order[0]]["spec"], cases[order[1]]["spec"]])
    pending = order[2:]
    finished = set()
    step = 0
    while job.active or pending:
        if pending and step % 3 == 1:
            job.add([cases[pending.pop(0)]["spec"]])
        finished.update(job.step(96 * (1 + step % 5)))
        step += 1
        assert step < 100000
    torch.cuda.synchronize()
    assert finished == set(range(len(cases)))
    for c in cases:
        _check(oracle, c)
    job.close()


def test_job_without_memo(gpu, oracle):
    import torch
    say("job no memo")
    cases = [_case(gpu, oracle, img, col) for img, col in _images(5)[:5]]
    job = gpu.WalkJob(gpu.SeqOptions(use_cache=False))
    job.add([c["spec"] for c in cases])
    while job.active:
        job.step(480)
    torch.cuda.synchronize()
    for c in cases:
        _check(oracle, c, use_cache=False)
    job.close()


def test_job_counters_match_one_launch(gpu, oracle):
    """The memo cache is restored exactly between slices: the walk counters
    (misses, line fills, HBM line loads) equal those of one uninterrupted
    launch of the same images."""
    import torch
    say("job counters")
    imgs = _images(6)
    a = [_case(gpu, oracle, img, col) for img, col in imgs]
    b = [_case(gpu, oracle, img, col) for img, col in imgs]
    opts = gpu.SeqOptions(memo_bits=5)
    gpu.seq_stats()
    gpu.walk_seq_batch([c["spec"] for c in a], opts)
    torch.cuda.synchronize()
    one = gpu.seq_stats()
    job = gpu.WalkJob(opts)
    job.add([c["spec"] for c in b])
    while job.active:
        job.step(240)
    torch.cuda.synchronize()
    sliced = gpu.seq_stats()
    assert one == sliced, (one, sliced)
    for c in b:
        _check(oracle, c)
    job.close()


def test_job_output_formats(gpu, oracle):
    import torch
    say("job rgba8 out")
    rng = np.random.default_rng(8)
    img = synth.sinusoid_image(150, 70, rng, channels=4)
    c = _case(gpu, oracle, img, 16, gpu.PIX_RGBA8)
    job = gpu.WalkJob()
    job.add([c["spec"]])
    while job.active:
        job.step(1024)
    torch.cuda.synchronize()
    exp = oracle.walk_seq(c["tree"], c["f"])
    got = H.from_device(c["dst"], (70, 150, 4), np.uint8)
    assert np.array_equal(got, H.scale_char_np(exp))
    job.close()


@pytest.mark.parametrize("budget_us", [20.0, 500.0])
def test_job_timed_steps(gpu, oracle, budget_us):
    """Time-budgeted steps: images stop wherever the clock says, at their own
    pace; the result is still one uninterrupted walk, and so are the counters."""
    import torch
    say("job timed", budget_us)
    imgs = _images(7)
    a = [_case(gpu, oracle, img, col) for img, col in imgs]
    b = [_case(gpu, oracle, img, col) for img, col in imgs]
    opts = gpu.SeqOptions(memo_bits=6)
    gpu.seq_stats()
    gpu.walk_seq_batch([c["spec"] for c in a], opts)
    torch.cuda.synchronize()
    one = gpu.seq_stats()
    job = gpu.WalkJob(opts)
    ids = job.add([c["spec"] for c in b])
    done, steps = [], 0
    while job.active:
        done += job.step_timed(budget_us)
        steps += 1
        assert steps < 100000
    torch.cuda.synchronize()
    assert sorted(done) == ids
    assert gpu.seq_stats() == one
    for c in b:
        _check(oracle, c)
    job.close()


def test_job_overlapped_steps(gpu, oracle):
    """Overlapped steps: launch, then collect the previous step.  Images join
    mid-flight; each result is one uninterrupted walk."""
    import torch
    say("job overlapped")
    cases = [_case(gpu, oracle, img, col) for img, col in _images(9)]
    job = gpu.WalkJob(gpu.SeqOptions(memo_bits=5))
    job.add([c["spec"] for c in cases[:3]])
    pending = list(range(3, len(cases)))
    finished, steps = set(), 0
    while job.active or pending:
        if pending and steps % 2 == 0:
            job.add([cases[pending.pop(0)]["spec"]])
        finished.update(job.step_overlapped(50.0))
        steps += 1
        assert steps < 100000
    torch.cuda.synchronize()
    assert finished == set(range(len(cases)))
    for c in cases:
        _check(oracle, c)
    job.close()


@pytest.mark.parametrize("mode", ["overlapped", "timed", "slice"])
def test_job_more_images_than_wavefronts(gpu, oracle, mode):
    """Seven images on two wavefronts: a bounded step walks only the images
    with the most visits left and the others wait their turn; every image
    still finishes as one uninterrupted walk."""
    import torch
    say("job waves=2", mode)
    cases = [_case(gpu, oracle, img, col) for img, col in _images(15)]
    job = gpu.WalkJob(gpu.SeqOptions(memo_bits=5, waves=2))
    ids = job.add([c["spec"] for c in cases])
    finished, steps = [], 0
    while job.active:
        if mode == "overlapped":
            finished += job.step_overlapped(40.0)
        elif mode == "timed":
            finished += job.step_timed(40.0)
        else:
            finished += job.step(480)
        steps += 1
        assert steps < 200000
    torch.cuda.synchronize()
    assert sorted(finished) == ids
    for c in cases:
        _check(oracle, c)
    job.close()


def test_job_keeps_what_its_images_use(gpu, oracle):
    """A job holds its images' tensors and palettes (through their specs)
    until they finish: with every other reference dropped and the freed sizes
    allocated again and scribbled over, the outputs still equal the oracle's."""
    import gc
    import torch
    say("job keeps its images' buffers")
    rng = np.random.default_rng(21)
    job = gpu.WalkJob()
    expect = []
    for w, h, col in [(96, 64, 16), (40, 31, 5)]:
        img = synth.sinusoid_image(w, h, rng, channels=4)
        tree, f = H.oracle_tree(oracle, img, col)
        exp = oracle.walk_seq(tree, f)
        ids = job.add([gpu.image_spec(H.to_device(img), H.out_alloc((h, w), gpu.PIX_RGBAF32),
                                      w, h, H.fmt_of(img), gpu.PIX_RGBAF32,
                                      gpu.Palette.from_tree(tree.data()))])
        expect.append((ids[0], job._refs[ids[0]][1], exp, (h, w)))
    outs = {i: d for i, d, _, _ in expect}
    expect = [(i, e, hw) for i, _, e, hw in expect]
    gc.collect()
    junk = [torch.full((n,), 255, dtype=torch.uint8, device="cuda") for n in (96 * 64 * 4, 40 * 31 * 4)]
    while job.active:
        job.step(480)
    torch.cuda.synchronize()
    for i, exp, (h, w) in expect:
        got = H.from_device(outs[i], (h, w, 4), np.float32)
        assert np.array_equal(got.view(np.uint32), exp.view(np.uint32))
    del junk
    job.close()
    assert not job._refs
