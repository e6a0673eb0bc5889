"""The serving layer (stipple_amd/serve.py): continuous batching and mixed
requests.  Every result must equal the same request run alone, and the
oracle: the mix changes when work runs, never what it computes."""
import numpy as np
import pytest

from stipple_amd import synth
from tests import helpers as H

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def test_pipeline_builds_and_walks_like_the_oracle(gpu, oracle):
    import torch
    from stipple_amd.serve import SeqPipeline, Still
    say("pipeline")
    rng = np.random.default_rng(31)
    sizes = [(1, 1), (3, 5), (17, 13), (120, 80), (257, 33), (64, 64), (90, 200)]
    stills, exp = [], []
    for k, (w, h) in enumerate(sizes):
        ch = 4 if k % 2 else 3
        img = synth.sinusoid_image(w, h, rng, channels=ch)
        col = [2, 5, 16, 40, 8, 256, 12][k]
        tree, f = H.oracle_tree(oracle, img, col)
        exp.append(oracle.walk_seq(tree, f))
        src = H.to_device(img)
        dst = H.out_alloc((h, w), gpu.PIX_RGBAF32)
        stills.append(Still(src, dst, w, h, H.fmt_of(img), gpu.PIX_RGBAF32, col,
                            assoc=(ch == 4), gray=synth.is_gray(img)))
    torch.cuda.synchronize()
    pipe = SeqPipeline(build_threads=3, budget_us=150.0)
    for st in stills:
        pipe.submit(st)
    pipe.run()
    torch.cuda.synchronize()
    for st, e, (w, h) in zip(stills, exp, sizes):
        assert not st.error, st.error
        got = H.from_device(st.dst, (h, w, 4), np.float32)
        assert np.array_equal(got.view(np.uint32), e.view(np.uint32)), (w, h)
        assert st.t_done >= st.t_ready >= st.t_submit
    pipe.close()


def _request(gpu, engine, frames, w, h, colors, assoc, rng):
    import torch
    from stipple_amd.serve import Request
    ch = 4 if assoc else 3
    imgs = [synth.sinusoid_image(w, h, rng, channels=ch) for _ in range(frames)]
    pairs = [(H.to_device(im), H.out_alloc((h, w), gpu.PIX_RGBAF32)) for im in imgs]
    torch.cuda.synchronize()
    fmt = gpu.PIX_RGBA8 if assoc else gpu.PIX_RGB8
    return Request(engine, pairs, w, h, fmt, gpu.PIX_RGBAF32, colors, assoc, block=32,
                   iterations=4, taps=24), imgs


def _alone(gpu, oracle, req, imgs):
    """The oracle's result for the request: palette from frame 0, each frame
    through the request's engine."""
    tree, _ = H.oracle_tree(oracle, imgs[0], req.colors, req.assoc)
    out = []
    for im in imgs:
        f = synth.widen_q16(im)
        if req.engine == "sequential":
            out.append(oracle.walk_seq(tree, f))
        elif req.engine == "blocks":
            out.append(oracle.walk_blocks(tree, f, req.block)[0])
        else:
            out.append(oracle.jacobi(tree, f, req.iterations, req.taps))
    return out


def test_server_mix_equals_alone(gpu, oracle):
    import torch
    from stipple_amd.serve import Server, latency_summary
    say("server mix")
    rng = np.random.default_rng(41)
    specs = [("sequential", 1, 130, 90, 16, True), ("blocks", 1, 200, 120, 8, False),
             ("approx", 1, 96, 64, 12, False), ("sequential", 3, 64, 48, 6, False),
             ("blocks", 4, 80, 60, 10, True), ("sequential", 1, 3, 5, 4, True),
             ("approx", 2, 50, 70, 16, True), ("sequential", 1, 211, 47, 40, False)]
    reqs, expect = [], []
    for engine, frames, w, h, col, assoc in specs:
        r, imgs = _request(gpu, engine, frames, w, h, col, assoc, rng)
        reqs.append(r)
        expect.append(_alone(gpu, oracle, r, imgs))
    arrivals = [(0.002 * k, r) for k, r in enumerate(reqs)]
    srv = Server(hbm_budget=2 << 30, build_threads=3, budget_us=300.0)
    done = srv.run(arrivals)
    torch.cuda.synchronize()
    assert len(done) == len(reqs)
    for r, exp in zip(reqs, expect):
        assert not r.error, r.error
        for (src, dst), e in zip(r.frames, exp):
            got = H.from_device(dst, (r.height, r.width, 4), np.float32)
            assert np.array_equal(got.view(np.uint32), e.view(np.uint32)), (r.engine, r.width, r.height)
    lat = latency_summary(done)
    say("latency", lat)
    assert lat["n"] == len(reqs) and lat["p99_ms"] >= lat["p50_ms"] > 0
    srv.close()
