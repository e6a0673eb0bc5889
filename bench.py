#!/usr/bin/env python3
"""Throughput of the exact sequential Riemersma walk on one MI355X.

Workload (the metric's batched configuration): 64 seeded RGBA8 stills of
uneven size (w, h uniform in [1, 2048], including 1x1, 3x5 and 17x13), alpha
from random opaque/transparent rectangles, RGB from seeded sinusoids plus
noise; palette size log-uniform on {2..256}; associated alpha.

One step = the whole hot path over the batch, inputs already in HBM:
octree classify -> reduce -> colormap on the device for every image (8 host
threads, one stream each, largest images first), the exact walk, and the pack
into RGBA8 outputs.  With the default pipeline chain each image joins one
resumable walk job as soon as its palette exists and the job advances every
admitted image by a few milliseconds of walking per launch (continuous
batching); the step ends when the last image is packed.

Reported beside it, in the same run:
  * plain: the same work one image per launch, synchronised per image (the
    shape of the reference's RiemersmaWalkCuda calls);
  * walk_only: the engine alone with resident palettes;
  * roofline of the walk kernel (algorithmic 8 B/visit: RGBA8 in + out),
    timed with HIP events on its own stream;
  * cpu_baseline: the reference's own CPU engine (oracle/_ref, built from the
    reference sources) -- or the restatement when that library is absent --
    on the host's cores, palette build + walk, one image per thread.

Prints ONE JSON line on stdout.
"""
from __future__ import annotations

import argparse
import json
import os
import queue
import sys
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mpixel/s per engine on 1 MI355X, batched vs plain; p50/p99 frame latency"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_batch(n, seed):
    from stipple_amd import synth
    return synth.still_batch(n, seed=seed)


def cpu_baseline(batch, threads_list):
    """Reference CPU engine (palette build + walk), one image per thread."""
    from oracle.oracle import Ref, Oracle
    from stipple_amd import synth
    use_ref = Ref.available()
    lib = Ref() if use_ref else Oracle()
    imgs = [synth.widen_q16(img) for img, _ in batch]
    pixels = sum(im.shape[0] * im.shape[1] for im in imgs)
    out = {}
    for threads in threads_list:
        work = [im.copy() for im in imgs]
        t0 = time.perf_counter()
        trees = [lib.tree_build(im, col, True, False) for im, (_, col) in zip(work, batch)]
        t1 = time.perf_counter()
        lib.walk_seq_many(trees, work, threads=threads)
        t2 = time.perf_counter()
        out[threads] = {"walk_s": t2 - t1, "palette_s": t1 - t0, "total_s": t2 - t0}
        del trees
    return ("reference" if use_ref else "port"), pixels, out


def _traffic(name, frames=None):
    """HBM bytes per launch measured by rocprofv3 --pmc on the same workload
    (profiles/<name>.json), or None; with `frames`, scaled from the profiled
    launch's frame count to this launch's (the bytes per frame are the same)."""
    path = os.path.join(ROOT, "profiles", name + ".json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        d = json.load(fh)
    t = d["traffic_bytes"]
    if frames and d.get("frames"):
        t = t * frames / d["frames"]
    return round(t)


def _ev_ms(fn, stream=None):
    import torch
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    a.record(stream)
    fn()
    b.record(stream)
    b.synchronize()
    return a.elapsed_time(b)


def _line(args, metric, value, unit, ms, dtype, config, roofline, cpu, extra=None):
    line = {"metric": metric, "value": round(value, 3), "unit": unit, "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 3),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": dtype,
            "data": "synthetic", "config": config, "roofline": roofline, "cpu_baseline": cpu}
    if extra:
        line.update(extra)
    if cpu:
        key = "gpu_vs_cpu_1core" if cpu.get("cores", 1) == 1 else f"gpu_vs_cpu_{cpu['cores']}_cores"
        line[key] = round(value / cpu["value"], 2)
    return line


def run_ramp(args, sa, torch, synth, dev):
    """640x480 RGB8 ramp + sine + noise, 16 colours: palette and exact walk of
    one image (the reference's CPU-runnable case)."""
    img = synth.ramp_image()
    h, w = img.shape[:2]
    src = torch.from_numpy(img.reshape(-1).copy()).cuda()
    dst = torch.empty(w * h * 3, dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()

    def step():
        pal = sa.Palette.build(src, w, h, sa.PIX_RGB8, 16, False, synth.is_gray(img))
        sa.walk_seq_batch([sa.image_spec(src, dst, w, h, sa.PIX_RGB8, sa.PIX_RGB8, pal)])
        torch.cuda.synchronize()
        return pal

    for _ in range(args.warmup):
        step()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        pal = step()
    t = (time.perf_counter() - t0) / args.steps
    spec = sa.image_spec(src, dst, w, h, sa.PIX_RGB8, sa.PIX_RGB8, pal)
    sa.timing_enable(True)
    sa.timing_collect()
    sa.walk_seq_batch([spec])
    torch.cuda.synchronize()
    walk_ms, _ = sa.timing_collect("seq_walk")
    sa.timing_enable(False)
    px = w * h
    cpu = None
    if not args.no_cpu:
        from oracle.oracle import Ref, Oracle
        lib = Ref() if Ref.available() else Oracle()
        f = synth.widen_q16(img)
        reps = 5
        t1 = time.perf_counter()
        for _ in range(reps):
            tr = lib.tree_build(f, 16, False, False)
            lib.walk_seq(tr, f)
        tc = (time.perf_counter() - t1) / reps
        cpu = {"value": round(px / tc / 1e6, 3), "unit": "Mpixel/s", "cores": 1,
               "kind": "reference" if Ref.available() else "port",
               "sample": f"the same image, {reps} runs: ColorTree::Build + RiemersmaWalkCpu"}
    achieved = 6.0 * px / (walk_ms * 1e-3) / 1e9
    roof = {"bound": "hbm", "achieved": round(achieved, 4), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 8), "traffic": None, "kernel": "seq_walk_kernel",
            "note": "one wavefront walks one image: the serial chain's latency bounds it, not HBM"}
    return _line(args, "Mpixel/s, one 640x480 image (palette + exact walk)", px / t / 1e6,
                 "Mpixel/s", t * 1e3, "f64",
                 {"workload": "640x480 RGB8 ramp/sine + N(0,8), 16-colour octree palette on the "
                              "device, exact sequential walk, RGB8 out", "device": dev},
                 roof, cpu, {"walk_ms": round(walk_ms, 3)})


def run_video(args, sa, torch, synth, dev):
    """1920x1080 plasma frames, a 16-colour palette per frame, block walk B=64:
    RGB8 in and out, or (--video-format yuv420) the video path's wire formats,
    planar 4:2:0 in and 4:4:4 out."""
    nf = args.frames or 32
    yuv = args.video_format == "yuv420"
    on_device = getattr(args, "device_gen", False) and not yuv
    params = synth.plasma_params(3)
    t0 = time.perf_counter()
    if on_device:
        # the same plasma formula made on the device (many frames)
        srcs = [synth.plasma_device(1920, 1080, t=t, seed=3).reshape(-1) for t in range(nf)]
        frames = [srcs[0].cpu().numpy().reshape(1080, 1920, 3)]
    else:
        frames = [synth.plasma_frame(t, params=params) for t in range(nf)]
    if yuv:
        frames = [synth.rgb8_to_yuv420(f) for f in frames]
    log(f"generated {nf} frames in {time.perf_counter() - t0:.1f}s")
    h, w = 1080, 1920
    fin = sa.PIX_YUV420 if yuv else sa.PIX_RGB8
    fout = sa.PIX_YUV444 if yuv else sa.PIX_RGB8
    if not on_device:
        srcs = [torch.from_numpy(f.reshape(-1).copy()).cuda() for f in frames]
    dsts = [torch.empty(w * h * 3, dtype=torch.uint8, device="cuda") for _ in srcs]
    torch.cuda.synchronize()

    def build():
        # every frame's palette in one batched device build
        return sa.Palette.build_batch([(srcs[k], w, h, fin, 16, False) for k in range(nf)])

    def step():
        pals = build()
        sa.walk_blocks([sa.image_spec(srcs[k], dsts[k], w, h, fin, fout, pals[k])
                        for k in range(nf)], block=64)
        torch.cuda.synchronize()
        return pals

    tuned = None
    if not args.no_tune:
        from stipple_amd import tune
        tp = build()
        tspecs = [sa.image_spec(srcs[k], dsts[k], w, h, fin, fout, tp[k]) for k in range(nf)]
        tuned = tune.tune("blocks.threads", lambda: sa.walk_blocks(tspecs, block=64),
                          outputs=lambda: dsts)
        log(f"tuned {tuned}")

    for _ in range(args.warmup):
        step()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        pals = step()
    t = (time.perf_counter() - t0) / args.steps
    specs = [sa.image_spec(srcs[k], dsts[k], w, h, fin, fout, pals[k]) for k in range(nf)]
    sa.timing_enable(True)
    sa.timing_collect()
    for _ in range(3):
        sa.walk_blocks(specs, block=64)
    torch.cuda.synchronize()
    walk_ms, launches = sa.timing_collect("blocks_walk")
    sa.timing_enable(False)
    walk_ms /= max(launches, 1)
    t0 = time.perf_counter()
    build()
    torch.cuda.synchronize()
    pal_ms = (time.perf_counter() - t0) * 1e3
    # plain: one frame per launch, palette + walk, synchronised per frame
    lat = []
    for k in range(0 if args.no_plain else nf):
        a = time.perf_counter()
        p = sa.Palette.build(srcs[k], w, h, fin, 16, False, False)
        sa.walk_blocks([sa.image_spec(srcs[k], dsts[k], w, h, fin, fout, p)], block=64)
        torch.cuda.synchronize()
        lat.append(time.perf_counter() - a)
    px = w * h
    piped = None
    if yuv:
        piped = _video_pipeline(sa, np, frames, w, h)
        log(f"pipeline {piped}")
    cpu = None
    if not args.no_cpu:
        from oracle.oracle import Ref, Oracle
        lib = Ref() if Ref.available() else Oracle()
        if yuv:
            # the CPU engine reads 4:4:4 only: upsampled beforehand, untimed
            f = Oracle().yuv_to_rgbaf(frames[0], w, h, sa.PIX_YUV420)[0]
        else:
            f = synth.widen_q16(frames[0])
        t1 = time.perf_counter()
        tr = lib.tree_build(f, 16, False, False)
        if yuv and isinstance(lib, Ref):
            lib.walk_blocks_fused(tr, 64, w, h, 1, batch=f, yuv_out=True)
        elif yuv:
            lib.walk_blocks_yuv444(tr, f, 64)
        else:
            lib.walk_blocks(tr, f, 64)
        tc = time.perf_counter() - t1
        cpu = {"value": round(px / tc / 1e6, 3), "unit": "Mpixel/s", "cores": 1,
               "kind": "reference" if Ref.available() else "port",
               "sample": "one 1920x1080 frame: ColorTree::Build + RiemersmaBlocksCpu B=64" +
                         (" with the fused 4:4:4 scatter (the frame widened from 4:2:0 "
                          "beforehand, untimed)" if yuv else "")}
    bpp = 4.5 if yuv else 6.0
    achieved = bpp * px * nf / (walk_ms * 1e-3) / 1e9
    roof = {"bound": "hbm", "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 6),
            "traffic": _traffic("blocks_yuv420_traffic" if yuv else "blocks_c3_traffic", nf),
            "kernel": "blocks_kernel",
            "note": (f"algorithmic 4.5 B/px (4:2:0 in + 4:4:4 out)" if yuv else
                     f"algorithmic 6 B/px (RGB8 in + out)") +
                    f" over one {nf}-frame launch; the kernel is FP64-VALU bound (about 99 + "
                    "15 x candidates non-FMA FP64 ops per px)"}
    workload = (f"{nf} 1920x1080 plasma frames as planar 4:2:0, a 16-colour octree palette per "
                "frame built on the device from the planes, block walk B=64, planar 4:4:4 out"
                if yuv else
                f"{nf} 1920x1080 RGB8 plasma frames, a 16-colour octree palette per "
                "frame built on the device, block walk B=64, RGB8 out" +
                (" (frames generated on the device)" if on_device else ""))
    return _line(args, "Mpixel/s and frames/s, 1920x1080 video, block engine B=64",
                 px * nf / t / 1e6, "Mpixel/s", t * 1e3, "f64",
                 {"workload": workload, "frames": nf, "device": dev}, roof, cpu,
                 {"fps": round(nf / t, 2), "walk_ms": round(walk_ms, 3), "tuned": tuned,
                  "pipeline": piped,
                  "walk_fps": round(nf / (walk_ms * 1e-3), 1), "palette_build_ms": round(pal_ms, 3),
                  "plain": None if not lat else {
                      "fps": round(nf / sum(lat), 2),
                      "latency_ms": {"p50": round(float(np.percentile(lat, 50)) * 1e3, 3),
                                     "p99": round(float(np.percentile(lat, 99)) * 1e3, 3)}}})


def _video_pipeline(sa, np, frames, w, h, n=128, batch=8, slots=3):
    """Host to host through the frame pipeline: 4:2:0 frames copied by the
    reader into pinned staging, uploaded, walked (each slot's captured chain,
    palette of frame 0 resident) and downloaded as 4:4:4, in order.  The rate
    includes both PCIe directions, so it is reported beside the line, never
    as its value."""
    import torch
    from stipple_amd.pipeline import FramePipeline
    src0 = torch.from_numpy(frames[0].copy()).cuda()
    pal = sa.Palette.build(src0, w, h, sa.PIX_YUV420, 16, False)
    pipe = FramePipeline(w, h, sa.PIX_YUV420, sa.PIX_YUV444, palette=pal, batch=batch,
                         slots=slots)

    def go(count):
        it = iter(range(count))

        def read(buf):
            k = next(it, None)
            if k is None:
                return False
            np.copyto(buf, frames[k % len(frames)])
            return True

        t0 = time.perf_counter()
        got = pipe.run(read, lambda i, buf: None)
        return got, time.perf_counter() - t0

    go(2 * batch * slots)
    pipe.latency_s.clear()
    got, t = go(n)
    lat = np.array(pipe.latency_s) * 1e3
    pipe.close()
    return {"frames": got, "batch": batch, "slots": slots, "fps": round(got / t, 1),
            "value": round(got * w * h / t / 1e6, 1), "unit": "Mpixel/s",
            "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 3),
                           "p99": round(float(np.percentile(lat, 99)), 3)},
            "note": "host-resident 4:2:0 in, host 4:4:4 out, PCIe both ways included"}


def run_screens(args, sa, torch, synth, dev):
    """The threshold-screen plugin dithers (bayer-ordered, void-and-cluster,
    void-and-cluster-fast, clustered-dot) over 1920x1080 RGB8 plasma frames
    with a resident 16-colour palette built on the device from frame 0."""
    nf = args.frames or 32
    params = synth.plasma_params(3)
    frames = [synth.plasma_frame(t, params=params) for t in range(nf)]
    h, w = frames[0].shape[:2]
    srcs = [torch.from_numpy(f.reshape(-1).copy()).cuda() for f in frames]
    dsts = [torch.empty(w * h * 3, dtype=torch.uint8, device="cuda") for _ in srcs]
    pal = sa.Palette.build(srcs[0], w, h, sa.PIX_RGB8, 16, False)
    specs = [sa.image_spec(srcs[k], dsts[k], w, h, sa.PIX_RGB8, sa.PIX_RGB8, pal)
             for k in range(nf)]
    stream = torch.cuda.current_stream()
    names = ["bayer-ordered", "void-and-cluster", "void-and-cluster-fast", "clustered-dot"]
    per = {}
    for screen, name in enumerate(names):
        for _ in range(args.warmup):
            sa.dither_screen(specs, screen)
        torch.cuda.synchronize()
        ms = _ev_ms(lambda: [sa.dither_screen(specs, screen) for _ in range(args.steps)],
                    stream) / args.steps
        per[name] = {"ms": round(ms, 4), "Mpixel_s": round(nf * w * h / ms / 1e3, 1),
                     "GB_s": round(6.0 * nf * w * h / ms / 1e6, 1)}
    # The error-diffusion plugins over the same frames.  floyd-steinberg and
    # atkinson pipeline a frame's rows in one wavefront; jarvis walks a frame
    # in raster order, so it is given more frames (one wavefront each) to fill
    # the device.
    diffuse = {}
    for kind, name in enumerate(["floyd-steinberg", "atkinson", "jarvis"]):
        sa.dither_diffuse(specs, kind)
        torch.cuda.synchronize()
        reps = max(1, min(args.steps, 3))
        ms = _ev_ms(lambda: [sa.dither_diffuse(specs, kind) for _ in range(reps)],
                    stream) / reps
        diffuse[name] = {"ms": round(ms, 3), "Mpixel_s": round(nf * w * h / ms / 1e3, 1),
                         "frames": nf}
        log(f"{name}: {ms:.2f} ms per {nf} frames")
    cpu_diffuse = None
    if not args.no_cpu:
        from oracle.oracle import Ref, Oracle
        f0 = synth.widen_q16(frames[0])
        cpu_diffuse = {}
        for kind, name in enumerate(["floyd-steinberg", "atkinson", "jarvis"]):
            t1 = time.perf_counter()
            if Ref.available():
                Ref().dither(name, pal.entries, f0)
            else:
                Oracle().dither_diffuse(pal.entries, f0, kind)
            cpu_diffuse[name] = round(w * h / (time.perf_counter() - t1) / 1e6, 3)
            diffuse[name]["vs_cpu_1core"] = round(diffuse[name]["Mpixel_s"] / cpu_diffuse[name], 1)
    t = per["bayer-ordered"]["ms"] * 1e-3
    px = w * h
    cpu = None
    if not args.no_cpu:
        from oracle.oracle import Ref, Oracle
        ent = pal.entries
        f = synth.widen_q16(frames[0])
        t1 = time.perf_counter()
        if Ref.available():
            Ref().dither("bayer-ordered", ent, f)
        else:
            Oracle().dither_screen(ent, f, 0)
        tc = time.perf_counter() - t1
        cpu = {"value": round(px / tc / 1e6, 3), "unit": "Mpixel/s", "cores": 1,
               "kind": "reference" if Ref.available() else "port",
               "sample": "one 1920x1080 frame through the reference's bayer-ordered plugin"}
    cl = per["clustered-dot"]
    roof = {"bound": "hbm", "achieved": cl["GB_s"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(cl["GB_s"] / HBM_PEAK_GBS, 6), "traffic": _traffic("screens_traffic"),
            "kernel": "k_screen<true> (clustered-dot)",
            "note": "algorithmic 6 B/px (RGB8 in + out); the OrderedPick screens are FP32-VALU "
                    "bound (16 candidates x ~14 ops per pixel), see per_screen"}
    return _line(args, "Mpixel/s per screen dither, 1920x1080 RGB8 frames",
                 px * nf / t / 1e6, "Mpixel/s", t * 1e3, "f32",
                 {"workload": f"{nf} 1920x1080 RGB8 plasma frames, bayer-ordered with a resident "
                              "16-colour device palette (value); every screen in per_screen",
                  "frames": nf, "device": dev}, roof, cpu,
                 {"per_screen": per, "diffusion": diffuse,
                  "cpu_diffusion_1core_Mpixel_s": cpu_diffuse})


def run_approx4k(args, sa, torch, synth, dev):
    """3840x2160 RGB16 radial gradients + salt and pepper, 256 colours from
    frame 0 held on the device, Jacobi 8 sweeps, taps 96, fp32."""
    nf = args.frames or 16
    t0 = time.perf_counter()
    if getattr(args, "device_gen", False):
        # the same radial-gradient formula made on the device (many frames)
        dev_frames = [synth.radial_device(t) for t in range(nf)]
        srcs = [f.reshape(-1).view(torch.uint8) for f in dev_frames]
        frames = [dev_frames[0].cpu().view(torch.int16).numpy().view(np.uint16)]
    else:
        frames = [synth.radial_frame16(t) for t in range(nf)]
        srcs = [torch.from_numpy(f.reshape(-1).view(np.uint8).copy()).cuda() for f in frames]
    log(f"generated {nf} frames in {time.perf_counter() - t0:.1f}s")
    h, w = frames[0].shape[:2]
    dsts = [torch.empty(w * h * 6, dtype=torch.uint8, device="cuda") for _ in srcs]
    torch.cuda.synchronize()
    pal = sa.Palette.build(srcs[0], w, h, sa.PIX_RGB16, 256, False, False)
    specs = [sa.image_spec(srcs[k], dsts[k], w, h, sa.PIX_RGB16, sa.PIX_RGB16, pal)
             for k in range(nf)]

    tuned = None
    if not args.no_tune:
        from stipple_amd import tune
        tuned = tune.tune("approx.per", lambda: sa.walk_approx(specs, iterations=8, taps=96),
                          outputs=lambda: dsts, reps=2)
        log(f"tuned {tuned}")
    # The per-frame chain with the palette resident: captured once as a graph
    # over the frame buffers and replayed each step.
    chain = sa.FrameChain.approx(specs, iterations=8, taps=96)

    def step():
        chain.run()
        torch.cuda.synchronize()

    def eager():
        sa.walk_approx(specs, iterations=8, taps=96)
        torch.cuda.synchronize()

    for _ in range(args.warmup):
        step()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    t = (time.perf_counter() - t0) / args.steps
    t0 = time.perf_counter()
    eager()
    t_eager = time.perf_counter() - t0
    sa.timing_enable(True)
    sa.timing_collect()
    eager()
    ap_ms, _ = sa.timing_collect("approx")
    sa.timing_enable(False)
    chain.close()
    px = w * h
    cpu = None
    tolerance = None
    if not args.no_cpu:
        # The CPU sample: a 960x540 crop of frame 0 through the restatement, with
        # frame 0's palette.  The same crop through the GPU engine gives the
        # stated tolerance (differing indices, mean per-channel Q16 error).
        from oracle.oracle import Oracle
        o = Oracle()
        f = synth.widen_q16(frames[0])
        tree = o.tree_build(f, 256, False, False)
        crop16 = np.ascontiguousarray(frames[0][:540, :960])
        crop = synth.widen_q16(crop16)
        threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
        t1 = time.perf_counter()
        exp = o.jacobi(tree, crop, 8, 96)
        tc = time.perf_counter() - t1
        cpu = {"value": round(crop.shape[0] * crop.shape[1] / tc / 1e6, 3), "unit": "Mpixel/s",
               "cores": min(threads, 16), "kind": "port",
               "sample": "a 960x540 crop of frame 0 through the CPU restatement of the Jacobi "
                         "engine (the reference has no CPU Jacobi), 8 sweeps, taps 96, the "
                         "sweep's positions split over host threads"}
        cpal = sa.Palette.from_tree(tree.data())
        csrc = torch.from_numpy(crop16.reshape(-1).view(np.uint8).copy()).cuda()
        cdst = torch.empty(540 * 960 * 16, dtype=torch.uint8, device="cuda")
        sa.walk_approx([sa.image_spec(csrc, cdst, 960, 540, sa.PIX_RGB16, sa.PIX_RGBAF32, cpal)],
                       iterations=8, taps=96)
        torch.cuda.synchronize()
        got = cdst.cpu().numpy().view(np.float32).reshape(540, 960, 4)
        differ = float(np.any(got[..., :3] != exp[..., :3], axis=-1).mean())
        q16 = float(np.abs(got[..., :3].astype(np.float64) - exp[..., :3]).mean())
        tolerance = {"differing_index_fraction": differ, "mean_q16_error": round(q16, 6),
                     "limit": 0.005, "within": differ <= 0.005,
                     "sample": "960x540 crop of frame 0, frame 0's 256-colour palette, "
                               "against the CPU restatement"}
        cpal.close()
    # Algorithmic bytes: per frame RGB16 in (6 B) and out (6 B); per sweep
    # the curve-order input read (6 B of RGB16) and the index in and out (2 B).
    algo = px * nf * (12 + 8 * 8)
    achieved = algo / (ap_ms * 1e-3) / 1e9
    roof = {"bound": "hbm", "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": _traffic("approx_c4_traffic", nf),
            "kernel": "approx (gather, 8 sweeps, scatter)",
            "note": "algorithmic 12 B/px per frame + 8 B/px per sweep (RGB16 in, index in "
                    "and out); traffic is the measured HBM bytes of the same launch; the sweep "
                    "is FP32-VALU bound at taps=96 (97 taps x 3 channels x multiply and add = "
                    "582 FP32 ops per px per sweep)"}
    return _line(args, "Mpixel/s and frames/s, 3840x2160 RGB16, Jacobi 8 sweeps",
                 px * nf / t / 1e6, "Mpixel/s", t * 1e3, "f32",
                 {"workload": f"{nf} 3840x2160 RGB16 radial-gradient frames with 5% salt and "
                              "pepper, one 256-colour palette from frame 0 resident on the device, "
                              "Jacobi 8 sweeps, taps 96, fp32, octree lookup" +
                              (" (frames generated on the device)"
                               if getattr(args, "device_gen", False) else ""),
                  "frames": nf, "device": dev}, roof, cpu,
                 {"fps": round(nf / t, 2), "approx_ms": round(ap_ms, 3),
                  "chain": "hipGraph replay of the 16-frame batch (sa_chain_run)", "tuned": tuned,
                  "eager_ms": round(t_eager * 1e3, 3),
                  "tolerance": tolerance})


def serving_mix(args, sa, torch, synth):
    """The serving mix: Poisson arrivals at --mix-rate requests/s for
    --mix-seconds through Server with a --mix-hbm-gb admission cap; 70% stills
    (square side uniform in [256, 2048], RGB8 plasma + noise), 30% segments of
    30 frames at 1280x720 whose palette (built from frame 0) stays on the
    device for all 30; colours uniform in {4..64}; engines 30% sequential,
    50% block (B=64), 20% Jacobi (8 sweeps, taps 96).  Inputs come from pools
    generated on the device beforehand (24 stills, 6 segments), so arrivals
    cost no host generation; outputs are RGB8.  One completed request of each
    (engine, kind) is checked against the same request run alone."""
    from stipple_amd.serve import Request, Server
    rng = np.random.default_rng(args.mix_seed)
    t0 = time.perf_counter()
    sides = [int(v) for v in rng.integers(256, 2049, 24)]
    still_pool = [synth.plasma_device(s_, s_, t=k, seed=11) for k, s_ in enumerate(sides)]
    seg_pool = [[synth.plasma_device(1280, 720, t=f, seed=20 + k) for f in range(30)]
                for k in range(6)]
    torch.cuda.synchronize()
    log(f"mix: input pools generated on the device in {time.perf_counter() - t0:.1f}s")
    sched, t = [], 0.0
    while True:
        t += float(rng.exponential(1.0 / args.mix_rate))
        if t >= args.mix_seconds:
            break
        kind = "still" if rng.random() < 0.7 else "segment"
        u = rng.random()
        engine = "sequential" if u < 0.3 else ("blocks" if u < 0.8 else "approx")
        colours = int(rng.integers(4, 65))
        idx = int(rng.integers(0, len(still_pool) if kind == "still" else len(seg_pool)))
        sched.append((t, kind, engine, colours, idx))
    # Sampled for the alone check: the first request of each (engine, kind)
    # after the first 10 s (the mix is then in full flow).
    sample, seen = set(), set()
    for k, (ta, kind, engine, _, _) in enumerate(sched):
        if ta >= min(10.0, args.mix_seconds / 2) and (engine, kind) not in seen:
            seen.add((engine, kind))
            sample.add(k)
    kept, stats = {}, []

    def make(k):
        ta, kind, engine, colours, idx = sched[k]
        frames = [still_pool[idx]] if kind == "still" else seg_pool[idx]
        h, w = frames[0].shape[:2]
        pairs = [(f, torch.empty(w * h * 3, dtype=torch.uint8, device="cuda")) for f in frames]
        req = Request(engine, pairs, w, h, sa.PIX_RGB8, sa.PIX_RGB8, colours, False, block=64,
                      iterations=8, taps=96)
        req.k, req.kind, req.px = k, kind, w * h * len(frames)
        return req

    def done(req):
        stats.append((req.k, req.kind, req.engine, req.px, req.t_arrive, req.t_done, req.error))
        if req.k in sample:
            kept[req.k] = req
        else:
            req.frames = None          # the outputs are complete: release them

    opts = sa.SeqOptions(memo_bits=args.mix_memo_bits)
    srv = Server(hbm_budget=int(args.mix_hbm_gb * (1 << 30)), build_threads=args.build_threads,
                 budget_us=args.budget_us, on_done=done, seq_options=opts)
    arrivals = [(sched[k][0], (lambda k=k: make(k))) for k in range(len(sched))]
    log(f"mix: {len(sched)} arrivals over {args.mix_seconds} s")
    srv.run(arrivals, stop_after=args.mix_seconds * 2 + 120, feed_until=args.mix_seconds * 1.5)
    torch.cuda.synchronize()
    t_first = srv.t0
    srv.close()
    errs = [e for *_, e in stats if e]
    if errs:
        raise RuntimeError(f"mix: {len(errs)} requests failed, first: {errs[0]}")
    lat = np.array([td - ta for _, _, _, _, ta, td, _ in stats])
    t_end = max(td for _, _, _, _, _, td, _ in stats)
    px = sum(p for _, _, _, p, _, _, _ in stats)

    def pct(sel):
        v = np.array([td - ta for _, kd, en, _, ta, td, _ in stats if sel(kd, en)])
        if v.size == 0:
            return None
        return {"n": int(v.size), "p50": round(float(np.percentile(v, 50)) * 1e3, 3),
                "p99": round(float(np.percentile(v, 99)) * 1e3, 3)}

    per = {f"{en}/{kd}": pct(lambda a, b, kd=kd, en=en: a == kd and b == en)
           for en in ("sequential", "blocks", "approx") for kd in ("still", "segment")}
    # The sampled requests again, each alone (its own palette build from frame 0
    # and one engine call): the mix must change when work runs, never what it
    # computes.  (tests/test_gpu_serve.py checks mixed results against the oracle.)
    checked, equal = 0, True
    for k, req in sorted(kept.items()):
        pal = sa.Palette.build(req.frames[0][0], req.width, req.height, sa.PIX_RGB8, req.colors,
                               False, False)
        pairs = [req.frames[0]] if req.kind == "still" else [req.frames[0], req.frames[-1]]
        alone = [torch.empty_like(dst) for _, dst in pairs]
        specs = [sa.image_spec(src, a, req.width, req.height, sa.PIX_RGB8, sa.PIX_RGB8, pal)
                 for (src, _), a in zip(pairs, alone)]
        if req.engine == "sequential":
            sa.walk_seq_batch(specs, sa.SeqOptions(memo_bits=args.mix_memo_bits))
        elif req.engine == "blocks":
            sa.walk_blocks(specs, block=64)
        else:
            sa.walk_approx(specs, iterations=8, taps=96)
        torch.cuda.synchronize()
        for (_, dst), a in zip(pairs, alone):
            equal = equal and bool(torch.equal(dst, a))
            checked += 1
        pal.close()
    res = {"seconds": args.mix_seconds, "rate_per_s": args.mix_rate, "hbm_cap_gb": args.mix_hbm_gb,
           "arrivals": len(sched), "completed": len(stats), "not_admitted": srv.dropped,
           "value": round(px / (t_end - t_first) / 1e6, 3), "unit": "Mpixel/s",
           "wall_s": round(t_end - t_first, 2), "pixels": int(px),
           "latency_ms": {"p50": round(float(np.percentile(lat, 50)) * 1e3, 3),
                          "p99": round(float(np.percentile(lat, 99)) * 1e3, 3)},
           "per_engine_kind": per, "admission_wait_s": round(srv.admission_wait_s, 2),
           "seq_memo_bits": args.mix_memo_bits or None,
           "checked_frames": checked, "same_as_alone": equal,
           "inputs": "pools generated on the device: 24 stills (side uniform in [256, 2048]) "
                     "and 6 segments of 30 1280x720 frames, RGB8 plasma + noise"}
    log(f"mix: {res['completed']} requests, {res['value']} Mpx/s, p50 {res['latency_ms']['p50']} "
        f"ms, p99 {res['latency_ms']['p99']} ms, {checked} frames checked, equal: {equal}")
    del still_pool, seg_pool, kept
    torch.cuda.synchronize()
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--images", type=int, default=64)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-plain", action="store_true",
                    help="skip the one-image (one-frame) per launch leg")
    ap.add_argument("--no-tune", action="store_true", help="skip the launch-shape autotuner")
    ap.add_argument("--build-threads", type=int, default=2,
                    help="host threads (one stream each) classifying palettes concurrently")
    ap.add_argument("--chain", choices=["pipeline", "batch"], default="pipeline",
                    help="pipeline: images join a resumable walk as their palettes are "
                         "classified (continuous batching); batch: classify all, then one "
                         "walk launch")
    ap.add_argument("--budget-us", type=float, default=16000.0,
                    help="pipeline: microseconds each image walks per step")
    ap.add_argument("--served-batches", type=int, default=8,
                    help="stills: batches submitted at once to the serving pipeline for the "
                         "served-throughput leg (0: skip)")
    ap.add_argument("--served-memo-bits", type=int, default=0,
                    help="stills: LDS memo size (seq.memo_bits, 4..10) for the served leg; "
                         "smaller memos let more images share a CU (0: library default)")
    ap.add_argument("--head", type=int, default=0,
                    help="batch chain: launch the walk of the N largest images before the "
                         "rest are classified (0: one walk launch after all palettes)")
    ap.add_argument("--video-format", choices=["rgb8", "yuv420"], default="rgb8",
                    help="video workload: RGB8 in and out, or planar 4:2:0 in and 4:4:4 out")
    ap.add_argument("--workload", choices=["stills", "ramp", "video", "approx4k", "screens", "mix"],
                    default="stills",
                    help="stills: 64 uneven RGBA8 stills, sequential engine (the headline); "
                         "ramp: one 640x480 image; video: 1080p frames, block engine; "
                         "approx4k: 4K RGB16 frames, Jacobi engine")
    ap.add_argument("--frames", type=int, default=0,
                    help="frames per step (0: 32 for video, one batched build; 16 for approx4k)")
    ap.add_argument("--mix-seconds", type=float, default=120.0,
                    help="stills: seconds of Poisson arrivals for the serving-mix leg (0: skip)")
    ap.add_argument("--mix-rate", type=float, default=200.0, help="serving mix: requests per second")
    ap.add_argument("--mix-hbm-gb", type=float, default=64.0, help="serving mix: HBM admission cap")
    ap.add_argument("--mix-memo-bits", type=int, default=8,
                    help="serving mix: LDS memo size of the sequential walks (4..10; 0: default)")
    ap.add_argument("--mix-seed", type=int, default=7)
    ap.add_argument("--engines", type=int, default=1,
                    help="stills: also measure the block engine on 300 1920x1080 frames and the "
                         "Jacobi engine on 120 3840x2160 RGB16 frames (frames generated on the "
                         "device) and report them in the same line (0: skip)")
    ap.add_argument("--device-gen", action="store_true",
                    help="video / approx4k: generate the frames on the device")
    ap.add_argument("--copy-calibrate", action="store_true",
                    help="first run a 256 MiB device copy (HBM counter calibration under rocprofv3)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the sequential walk shards by image with no "
              f"exchange, and this build runs and reports a single MI355X (got --gpus {args.gpus}).",
              file=sys.stderr)
        sys.exit(2)

    import torch
    import stipple_amd as sa
    from stipple_amd import synth

    sa.init(0)
    dev = sa.device_name()
    log("device:", dev)
    if args.copy_calibrate:
        a = torch.empty(256 << 20, dtype=torch.uint8, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    if args.workload == "mix":
        mix = serving_mix(args, sa, torch, synth)
        print(json.dumps(_line(args, "Mpixel/s and p50/p99 request latency, serving mix",
                               mix["value"], "Mpixel/s", mix["wall_s"] * 1e3, "f64",
                               {"workload": "serving mix: Poisson arrivals, stills and 30-frame "
                                            "segments, sequential/block/Jacobi engines",
                                "device": dev}, None, None, {"serving_mix": mix})), flush=True)
        return
    other = {"ramp": run_ramp, "video": run_video, "approx4k": run_approx4k,
             "screens": run_screens}.get(args.workload)
    if other:
        print(json.dumps(other(args, sa, torch, synth, dev)), flush=True)
        return
    t0 = time.perf_counter()
    batch = make_batch(args.images, args.seed)
    log(f"generated {len(batch)} images in {time.perf_counter() - t0:.1f}s")
    pixels = sum(img.shape[0] * img.shape[1] for img, _ in batch)

    srcs, dsts = [], []
    for img, _ in batch:
        srcs.append(torch.from_numpy(img.reshape(-1).view(np.uint8).copy()).cuda())
        dsts.append(torch.empty(img.shape[0] * img.shape[1] * 4, dtype=torch.uint8, device="cuda"))
    torch.cuda.synchronize()

    grays = [synth.is_gray(img) for img, _ in batch]
    order = sorted(range(len(batch)), key=lambda k: -batch[k][0].shape[0] * batch[k][0].shape[1])
    build_streams = queue.Queue()
    for _ in range(args.build_threads):
        build_streams.put(torch.cuda.Stream())
    pool = ThreadPoolExecutor(max_workers=args.build_threads)
    walk_streams = [torch.cuda.Stream(), torch.cuda.Stream()]

    def build_one(k):
        img, col = batch[k]
        st = build_streams.get()
        try:
            return sa.Palette.build(srcs[k], img.shape[1], img.shape[0], sa.PIX_RGBA8, col, True,
                                    grays[k], stream=st)
        finally:
            build_streams.put(st)

    def build_palettes(ks=None):
        ks = list(range(len(batch))) if ks is None else ks
        return dict(zip(ks, pool.map(build_one, ks)))

    def specs_for(pals, ks=None):
        ks = list(range(len(batch))) if ks is None else ks
        return [sa.image_spec(srcs[k], dsts[k], batch[k][0].shape[1], batch[k][0].shape[0],
                              sa.PIX_RGBA8, sa.PIX_RGBA8, pals[k]) for k in ks]

    from stipple_amd.serve import SeqPipeline, Still
    pipe = SeqPipeline(build_threads=args.build_threads, budget_us=args.budget_us) \
        if args.chain == "pipeline" else None
    last_lat = []

    def step_pipeline():
        # Largest images are classified first; each joins the walk as soon as
        # its palette exists, and every launch lets each admitted image walk
        # for a time budget (continuous batching on one resumable walk job).
        stills = [Still(srcs[k], dsts[k], batch[k][0].shape[1], batch[k][0].shape[0],
                        sa.PIX_RGBA8, sa.PIX_RGBA8, batch[k][1], True, grays[k]) for k in order]
        t0 = time.perf_counter()
        for st in stills:
            st.t_submit = t0
            pipe.submit(st)
        pipe.run()
        torch.cuda.synchronize()
        last_lat[:] = [st.t_done - st.t_submit for st in stills]
        errs = [st.error for st in stills if st.error]
        if errs:
            raise RuntimeError(errs[0])
        pals = [None] * len(batch)
        for k, st in zip(order, stills):
            pals[k] = st.palette
        return pals

    def step():
        if pipe is not None:
            return step_pipeline()
        # Palettes are classified concurrently (one stream per host thread); the
        # walk is one batched launch, largest images first.
        if args.head > 0:
            head, tail = order[:args.head], order[args.head:]
            pals = build_palettes(head)
            sa.walk_seq_batch(specs_for(pals, head), stream=walk_streams[0])
            if tail:
                pals.update(build_palettes(tail))
                sa.walk_seq_batch(specs_for(pals, tail), stream=walk_streams[1])
        else:
            pals = build_palettes(order)
            sa.walk_seq_batch(specs_for(pals, order))
        torch.cuda.synchronize()
        return [pals[k] for k in range(len(batch))]

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        pals = step()
    torch.cuda.synchronize()
    t_batched = (time.perf_counter() - t0) / args.steps
    log(f"batched chain: {t_batched * 1e3:.1f} ms/step  {pixels / t_batched / 1e6:.1f} Mpx/s")
    chain_lat = None
    if last_lat:
        chain_lat = {"p50": round(float(np.percentile(last_lat, 50)) * 1e3, 3),
                     "p99": round(float(np.percentile(last_lat, 99)) * 1e3, 3)}
        log(f"per-image latency in the batch: p50 {chain_lat['p50']} ms, p99 {chain_lat['p99']} ms, "
            f"{pipe.steps} walk steps so far")

    # Served: several copies of the batch submitted at once to the pipeline;
    # their images walk side by side (continuous batching), so the largest
    # image's serial walk no longer bounds the throughput.  Every output must
    # equal the same image's output walked in its own batch.
    served = None
    if pipe is not None and args.served_batches > 0:
        nb_s = args.served_batches
        sdsts = [[torch.empty_like(d) for d in dsts] for _ in range(nb_s)]
        torch.cuda.synchronize()
        stills = []
        for b in range(nb_s):
            for k in range(len(batch)):
                stills.append((k, Still(srcs[k], sdsts[b][k], batch[k][0].shape[1],
                                        batch[k][0].shape[0], sa.PIX_RGBA8, sa.PIX_RGBA8,
                                        batch[k][1], True, grays[k])))
        stills.sort(key=lambda e: -batch[e[0]][0].shape[0] * batch[e[0]][0].shape[1])
        memo_before = sa.tune_get("seq.memo_bits")
        if args.served_memo_bits:
            sa.tune_set("seq.memo_bits", args.served_memo_bits)
        t0 = time.perf_counter()
        for _, st in stills:
            st.t_submit = t0
            pipe.submit(st)
        pipe.run()
        torch.cuda.synchronize()
        t_served = time.perf_counter() - t0
        sa.tune_set("seq.memo_bits", memo_before)
        errs = [st.error for _, st in stills if st.error]
        if errs:
            raise RuntimeError(errs[0])
        same = all(torch.equal(sdsts[b][k], dsts[k]) for b in range(nb_s) for k in range(len(batch)))
        lat = [st.t_done - st.t_submit for _, st in stills]
        served = {"batches": nb_s, "images": nb_s * len(batch),
                  "memo_bits": args.served_memo_bits or None,
                  "value": round(nb_s * pixels / t_served / 1e6, 3), "ms": round(t_served * 1e3, 2),
                  "latency_ms": {"p50": round(float(np.percentile(lat, 50)) * 1e3, 3),
                                 "p99": round(float(np.percentile(lat, 99)) * 1e3, 3)},
                  "same_as_alone": same}
        log(f"served: {nb_s} batches at once, {served['value']} Mpx/s, "
            f"{t_served * 1e3:.0f} ms, outputs equal to the batch's own: {same}")
        del sdsts, stills
        torch.cuda.synchronize()

    mix = serving_mix(args, sa, torch, synth) if args.mix_seconds > 0 else None

    # The block and Jacobi engines at their workloads' full sizes.
    engines = None
    if args.engines:
        engines = {}
        for name, fn, nf in (("block_video_1080p", run_video, 300),
                             ("jacobi_4k_rgb16", run_approx4k, 120)):
            sub = argparse.Namespace(**vars(args))
            sub.frames, sub.steps, sub.warmup, sub.device_gen = nf, 3, 1, True
            sub.video_format = "rgb8"
            r = fn(sub, sa, torch, synth, dev)
            keep = ("value", "unit", "ms_per_step", "fps", "roofline", "cpu_baseline",
                    "tolerance", "plain", "walk_ms", "approx_ms", "palette_build_ms", "tuned",
                    "config", "dtype", "steps", "warmup")
            engines[name] = {k: r[k] for k in keep if k in r}
            for k in r:
                if k.startswith("gpu_vs_cpu"):
                    engines[name][k] = r[k]
            log(f"{name}: {r['value']} Mpx/s")
            torch.cuda.empty_cache()

    # Engine alone, palettes resident, with the walk kernel timed by events.
    specs = specs_for(pals)
    sa.timing_enable(True)
    sa.timing_collect()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        sa.walk_seq_batch(specs)
    torch.cuda.synchronize()
    t_walk = (time.perf_counter() - t0) / args.steps
    walk_ms, walk_launches = sa.timing_collect("seq_walk")
    gather_ms, _ = sa.timing_collect("seq_gather")
    pack_ms, _ = sa.timing_collect("seq_pack")
    sa.timing_collect()
    sa.timing_enable(False)
    walk_ms /= max(walk_launches, 1)
    gather_ms /= max(walk_launches, 1)
    pack_ms /= max(walk_launches, 1)
    log(f"walk only: {t_walk * 1e3:.1f} ms  kernel {walk_ms:.1f} ms  gather {gather_ms:.2f} pack {pack_ms:.2f}")

    # Palette build alone: the 64 palettes as one batched device build, and
    # one image at a time.
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    built = sa.Palette.build_batch([(srcs[k], batch[k][0].shape[1], batch[k][0].shape[0],
                                     sa.PIX_RGBA8, batch[k][1], True, grays[k]) for k in order])
    torch.cuda.synchronize()
    t_pal = time.perf_counter() - t0
    for p in built:
        p.close()
    t0 = time.perf_counter()
    for k in range(len(batch)):
        img, col = batch[k]
        sa.Palette.build(srcs[k], img.shape[1], img.shape[0], sa.PIX_RGBA8, col, True, grays[k])
    torch.cuda.synchronize()
    t_pal_serial = time.perf_counter() - t0

    plain = None
    if not args.no_plain:
        # One image per launch, synchronised each time.
        torch.cuda.synchronize()
        lat = []
        t0 = time.perf_counter()
        for k, (src, dst, (img, col)) in enumerate(zip(srcs, dsts, batch)):
            a = time.perf_counter()
            p = sa.Palette.build(src, img.shape[1], img.shape[0], sa.PIX_RGBA8, col, True,
                                 synth.is_gray(img))
            sa.walk_seq_batch([sa.image_spec(src, dst, img.shape[1], img.shape[0], sa.PIX_RGBA8,
                                             sa.PIX_RGBA8, p)])
            torch.cuda.synchronize()
            lat.append(time.perf_counter() - a)
        t_plain = time.perf_counter() - t0
        plain = {"value": round(pixels / t_plain / 1e6, 3), "ms": round(t_plain * 1e3, 2),
                 "latency_ms": {"p50": round(float(np.percentile(lat, 50)) * 1e3, 3),
                                "p99": round(float(np.percentile(lat, 99)) * 1e3, 3)}}
        log(f"plain: {t_plain * 1e3:.1f} ms  {pixels / t_plain / 1e6:.1f} Mpx/s")

    cpu = None
    if not args.no_cpu:
        nproc = min(os.cpu_count() or 1, 16)
        kind, cpx, res = cpu_baseline(batch, [1, nproc])
        one = res[1]
        allc = res[nproc]
        cpu = {"value": round(cpx / one["total_s"] / 1e6, 3), "unit": "Mpixel/s", "cores": 1,
               "kind": kind,
               "sample": f"the same {len(batch)} images ({cpx} px): ColorTree::Build + RiemersmaWalkCpu, one thread",
               "walk_only_value": round(cpx / one["walk_s"] / 1e6, 3),
               "all_cores": {"cores": nproc, "value": round(cpx / allc["total_s"] / 1e6, 3),
                             "walk_only_value": round(cpx / allc["walk_s"] / 1e6, 3)}}
        log(f"cpu ({kind}): 1 core {cpu['value']} Mpx/s, {nproc} cores {cpu['all_cores']['value']} Mpx/s")

    algo_bytes = 8.0 * pixels  # RGBA8 in + RGBA8 out per visit
    achieved = algo_bytes / (walk_ms * 1e-3) / 1e9
    # HBM bytes per 64-image walk launch, measured with rocprofv3 --pmc
    # FETCH_SIZE / WRITE_SIZE on this same workload (scaled by a copy kernel).
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "seq_walk_c2_traffic.json")
    if os.path.exists(tpath) and args.images == 64 and args.seed == 2:
        with open(tpath) as fh:
            traffic = round(json.load(fh)["traffic_bytes"])
    line = {
        "metric": METRIC,
        "value": round(pixels / t_batched / 1e6, 3),
        "unit": "Mpixel/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_batched * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{len(batch)} seeded RGBA8 stills, w,h uniform in [1,2048] "
                               "(incl. 1x1, 3x5, 17x13), octree palette 2..256 colours built on "
                               "the device, exact sequential Riemersma walk, RGBA8 out",
                   "images": len(batch), "pixels": pixels, "engine": "sequential",
                   "form": "batched", "device": dev},
        "walk_only": {"value": round(pixels / t_walk / 1e6, 3), "ms": round(t_walk * 1e3, 3),
                      "kernel_ms": round(walk_ms, 3), "gather_ms": round(gather_ms, 3),
                      "pack_ms": round(pack_ms, 3)},
        "chain": args.chain,
        "latency_ms": chain_lat,
        "served": served,
        "serving_mix": mix,
        "engines": engines,
        "palette_build_ms": round(t_pal * 1e3, 3),
        "palette_build_serial_ms": round(t_pal_serial * 1e3, 3),
        "plain": plain,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": traffic,
                     "kernel": "seq_walk_kernel", "note": "latency-bound serial chain; achieved is "
                     "algorithmic (8 B/visit) over the whole 64-image launch; traffic is the "
                     "launch's measured HBM bytes (profiles/seq_walk_c2_batch64.md)"},
        "cpu_baseline": cpu,
    }
    if cpu:
        line["gpu_vs_cpu_1core"] = round(line["value"] / cpu["value"], 2)
    if plain:
        line["batched_vs_plain"] = round(line["value"] / plain["value"], 2)
    if served:
        if cpu:
            served["vs_cpu_1core"] = round(served["value"] / cpu["value"], 2)
            served["vs_cpu_all_cores"] = round(served["value"] / cpu["all_cores"]["value"], 2)
        if plain:
            served["vs_plain"] = round(served["value"] / plain["value"], 2)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
