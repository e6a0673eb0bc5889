"""Continuous batching for the dither engines.

A sequential walk is a serial chain per image, so the device is only busy
when many images walk at once.  `SeqPipeline` keeps one resumable walk job
(sa_walk_job_*) stepping on its own stream: in every step each admitted image
walks for a time budget (a few milliseconds, at its own pace), images whose
palette has just been classified join at the next step, and images that
finish leave with their output packed.  Short steps also bound how long any
other stream's work can wait behind the walk on a shared hardware queue.

Palettes are classified by a `PaletteBuilder`: whatever has arrived is built
as one batch (sa_palette_build_batch: one launch per kernel for all images),
and requests that arrive meanwhile join the next batch.  A request that finds
a builder thread free (fewer requests queued or building than threads) is
built on its own, so the first images of a burst -- the largest, when a batch
is submitted largest first -- start walking after their own builds rather
than after a whole batch's.  A few builder threads, one stream each, keep batches in flight
while another reads back.

`Server` admits mixed requests -- stills and video segments, for the
sequential, block and Jacobi engines -- against an HBM budget, runs the block
and Jacobi engines on a stream of their own, keeps a segment's palette
resident across its frames, and records submit-to-done latency per request.

Every result is the engine's own: the same image gives the same bytes alone
and in any mix (tests/test_gpu_serve.py).
"""
from __future__ import annotations

import collections
import queue
import threading
import time
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L


class PaletteBuilder:
    """Continuous batching of the octree build.  submit() queues one image;
    its callback(palette, error) runs on a builder thread once the batch that
    took it has built."""

    def __init__(self, threads=2, max_batch_px=1 << 25, max_items=128):
        import torch
        self.max_px = int(max_batch_px)
        self.max_items = int(max_items)
        # Requests in submission order; an item that does not fit a batch goes
        # back to the front, so it leads the next batch.
        self._q = collections.deque()
        self._cv = threading.Condition()
        self._stop = False
        self._building = 0               # batches taken and not yet built
        self._nthreads = max(1, threads)
        self.batches = 0
        self._threads = []
        for _ in range(max(1, threads)):
            th = threading.Thread(target=self._loop, args=(torch.cuda.Stream(),), daemon=True)
            th.start()
            self._threads.append(th)

    def submit(self, src, width, height, src_fmt, colors, assoc, gray, callback):
        with self._cv:
            # A free builder thread: this request is built alone (see the module
            # docstring).
            solo = len(self._q) + self._building < self._nthreads
            self._q.append((src, width, height, src_fmt, colors, assoc, gray, callback, solo))
            self._cv.notify()

    def _take(self):
        with self._cv:
            while not self._q and not self._stop:
                self._cv.wait()
            if not self._q:
                return None  # stopped and drained
            first = self._q.popleft()
            self._building += 1
            got, px = [first], first[1] * first[2]
            while self._q and len(got) < self.max_items and not first[8]:
                nxt = self._q[0]
                if px + nxt[1] * nxt[2] > self.max_px:
                    break
                got.append(self._q.popleft())
                px += nxt[1] * nxt[2]
            return got

    def _loop(self, stream):
        while True:
            got = self._take()
            if got is None:
                return
            items = [g[:7] for g in got]
            try:
                pals, status = L.Palette.build_batch(items, stream=stream, strict=False)
                self.batches += 1
                errs = [None if p is not None else f"palette build failed (SA code {c})"
                        for p, c in zip(pals, status)]
            except L.EngineError as e:
                pals, errs = [None] * len(got), [str(e)] * len(got)
            finally:
                with self._cv:
                    self._building -= 1
            for g, pal, err in zip(got, pals, errs):
                g[7](pal, err)

    def close(self):
        with self._cv:
            self._stop = True
            self._cv.notify_all()
        for th in self._threads:
            th.join()


@dataclass
class Still:
    """One image: src/dst are contiguous device tensors in src_fmt/dst_fmt."""
    src: object
    dst: object
    width: int
    height: int
    src_fmt: int
    dst_fmt: int
    colors: int
    assoc: bool = False
    gray: bool = False
    palette: object = None          # a resident Palette; built on the device when None
    # filled in by the pipeline
    id: int = -1
    t_submit: float = 0.0
    t_ready: float = 0.0            # palette available
    t_done: float = 0.0
    error: str = ""

    @property
    def pixels(self):
        return self.width * self.height


class SeqPipeline:
    """Sequential engine with continuous batching (see the module docstring)."""

    def __init__(self, build_threads=2, budget_us=8000.0, options=None, on_done=None,
                 slice_visits=0, builder=None):
        import torch
        self.budget_us = float(budget_us)
        self.slice = int(slice_visits)      # > 0: fixed visit slices instead of a time budget
        self.options = options or L.SeqOptions()
        self.on_done = on_done
        self._own_builder = builder is None
        self.builder = builder or PaletteBuilder(threads=build_threads)
        self._walk_stream = torch.cuda.Stream()
        self._ready = queue.Queue()
        self._live = {}                  # job id -> Still
        self._pending = 0                # submitted, not yet joined the walk
        self._lock = threading.Lock()
        self._job = L.WalkJob(self.options)
        self._finishing = []             # (event after the pack, finished images)
        self.steps = 0

    # -- submission (any thread) ------------------------------------------------
    def submit(self, st: Still):
        st.t_submit = st.t_submit or time.perf_counter()
        with self._lock:
            self._pending += 1
        if st.palette is not None:
            st.t_ready = st.t_submit
            self._ready.put(st)
        else:
            def built(pal, err, st=st):
                st.palette = pal
                st.error = err or ""
                st.t_ready = time.perf_counter()
                self._ready.put(st)
            self.builder.submit(st.src, st.width, st.height, st.src_fmt, st.colors, st.assoc,
                                st.gray, built)

    # -- the stepping loop (one thread) ------------------------------------------
    def _admit(self, block):
        got = []
        try:
            got.append(self._ready.get(timeout=0.05) if block else self._ready.get_nowait())
            while True:
                got.append(self._ready.get_nowait())
        except queue.Empty:
            pass
        ok = []
        for st in got:
            with self._lock:
                self._pending -= 1
            if st.error:
                st.t_done = time.perf_counter()
                if self.on_done:
                    self.on_done(st)
                continue
            ok.append(st)
        if ok:
            specs = [L.image_spec(st.src, st.dst, st.width, st.height, st.src_fmt, st.dst_fmt,
                                  st.palette) for st in ok]
            ids = self._job.add(specs, stream=self._walk_stream)
            for i, st in zip(ids, ok):
                st.id = i
                self._live[i] = st
        return len(got)

    def idle(self):
        with self._lock:
            return self._pending == 0 and not self._live

    def _finish(self, block):
        """Reports the images whose packed outputs are complete."""
        keep = []
        for ev, sts in self._finishing:
            if block:
                ev.synchronize()
            if block or ev.query():
                now = time.perf_counter()
                for st in sts:
                    st.t_done = now
                    if self.on_done:
                        self.on_done(st)
            else:
                keep.append((ev, sts))
        self._finishing = keep

    def run(self, until_idle=True, stop=None):
        """Steps until nothing is pending or live (or `stop` is set)."""
        import torch
        self._finishing = []
        while True:
            if stop is not None and stop.is_set():
                break
            # Wait for a palette only when nothing walks and something is still
            # being classified (waiting otherwise would idle out the timeout at
            # the end of every run).
            with self._lock:
                waiting = self._pending > 0
            self._admit(block=not self._live and waiting)
            self._finish(block=False)
            if not self._live:
                self._finish(block=True)
                if until_idle and self.idle():
                    break
                continue
            if self.slice > 0:
                done = self._job.step(self.slice, stream=self._walk_stream)
            else:
                # Every image walks for the budget at its own pace; the host
                # collects the previous step while this one runs.
                done = self._job.step_overlapped(self.budget_us, stream=self._walk_stream)
            self.steps += 1
            if done:
                ev = torch.cuda.Event()
                ev.record(self._walk_stream)
                self._finishing.append((ev, [self._live.pop(i) for i in done]))
        self._finish(block=True)

    def close(self):
        if self._own_builder:
            self.builder.close()
        self._job.close()


# ---- the serving loop ------------------------------------------------------------------


@dataclass
class Request:
    """A dither job: a still (frames == 1) or a video segment sharing one
    palette (built from frame 0 and kept on the device for all its frames)."""
    engine: str                      # "sequential" | "blocks" | "approx"
    frames: list                     # [(src, dst)] device tensors, one geometry
    width: int
    height: int
    src_fmt: int
    dst_fmt: int
    colors: int
    assoc: bool = False
    block: int = 64
    iterations: int = 8
    taps: int = 96
    t_arrive: float = 0.0
    t_start: float = 0.0
    t_done: float = 0.0
    error: str = ""
    palette: object = None
    bytes: int = 0                   # device footprint charged at admission
    _left: int = 0
    _lock: object = field(default_factory=threading.Lock)

    @property
    def pixels(self):
        return self.width * self.height * len(self.frames)


def footprint(req: Request) -> int:
    """Device bytes a request holds while admitted (its frames' scratch)."""
    px = req.width * req.height
    if req.engine == "sequential":
        memo = (2 << 24) if req.assoc else (2 << 18)
        # staged inputs: 4 B per pixel for 8-bit sources, 8 for 16-bit, 16 for float
        stage = {1: 4, 2: 4, 3: 8, 4: 8}.get(req.src_fmt, 16)
        return len(req.frames) * (px * (stage + 2) + memo + (132 << 10))
    if req.engine == "approx":
        return len(req.frames) * px * (16 + 2) + (1 << 20)
    return len(req.frames) * px * 16 + (1 << 20)


class Server:
    """Admits requests by HBM footprint, overlaps them on the device and
    records latency.  `run(arrivals)` takes (t_offset, Request) pairs.

    Palettes of every engine come from one PaletteBuilder.  Block and Jacobi
    requests are launched in arrival-of-palette order on a stream of their
    own without waiting for each other; a finisher thread stamps each one done
    when its event completes."""

    def __init__(self, hbm_budget=64 << 30, build_threads=2, budget_us=8000.0, slice_visits=0,
                 on_done=None, seq_options=None):
        import torch
        self.budget = hbm_budget
        self.used = 0
        self.done = []
        self.on_done = on_done           # called with each finished Request (any thread)
        self.admission_wait_s = 0.0      # time arrivals spent waiting for the HBM budget
        self._cv = threading.Condition()
        self.builder = PaletteBuilder(threads=build_threads)
        self._seq = SeqPipeline(budget_us=budget_us, on_done=self._seq_done,
                                slice_visits=slice_visits, builder=self.builder,
                                options=seq_options)
        self._fast_stream = torch.cuda.Stream()
        self._fast = ThreadPoolExecutor(max_workers=1)
        self._events = queue.Queue()
        self._finisher = threading.Thread(target=self._finish_fast, daemon=True)
        self._finisher.start()

    def _release(self, req):
        if self.on_done:
            self.on_done(req)
        with self._cv:
            self.used -= req.bytes
            self.done.append(req)
            self._cv.notify_all()

    def _seq_done(self, st):
        req = st.request
        if st.error:
            req.error = st.error
        with req._lock:
            req._left -= 1
            last = req._left == 0
        if last:
            req.t_done = st.t_done
            self._release(req)

    def _fail(self, req, err):
        req.error = err
        req.t_done = time.perf_counter()
        self._release(req)

    def _launch_fast(self, req):
        """Block / Jacobi work on the fast stream; completion is stamped by the finisher."""
        import torch
        try:
            specs = [L.image_spec(src, dst, req.width, req.height, req.src_fmt, req.dst_fmt,
                                  req.palette) for src, dst in req.frames]
            if req.engine == "blocks":
                L.walk_blocks(specs, block=req.block, stream=self._fast_stream)
            else:
                L.walk_approx(specs, iterations=req.iterations, taps=req.taps,
                              stream=self._fast_stream)
            ev = torch.cuda.Event()
            ev.record(self._fast_stream)
            self._events.put((ev, req))
        except L.EngineError as e:
            self._fail(req, str(e))

    def _finish_fast(self):
        while True:
            got = self._events.get()
            if got is None:
                return
            ev, req = got
            ev.synchronize()
            req.t_done = time.perf_counter()
            self._release(req)

    def _start(self, req):
        req.t_start = time.perf_counter()
        if req.engine == "sequential":
            req._left = len(req.frames)
            if len(req.frames) == 1:
                st = Still(req.frames[0][0], req.frames[0][1], req.width, req.height,
                           req.src_fmt, req.dst_fmt, req.colors, req.assoc)
                st.request = req
                self._seq.submit(st)
                return

            def seg(pal, err):
                if err:
                    self._fail(req, err)
                    return
                req.palette = pal
                for src, dst in req.frames:
                    st = Still(src, dst, req.width, req.height, req.src_fmt, req.dst_fmt,
                               req.colors, req.assoc, palette=pal)
                    st.request = req
                    self._seq.submit(st)
        else:
            def seg(pal, err):
                if err:
                    self._fail(req, err)
                    return
                req.palette = pal
                self._fast.submit(self._launch_fast, req)
        src0 = req.frames[0][0]
        self.builder.submit(src0, req.width, req.height, req.src_fmt, req.colors, req.assoc,
                            False, seg)

    def run(self, arrivals, stop_after=None, feed_until=None):
        """Feeds (t_offset, Request) pairs at their times, admitting each when
        its footprint fits the budget (first come, first served); returns the
        finished requests.  The second element may instead be a callable that
        returns the Request when it arrives (so a long arrival schedule does not
        hold every request's buffers at once).  With `feed_until` (seconds),
        arrivals not admitted by then are not started; `self.dropped` counts
        them."""
        arrivals = sorted(arrivals, key=lambda a: a[0])
        stop = threading.Event()
        driver = threading.Thread(target=self._seq.run, kwargs={"until_idle": False, "stop": stop})
        driver.start()
        t0 = time.perf_counter()
        self.t0 = t0
        self.dropped = 0
        started = 0
        try:
            for t_off, req in arrivals:
                now = time.perf_counter() - t0
                if feed_until is not None and now > feed_until:
                    self.dropped = len(arrivals) - started
                    break
                if t_off > now:
                    time.sleep(t_off - now)
                if callable(req):
                    req = req()
                req.t_arrive = t0 + t_off
                req.bytes = footprint(req)
                with self._cv:
                    w0 = time.perf_counter()
                    while self.used + req.bytes > self.budget and self.used > 0:
                        self._cv.wait(0.05)
                    self.admission_wait_s += time.perf_counter() - w0
                    self.used += req.bytes
                self._start(req)
                started += 1
            with self._cv:
                while len(self.done) < started:
                    self._cv.wait(0.05)
                    if stop_after and time.perf_counter() - t0 > stop_after:
                        break
        finally:
            stop.set()
            driver.join()
        return list(self.done)

    def close(self):
        self._fast.shutdown(wait=True)
        self._events.put(None)
        self._finisher.join()
        self._seq.close()
        self.builder.close()


def latency_summary(reqs):
    lat = np.array([r.t_done - r.t_arrive for r in reqs if r.t_done > 0])
    if lat.size == 0:
        return {}
    return {"n": int(lat.size), "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 3),
            "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 3),
            "max_ms": round(float(lat.max()) * 1e3, 3)}
