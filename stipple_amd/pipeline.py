"""The frame pipeline: reader -> device slots -> ordered writer.

The shape of the reference's video batch pipeline (src/rd_video.cpp:2785-3036,
:3967-4705): a reader fills free slots with batches of frames, the engine
dithers each slot, and a writer hands the frames back strictly in frame order
(each batch carries the index of its first frame, src/rd_video.cpp:2803-2811,
:4590-4606), returning the slot to the reader.

Built for one GPU:
  * A slot is a batch of frames resident in HBM (source and result) plus
    page-locked host staging for both directions (hipHostMalloc through torch),
    allocated once.  The reader fills the staging in place -- `read(buf)`
    writes one frame into `buf` -- so no frame is copied on the host.
  * Copies in, the walk and copies out run on three HIP streams, ordered by
    events, so one slot's upload and another's download overlap a third's
    walk.
  * With a resident palette (a video segment's), every slot's walk is captured
    once as a hipGraph over its own buffers (sa_chain_*): a batch is one graph
    launch.  Without one, each batch builds its frames' palettes on the device
    in one batched build and walks eagerly.  A short last batch walks eagerly.
  * The writer thread waits for a slot's download, calls `write(index, buf)`
    for its frames in order (`buf` stays valid during the call) and frees the
    slot.

Every frame's result is the engine's result for that frame alone; the test
checks it byte for byte.
"""
from __future__ import annotations

import queue
import threading
import time


from . import _lib as L

_BYTES = {L.PIX_RGBAF32: 16, L.PIX_RGB8: 3, L.PIX_RGBA8: 4, L.PIX_RGB16: 6, L.PIX_RGBA16: 8,
          L.PIX_INDEX8: 1, L.PIX_INDEX16: 2}


def frame_bytes(fmt, width, height):
    """Bytes of one frame in a pixel format, planar YUV included."""
    if fmt == L.PIX_YUV444:
        return 3 * width * height
    if fmt == L.PIX_YUV420:
        return width * height + 2 * (width // 2) * (height // 2)
    return _BYTES[fmt] * width * height


class _Slot:
    def __init__(self, torch, nb_in, nb_out, batch):
        self.dev_in = torch.empty(nb_in * batch, dtype=torch.uint8, device="cuda")
        self.dev_out = torch.empty(nb_out * batch, dtype=torch.uint8, device="cuda")
        self.host_in = torch.empty(nb_in * batch, dtype=torch.uint8, pin_memory=True)
        self.host_out = torch.empty(nb_out * batch, dtype=torch.uint8, pin_memory=True)
        self.np_in = self.host_in.numpy()
        self.np_out = self.host_out.numpy()
        self.uploaded = torch.cuda.Event()
        self.walked = torch.cuda.Event()
        self.downloaded = torch.cuda.Event()
        self.chain = None
        self.first = 0
        self.count = 0
        self.t_read = []  # per frame: when the reader finished it


class FramePipeline:
    """Frames of one geometry through the block engine (or the Jacobi engine,
    engine="approx"), `batch` frames per slot, `slots` slots in flight.

    palette: a resident Palette for every frame (a segment's), or None to
    build each frame's own palette of `colors` on the device."""

    def __init__(self, width, height, src_fmt, dst_fmt, palette=None, colors=16, batch=8,
                 slots=3, block=64, engine="blocks", iterations=20, taps=96, diffusion=1.0):
        import torch
        if batch < 1 or slots < 1:
            raise ValueError("batch and slots must be >= 1")
        if engine not in ("blocks", "approx"):
            raise ValueError("engine is 'blocks' or 'approx'")
        self.torch = torch
        self.w, self.h = width, height
        self.src_fmt, self.dst_fmt = src_fmt, dst_fmt
        self.palette, self.colors = palette, colors
        self.batch, self.block, self.engine = batch, block, engine
        self.iterations, self.taps, self.diffusion = iterations, taps, diffusion
        self.nb_in = frame_bytes(src_fmt, width, height)
        self.nb_out = frame_bytes(dst_fmt, width, height)
        self.s_in = torch.cuda.Stream()
        self.s_walk = torch.cuda.Stream()
        self.s_out = torch.cuda.Stream()
        self.slots = [_Slot(torch, self.nb_in, self.nb_out, batch) for _ in range(slots)]
        if palette is not None:
            for sl in self.slots:
                sl.chain = self._capture(sl)
        self.latency_s = []

    def _specs(self, sl, n, pals):
        w, h = self.w, self.h
        return [L.image_spec(sl.dev_in[k * self.nb_in:(k + 1) * self.nb_in],
                             sl.dev_out[k * self.nb_out:(k + 1) * self.nb_out],
                             w, h, self.src_fmt, self.dst_fmt, pals[k]) for k in range(n)]

    def _capture(self, sl):
        specs = self._specs(sl, self.batch, [self.palette] * self.batch)
        if self.engine == "blocks":
            return L.FrameChain.blocks(specs, block=self.block, diffusion=self.diffusion,
                                       stream=self.s_walk)
        return L.FrameChain.approx(specs, iterations=self.iterations, taps=self.taps,
                                   diffusion=self.diffusion, stream=self.s_walk)

    def _walk(self, sl):
        """The slot's batch on the walk stream, after its upload."""
        n = sl.count
        self.s_walk.wait_event(sl.uploaded)
        if sl.chain is not None and n == self.batch:
            sl.chain.run(stream=self.s_walk)
        else:
            if self.palette is not None:
                pals = [self.palette] * n
            else:
                pals = L.Palette.build_batch(
                    [(sl.dev_in[k * self.nb_in:(k + 1) * self.nb_in], self.w, self.h,
                      self.src_fmt, self.colors, False) for k in range(n)], stream=self.s_walk)
                sl.pals = pals  # alive until the slot is reused
            specs = self._specs(sl, n, pals)
            if self.engine == "blocks":
                L.walk_blocks(specs, block=self.block, diffusion=self.diffusion,
                              stream=self.s_walk)
            else:
                L.walk_approx(specs, iterations=self.iterations, taps=self.taps,
                              diffusion=self.diffusion, stream=self.s_walk)
        sl.walked.record(self.s_walk)
        nbytes = n * self.nb_out
        self.s_out.wait_event(sl.walked)
        with self.torch.cuda.stream(self.s_out):
            sl.host_out[:nbytes].copy_(sl.dev_out[:nbytes], non_blocking=True)
        sl.downloaded.record(self.s_out)

    def run(self, read, write):
        """read(buf) fills `buf` (a uint8 array of one frame's bytes in the
        source format) and returns True, or returns False at the end;
        write(index, buf) receives every result in frame order.  Returns the
        number of frames."""
        free = queue.Queue()
        for sl in self.slots:
            free.put(sl)
        ready = queue.Queue()
        done = {}
        cv = threading.Condition()
        err = []
        state = {"total": None}

        def reader():
            try:
                index = 0
                while True:
                    sl = free.get()
                    sl.first, sl.count, sl.t_read = index, 0, []
                    while sl.count < self.batch:
                        off = sl.count * self.nb_in
                        if not read(sl.np_in[off:off + self.nb_in]):
                            break
                        sl.count += 1
                        sl.t_read.append(time.perf_counter())
                    index += sl.count
                    if sl.count:
                        nbytes = sl.count * self.nb_in
                        with self.torch.cuda.stream(self.s_in):
                            sl.dev_in[:nbytes].copy_(sl.host_in[:nbytes], non_blocking=True)
                        sl.uploaded.record(self.s_in)
                        ready.put(sl)
                    if sl.count < self.batch:
                        ready.put(None)
                        with cv:
                            state["total"] = index
                            cv.notify_all()
                        return
            except BaseException as e:  # noqa: BLE001 -- reported by run()
                err.append(e)
                ready.put(None)

        def writer():
            try:
                nxt = 0
                while True:
                    with cv:
                        while nxt not in done and not (state["total"] is not None and
                                                       nxt >= state["total"]) and not err:
                            cv.wait()
                        if err or (nxt not in done):
                            return
                        sl = done.pop(nxt)
                    sl.downloaded.synchronize()
                    now = time.perf_counter()
                    for k in range(sl.count):
                        write(sl.first + k, sl.np_out[k * self.nb_out:(k + 1) * self.nb_out])
                        self.latency_s.append(now - sl.t_read[k])
                    nxt += sl.count
                    free.put(sl)
            except BaseException as e:  # noqa: BLE001 -- reported by run()
                err.append(e)
                with cv:
                    cv.notify_all()

        tr = threading.Thread(target=reader, daemon=True)
        tw = threading.Thread(target=writer, daemon=True)
        tr.start()
        tw.start()
        try:
            while not err:
                sl = ready.get()
                if sl is None:
                    break
                self._walk(sl)
                with cv:
                    done[sl.first] = sl
                    cv.notify_all()
        except BaseException as e:  # noqa: BLE001
            err.append(e)
        with cv:
            cv.notify_all()
        tw.join()
        if err:
            raise err[0]  # the reader may still wait for a slot; it is a daemon
        tr.join()
        return state["total"] or 0

    def close(self):
        for sl in self.slots:
            if sl.chain is not None:
                sl.chain.close()
                sl.chain = None
