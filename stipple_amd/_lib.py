"""ctypes binding of include/stipple_amd.h.

Loads the in-tree libstipple_amd.so; a missing or unloadable library is an
error (ImportError at first use), never a silent fallback.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "libstipple_amd.so")

PIX_RGBAF32, PIX_RGB8, PIX_RGBA8, PIX_RGB16, PIX_RGBA16, PIX_INDEX8, PIX_INDEX16 = range(7)
# Planar 8-bit YUV (block engine only): 4:4:4 in and out, 4:2:0 in.
PIX_YUV444, PIX_YUV420 = 7, 8


# Return codes of the C-ABI (SA_* in include/stipple_amd.h).
SA_OK, SA_EINVAL, SA_EREFUSED, SA_EDEVICE, SA_ENOMEM, SA_EUNSUPPORTED = range(6)


class EngineError(RuntimeError):
    """An engine refused its input or failed on the device (code, message)."""

    def __init__(self, code, message):
        super().__init__(f"[{code}] {message}")
        self.code = code
        self.message = message


class _Image(C.Structure):
    _fields_ = [("src", C.c_void_p), ("dst", C.c_void_p), ("width", C.c_int64),
                ("height", C.c_int64), ("src_fmt", C.c_int32), ("dst_fmt", C.c_int32),
                ("palette", C.c_void_p)]


class SaBuild(C.Structure):
    """sa_build: one image of a batched palette build."""
    _fields_ = [("src", C.c_void_p), ("src_fmt", C.c_int), ("width", C.c_int64),
                ("height", C.c_int64), ("colors", C.c_int), ("assoc", C.c_int), ("gray", C.c_int),
                ("keep_tree", C.c_int)]


class SeqOptions(C.Structure):
    _fields_ = [("diffusion", C.c_double), ("use_cache", C.c_int32),
                ("memo_bits", C.c_int32), ("waves", C.c_int32), ("reserved", C.c_int32)]

    def __init__(self, diffusion=1.0, use_cache=True, memo_bits=0, waves=0):
        super().__init__(diffusion, int(use_cache), memo_bits, waves, 0)


_lib = None
_lock = threading.Lock()


def library_path():
    return _SO


def lib():
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(_SO):
                raise ImportError(
                    f"{_SO} is not built; run __graft_entry__.build() (make -C stipple_amd/csrc)")
            L = C.CDLL(_SO)
            vp, i32, i64, dbl, cp, sz = C.c_void_p, C.c_int, C.c_int64, C.c_double, C.c_char_p, C.c_size_t
            sig = {
                "sa_available": (i32, []),
                "sa_init": (i32, [i32, cp, sz]),
                "sa_device_name": (i32, [cp, sz]),
                "sa_palette_create": (i32, [vp, i32, i32, vp, vp, vp, vp, i64, C.POINTER(vp), cp, sz]),
                "sa_palette_destroy": (i32, [vp]),
                "sa_palette_count": (i32, [vp]),
                "sa_curve_order": (i32, [i64, i64, vp, vp, cp, sz]),
                "sa_palette_build": (i32, [vp, i32, i64, i64, i32, i32, i32, i32, vp,
                                           C.POINTER(vp), cp, sz]),
                "sa_palette_build_batch": (i32, [C.POINTER(SaBuild), i32, vp, C.POINTER(vp),
                                                 C.POINTER(i32), cp, sz]),
                "sa_palette_info": (i32, [vp, C.POINTER(i32), C.POINTER(i32), C.POINTER(i32),
                                          C.POINTER(i64), C.POINTER(i32)]),
                "sa_palette_entries": (i32, [vp, vp, cp, sz]),
                "sa_palette_tree": (i32, [vp, vp, vp, vp, vp, vp, vp, vp, vp, cp, sz]),
                "sa_walk_seq_batch": (i32, [C.POINTER(_Image), i32, C.POINTER(SeqOptions), vp, cp, sz]),
                "sa_walk_blocks": (i32, [C.POINTER(_Image), i32, i32, dbl, vp, cp, sz]),
                "sa_walk_blocks_noalloc": (i32, [C.POINTER(_Image), i32, i32, dbl, vp, cp, sz]),
                "sa_video_montage": (i32, [vp, i32, i32, i64, i64, i32, i32, i32, vp,
                                           C.POINTER(i64), C.POINTER(i64), C.POINTER(i32), vp,
                                           cp, sz]),
                "sa_video_palette": (i32, [vp, i32, i32, i64, i64, i32, i32, i32, i32, i32, vp,
                                           C.POINTER(vp), cp, sz]),
                "sa_palette_write_txt": (i32, [vp, i32, cp, cp, sz]),
                "sa_palette_read_txt": (i32, [cp, i32, vp, C.POINTER(i32), C.POINTER(i32), cp,
                                              sz]),
                "sa_walk_approx": (i32, [C.POINTER(_Image), i32, i32, i32, i32, i32, dbl, vp, cp, sz]),
                "sa_dither_screen": (i32, [C.POINTER(_Image), i32, i32, dbl, vp, cp, sz]),
                "sa_dither_diffuse": (i32, [C.POINTER(_Image), i32, i32, dbl, vp, cp, sz]),
                "sa_seq_stats": (i32, [vp, cp, sz]),
                "sa_inverse_filter": (i32, [i32, dbl, vp, cp, sz]),
                "sa_walk_job_create": (i32, [C.POINTER(SeqOptions), C.POINTER(vp), cp, sz]),
                "sa_walk_job_add": (i32, [vp, C.POINTER(_Image), i32, vp, vp, cp, sz]),
                "sa_walk_job_step": (i32, [vp, i64, vp, vp, i32, C.POINTER(i32), C.POINTER(i32),
                                           cp, sz]),
                "sa_walk_job_step_timed": (i32, [vp, dbl, vp, vp, i32, C.POINTER(i32),
                                                 C.POINTER(i32), cp, sz]),
                "sa_walk_job_step_overlapped": (i32, [vp, dbl, vp, vp, i32, C.POINTER(i32),
                                                      C.POINTER(i32), cp, sz]),
                "sa_walk_job_destroy": (i32, [vp]),
                "sa_chain_create_blocks": (i32, [C.POINTER(_Image), i32, i32, dbl, vp,
                                                 C.POINTER(vp), cp, sz]),
                "sa_chain_create_approx": (i32, [C.POINTER(_Image), i32, i32, i32, i32, i32, dbl,
                                                 vp, C.POINTER(vp), cp, sz]),
                "sa_chain_run": (i32, [vp, vp, cp, sz]),
                "sa_chain_destroy": (i32, [vp]),
                "sa_tune_set": (i32, [cp, i32, cp, sz]),
                "sa_tune_get": (i32, [cp, C.POINTER(i32)]),
                "sa_timing_enable": (i32, [i32]),
                "sa_timing_collect": (i32, [cp, C.POINTER(dbl), C.POINTER(i32), cp, sz]),
            }
            for name, (res, args) in sig.items():
                fn = getattr(L, name)
                fn.restype = res
                fn.argtypes = args
            _lib = L
    return _lib


def _check(rc, buf):
    if rc != 0:
        raise EngineError(rc, buf.value.decode(errors="replace"))


def _err():
    return C.create_string_buffer(1024)


def available():
    return bool(lib().sa_available())


def init(device=0):
    e = _err()
    _check(lib().sa_init(device, e, len(e)), e)


def device_name():
    b = C.create_string_buffer(256)
    _check(lib().sa_device_name(b, len(b)), b)
    return b.value.decode()


def _np_ptr(a):
    return a.ctypes.data_as(C.c_void_p)


class Palette:
    """A palette and the octree its lookups search, resident on the device.

    Mirrors the (Palette, ColorTree) pair the reference engines take
    (include/rd_riemersma.h:63-67, include/rd_octree.h:54-123)."""

    def __init__(self, entries, assoc, child16, parent, color_number, number_unique):
        entries = np.ascontiguousarray(entries, np.float64).reshape(-1, 4)
        child16 = np.ascontiguousarray(child16, np.int32).reshape(-1, 16)
        parent = np.ascontiguousarray(parent, np.int32)
        color_number = np.ascontiguousarray(color_number, np.uint32)
        number_unique = np.ascontiguousarray(number_unique, np.uint64)
        h = C.c_void_p()
        e = _err()
        rc = lib().sa_palette_create(_np_ptr(entries), entries.shape[0], int(bool(assoc)),
                                     _np_ptr(child16), _np_ptr(parent), _np_ptr(color_number),
                                     _np_ptr(number_unique), child16.shape[0], C.byref(h), e, len(e))
        _check(rc, e)
        self.handle = h.value
        self.count = entries.shape[0]
        self.assoc = bool(assoc)
        self._entries = entries

    @classmethod
    def build(cls, src, width, height, src_fmt, colors, assoc, gray=False, keep_tree=False,
              stream=None):
        """ColorTree::Build on the device from a device image (sa_palette_build)."""
        h = C.c_void_p()
        e = _err()
        rc = lib().sa_palette_build(_dev_ptr(src), src_fmt, width, height, colors, int(bool(assoc)),
                                    int(bool(gray)), int(bool(keep_tree)), _stream_ptr(stream),
                                    C.byref(h), e, len(e))
        _check(rc, e)
        return cls._built(h.value, assoc)

    @classmethod
    def _built(cls, handle, assoc):
        self = cls.__new__(cls)
        self.handle = handle
        self.assoc = bool(assoc)
        self.count = self.info()["count"]
        self._entries = None
        return self

    @classmethod
    def build_batch(cls, items, stream=None, strict=True):
        """ColorTree::Build for many device images in one batched build
        (sa_palette_build_batch).  `items` are (src, width, height, src_fmt,
        colors, assoc[, gray[, keep_tree]]) tuples.  Returns one Palette per
        item; with strict, raises EngineError naming the first item that
        failed, otherwise returns (palettes with None where an item failed,
        the SA_* status of every item)."""
        n = len(items)
        arr = (SaBuild * max(n, 1))()
        keep = []
        for k, it in enumerate(items):
            src, w, h, fmt, colors, assoc = it[:6]
            gray = it[6] if len(it) > 6 else False
            kt = it[7] if len(it) > 7 else False
            keep.append(src)
            arr[k] = SaBuild(_dev_ptr(src), fmt, w, h, colors, int(bool(assoc)), int(bool(gray)),
                             int(bool(kt)))
        out = (C.c_void_p * max(n, 1))()
        status = (C.c_int * max(n, 1))()
        e = _err()
        rc = lib().sa_palette_build_batch(arr, n, _stream_ptr(stream), out, status, e, len(e))
        pals = [cls._built(out[k], items[k][5]) if out[k] else None for k in range(n)]
        if not strict:
            return pals, [int(status[k]) for k in range(n)]
        if rc != 0:
            for p in pals:
                if p is not None:
                    p.close()
            _check(rc, e)
        return pals

    @classmethod
    def from_video(cls, frames, width, height, src_fmt, colors, want=None, tile=0, cols=0,
                   keep_tree=False, stream=None):
        """VideoBuildPalette's montage palette (sa_video_palette): `frames` are
        the sampled device frames in order, `want` the montage's cell count
        (default: len(frames))."""
        ptrs, keep = _frame_ptrs(frames)
        h = C.c_void_p()
        e = _err()
        rc = lib().sa_video_palette(ptrs, len(frames), src_fmt, width, height,
                                    want or len(frames), tile, cols, colors, int(bool(keep_tree)),
                                    _stream_ptr(stream), C.byref(h), e, len(e))
        _check(rc, e)
        return cls._built(h.value, False)

    @classmethod
    def import_txt(cls, path, expect_colors=0, keep_tree=False):
        """ImReadPalette's .txt form: the file's entries (narrowed, canonically
        sorted) as a count x 1 strip, its tree built on the device at `count`
        colours, the tree's colormap adopted.  Raises EngineError when the
        tree reduces the entries (duplicate or unreachable colours)."""
        import torch
        ent, _assoc = read_palette_txt(path, expect_colors)
        n = ent.shape[0]
        strip = torch.from_numpy(ent.astype(np.float32).reshape(-1).view(np.uint8).copy()).cuda()
        gray = bool(np.all((ent[:, 0] == ent[:, 1]) & (ent[:, 1] == ent[:, 2])))
        pal = cls.build(strip, n, 1, PIX_RGBAF32, n, False, gray, keep_tree)
        torch.cuda.synchronize()
        if pal.count != n:
            got = pal.count
            pal.close()
            raise EngineError(SA_EINVAL, f"{path}: the tree reduced {n} colours to {got}; the "
                                         "palette holds duplicate or unreachable entries")
        return pal

    @property
    def entries(self):
        if self._entries is None:
            self._entries = self.colormap()
        return self._entries

    def info(self):
        c, a, d, ex = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        n = C.c_int64()
        lib().sa_palette_info(self.handle, C.byref(c), C.byref(a), C.byref(d), C.byref(n),
                              C.byref(ex))
        return {"count": c.value, "assoc": bool(a.value), "depth": d.value, "nodes": n.value,
                "exact_root": bool(ex.value)}

    def colormap(self):
        out = np.zeros((self.info()["count"], 4), np.float64)
        e = _err()
        _check(lib().sa_palette_entries(self.handle, _np_ptr(out), e, len(e)), e)
        return out

    def tree(self):
        """Node arrays (needs keep_tree=True at build)."""
        n = self.info()["nodes"]
        child = np.zeros((n, 16), np.int32)
        parent = np.zeros(n, np.int32)
        cn = np.zeros(n, np.uint32)
        nu = np.zeros(n, np.uint64)
        tot = np.zeros((n, 4), np.float64)
        qe = np.zeros(n, np.float64)
        ids = np.zeros(n, np.uint32)
        lev = np.zeros(n, np.uint32)
        e = _err()
        _check(lib().sa_palette_tree(self.handle, _np_ptr(child), _np_ptr(parent), _np_ptr(cn),
                                     _np_ptr(nu), _np_ptr(tot), _np_ptr(qe), _np_ptr(ids),
                                     _np_ptr(lev), e, len(e)), e)
        return {"child": child, "parent": parent, "color_number": cn, "number_unique": nu,
                "total": tot, "qe": qe, "id": ids, "level": lev}

    @classmethod
    def from_tree(cls, tree):
        """From any object with colormap/child/parent/color_number/number_unique/assoc
        arrays (e.g. a TreeData)."""
        return cls(tree.colormap, tree.assoc, tree.child, tree.parent, tree.color_number,
                   tree.number_unique)

    def close(self):
        if getattr(self, "handle", None):
            lib().sa_palette_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _frame_ptrs(frames):
    arr = (C.c_void_p * max(len(frames), 1))(*[_dev_ptr(f) for f in frames])
    return arr, frames


def video_montage(frames, width, height, src_fmt, want=None, tile=0, cols=0, stream=None):
    """The palette montage of sampled device frames (sa_video_montage): returns
    (montage as an RGBA16 device tensor of mh x mw x 4 uint16 bytes, mw, mh,
    gray)."""
    import torch
    want = want or len(frames)
    mw, mh = C.c_int64(), C.c_int64()
    e = _err()
    ptrs, _keep = _frame_ptrs(frames)
    _check(lib().sa_video_montage(ptrs, len(frames), src_fmt, width, height, want, tile, cols,
                                  None, C.byref(mw), C.byref(mh), None, None, e, len(e)), e)
    out = torch.empty(8 * mw.value * mh.value, dtype=torch.uint8, device="cuda")
    gray = C.c_int()
    _check(lib().sa_video_montage(ptrs, len(frames), src_fmt, width, height, want, tile, cols,
                                  _dev_ptr(out), C.byref(mw), C.byref(mh), C.byref(gray),
                                  _stream_ptr(stream), e, len(e)), e)
    return out, mw.value, mh.value, bool(gray.value)


def write_palette_txt(entries, path):
    """ImWritePalette's .txt form (sa_palette_write_txt)."""
    ent = np.ascontiguousarray(entries, np.float64).reshape(-1, 4)
    e = _err()
    _check(lib().sa_palette_write_txt(_np_ptr(ent), ent.shape[0], os.fsencode(path), e, len(e)), e)


def read_palette_txt(path, expect_colors=0):
    """ImReadPalette's .txt parse (sa_palette_read_txt): (entries n x 4, assoc)."""
    buf = np.zeros((65536, 4), np.float64)
    n, a = C.c_int(), C.c_int()
    e = _err()
    _check(lib().sa_palette_read_txt(os.fsencode(path), expect_colors, _np_ptr(buf), C.byref(n),
                                     C.byref(a), e, len(e)), e)
    return buf[:n.value].copy(), bool(a.value)


def _dev_ptr(t):
    import torch
    assert isinstance(t, torch.Tensor) and t.is_cuda and t.is_contiguous(), "device tensor required"
    return t.data_ptr()


def image_spec(src, dst, width, height, src_fmt, dst_fmt, palette):
    """One image for the engines; src/dst are contiguous torch device tensors.
    The spec keeps them and the palette referenced (`_keep`), so a job or a
    chain that holds the spec keeps what its device work reads and writes."""
    spec = _Image(_dev_ptr(src), _dev_ptr(dst), width, height, src_fmt, dst_fmt, palette.handle)
    spec._keep = (src, dst, palette)
    return spec


def _stream_ptr(stream):
    if stream is None:
        import torch
        stream = torch.cuda.current_stream()
    return stream.cuda_stream


def walk_seq_batch(images, options=None, stream=None):
    """The exact sequential walk over a batch (sa_walk_seq_batch)."""
    arr = (_Image * len(images))(*images)
    opt = options or SeqOptions()
    e = _err()
    rc = lib().sa_walk_seq_batch(arr, len(images), C.byref(opt), _stream_ptr(stream), e, len(e))
    _check(rc, e)


class WalkJob:
    """A resumable batch of sequential walks (sa_walk_job_*): images join with
    add() while others are on the device; each step() advances every unfinished
    image by a slice of visits and packs the ones that finish.  Every image is
    bit-exact with one uninterrupted walk."""

    def __init__(self, options=None):
        opt = options or SeqOptions()
        h = C.c_void_p()
        e = _err()
        _check(lib().sa_walk_job_create(C.byref(opt), C.byref(h), e, len(e)), e)
        self.handle = h.value
        self.active = 0
        # What each unfinished image's device work uses (its spec's tensors and
        # palette), and what finished in the last step: released one call
        # later, when the stream has passed that step.
        self._refs = {}
        self._done_refs = []

    def _finished(self, ids, synced):
        released = [self._refs.pop(i, None) for i in ids]
        if synced:
            self._done_refs = []
        else:
            self._done_refs = released

    def add(self, images, stream=None):
        """Adds image_spec()s; returns their ids."""
        if not images:
            return []
        arr = (_Image * len(images))(*images)
        ids = np.zeros(len(images), np.int64)
        e = _err()
        _check(lib().sa_walk_job_add(self.handle, arr, len(images), _np_ptr(ids),
                                     _stream_ptr(stream), e, len(e)), e)
        self.active += len(images)
        for i, img in zip(ids, images):
            self._refs[int(i)] = getattr(img, "_keep", None)
        return [int(i) for i in ids]

    def step(self, visits, stream=None):
        """One launch; returns the ids finished by it (outputs ready when the
        stream reaches this point)."""
        cap = max(self.active, 1)
        done = np.zeros(cap, np.int64)
        nd = C.c_int()
        act = C.c_int()
        e = _err()
        _check(lib().sa_walk_job_step(self.handle, int(visits), _stream_ptr(stream), _np_ptr(done),
                                      cap, C.byref(nd), C.byref(act), e, len(e)), e)
        self.active = act.value
        ids = [int(i) for i in done[:nd.value]]
        self._finished(ids, synced=False)
        return ids

    def step_timed(self, budget_us, stream=None):
        """One launch in which every image walks for about `budget_us`
        microseconds; synchronises the stream and returns the finished ids."""
        cap = max(self.active, 1)
        done = np.zeros(cap, np.int64)
        nd = C.c_int()
        act = C.c_int()
        e = _err()
        _check(lib().sa_walk_job_step_timed(self.handle, float(budget_us), _stream_ptr(stream),
                                            _np_ptr(done), cap, C.byref(nd), C.byref(act), e,
                                            len(e)), e)
        self.active = act.value
        ids = [int(i) for i in done[:nd.value]]
        self._finished(ids, synced=True)
        return ids

    def step_overlapped(self, budget_us, stream=None):
        """Launches a timed step, then returns the ids that finished in the
        previous one (their outputs are ready when the stream reaches this
        point); the host waits only for the previous step."""
        cap = max(self.active, 1)
        done = np.zeros(cap, np.int64)
        nd = C.c_int()
        act = C.c_int()
        e = _err()
        _check(lib().sa_walk_job_step_overlapped(self.handle, float(budget_us),
                                                 _stream_ptr(stream), _np_ptr(done), cap,
                                                 C.byref(nd), C.byref(act), e, len(e)), e)
        self.active = act.value
        ids = [int(i) for i in done[:nd.value]]
        self._finished(ids, synced=True)  # those finished in the step already waited for
        return ids

    def close(self):
        if getattr(self, "handle", None):
            lib().sa_walk_job_destroy(self.handle)  # synchronises the device first
            self.handle = None
        self._refs = {}
        self._done_refs = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def inverse_filter(taps, diffusion=1.0):
    """g[0..taps] of the Jacobi engine (sa_inverse_filter)."""
    g = np.zeros(taps + 1, np.float64)
    e = _err()
    _check(lib().sa_inverse_filter(taps, diffusion, _np_ptr(g), e, len(e)), e)
    return g


def walk_blocks(frames, block=64, diffusion=1.0, stream=None):
    """The block-parallel walk over frames of one geometry (sa_walk_blocks)."""
    arr = (_Image * len(frames))(*frames)
    e = _err()
    rc = lib().sa_walk_blocks(arr, len(frames), block, diffusion, _stream_ptr(stream), e, len(e))
    _check(rc, e)


def walk_blocks_noalloc(frames, block=64, diffusion=1.0, stream=None):
    """walk_blocks with the library's fixed device scratch (sa_walk_blocks_noalloc):
    no device allocation; calls that overlap on one device race."""
    arr = (_Image * len(frames))(*frames)
    e = _err()
    rc = lib().sa_walk_blocks_noalloc(arr, len(frames), block, diffusion, _stream_ptr(stream), e,
                                      len(e))
    _check(rc, e)


SCREEN_BAYER8, SCREEN_VOID64, SCREEN_VOID32, SCREEN_CLUSTERED8 = range(4)
SCREENS = {"bayer-ordered": SCREEN_BAYER8, "void-and-cluster": SCREEN_VOID64,
           "void-and-cluster-fast": SCREEN_VOID32, "clustered-dot": SCREEN_CLUSTERED8}


DIFFUSE_FLOYD, DIFFUSE_ATKINSON, DIFFUSE_JARVIS = range(3)
DIFFUSERS = {"floyd-steinberg": DIFFUSE_FLOYD, "atkinson": DIFFUSE_ATKINSON,
             "jarvis": DIFFUSE_JARVIS}


def dither_diffuse(frames, kind, diffusion=1.0, stream=None):
    """An error-diffusion plugin dither (sa_dither_diffuse) over frames of one
    geometry; `kind` is DIFFUSE_* or a plugin name."""
    if isinstance(kind, str):
        kind = DIFFUSERS[kind]
    arr = (_Image * len(frames))(*frames)
    e = _err()
    _check(lib().sa_dither_diffuse(arr, len(frames), kind, diffusion, _stream_ptr(stream), e,
                                   len(e)), e)


def dither_screen(frames, screen, diffusion=1.0, stream=None):
    """A threshold-screen dither of the plugin seam over device frames
    (sa_dither_screen); `screen` is a SCREEN_* value or a plugin name."""
    if isinstance(screen, str):
        screen = SCREENS[screen]
    arr = (_Image * max(len(frames), 1))(*frames)
    e = _err()
    _check(lib().sa_dither_screen(arr, len(frames), screen, diffusion, _stream_ptr(stream), e,
                                  len(e)), e)


def walk_approx(frames, iterations=20, taps=96, fp64=False, use_tree=True, diffusion=1.0,
                stream=None):
    """The Jacobi approximation over frames of one geometry (sa_walk_approx)."""
    arr = (_Image * len(frames))(*frames)
    e = _err()
    rc = lib().sa_walk_approx(arr, len(frames), iterations, taps, int(bool(fp64)),
                              int(bool(use_tree)), diffusion, _stream_ptr(stream), e, len(e))
    _check(rc, e)


class FrameChain:
    """An engine's work over a fixed batch of frame buffers, captured once
    into a hipGraph (sa_chain_*): refill the src buffers, then run() replays
    the batch with one launch."""

    def __init__(self, handle, frames=()):
        self.handle = handle
        # the frames' tensors and palettes, which every replay reads and writes
        self._keep = [getattr(f, "_keep", None) for f in frames]

    @classmethod
    def blocks(cls, frames, block=64, diffusion=1.0, stream=None):
        arr = (_Image * len(frames))(*frames)
        h = C.c_void_p()
        e = _err()
        rc = lib().sa_chain_create_blocks(arr, len(frames), block, diffusion, _stream_ptr(stream),
                                          C.byref(h), e, len(e))
        _check(rc, e)
        return cls(h.value, frames)

    @classmethod
    def approx(cls, frames, iterations=20, taps=96, fp64=False, use_tree=True, diffusion=1.0,
               stream=None):
        arr = (_Image * len(frames))(*frames)
        h = C.c_void_p()
        e = _err()
        rc = lib().sa_chain_create_approx(arr, len(frames), iterations, taps, int(bool(fp64)),
                                          int(bool(use_tree)), diffusion, _stream_ptr(stream),
                                          C.byref(h), e, len(e))
        _check(rc, e)
        return cls(h.value, frames)

    def run(self, stream=None):
        e = _err()
        _check(lib().sa_chain_run(self.handle, _stream_ptr(stream), e, len(e)), e)

    def close(self):
        if getattr(self, "handle", None):
            lib().sa_chain_destroy(self.handle)
            self.handle = None
        self._keep = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def curve_order(width, height, stream=None):
    """Device Hilbert visit order (int32 torch tensor of width*height)."""
    import torch
    out = torch.empty(width * height, dtype=torch.int32, device="cuda")
    e = _err()
    rc = lib().sa_curve_order(width, height, out.data_ptr(), _stream_ptr(stream), e, len(e))
    _check(rc, e)
    return out


def tune_set(knob, value):
    """Sets a launch-shape knob (sa_tune_set); no knob changes a result."""
    e = _err()
    _check(lib().sa_tune_set(knob.encode(), int(value), e, len(e)), e)


def tune_get(knob):
    v = C.c_int()
    rc = lib().sa_tune_get(knob.encode(), C.byref(v))
    if rc != 0:
        raise EngineError(rc, f"unknown knob {knob}")
    return v.value


def timing_enable(on=True):
    """Bracket engine kernels with HIP events (sa_timing_enable)."""
    lib().sa_timing_enable(int(bool(on)))


def timing_collect(name=None):
    """(total_ms, launches) of kernel `name` since the last collect."""
    ms = C.c_double()
    n = C.c_int()
    e = _err()
    _check(lib().sa_timing_collect(name.encode() if name else None, C.byref(ms), C.byref(n), e,
                                   len(e)), e)
    return ms.value, n.value


def seq_stats():
    """Walk counters since the last call: misses, line fills, HBM line loads."""
    out = np.zeros(4, np.uint64)
    e = _err()
    _check(lib().sa_seq_stats(_np_ptr(out), e, len(e)), e)
    return {"misses": int(out[0]), "fills": int(out[1]), "loads": int(out[2])}
