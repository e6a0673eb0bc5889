"""ctypes bindings for the CPU checker (oracle/liboracle.so) and, where it was
built, the reference's own CPU sources (oracle/_ref/libref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py.  The product package never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(_HERE, "liboracle.so")
REF_SO = os.path.join(_HERE, "_ref", "libref.so")

_p = C.c_void_p
_i64 = C.c_int64
_i32 = C.c_int
_dbl = C.c_double


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def _rgba_f32(img):
    a = np.ascontiguousarray(img, dtype=np.float32)
    assert a.ndim == 3 and a.shape[2] == 4, a.shape
    return a


class TreeData:
    """Plain arrays describing a built octree (either implementation)."""

    def __init__(self, colormap, child, parent, color_number, number_unique,
                 total, qe, ids, level, depth, assoc):
        self.colormap = colormap            # (count, 4) float64
        self.child = child                  # (nodes, 16) int32
        self.parent = parent                # (nodes,) int32
        self.color_number = color_number    # (nodes,) uint32
        self.number_unique = number_unique  # (nodes,) uint64
        self.total = total                  # (nodes, 4) float64
        self.qe = qe                        # (nodes,) float64
        self.ids = ids
        self.level = level
        self.depth = depth
        self.assoc = assoc

    @property
    def count(self):
        return self.colormap.shape[0]

    @property
    def node_count(self):
        return self.parent.shape[0]


class _Lib:
    prefix = ""

    def __init__(self, path):
        self.lib = C.CDLL(path)

    def _f(self, name, res, *args):
        fn = getattr(self.lib, self.prefix + name)
        fn.restype = res
        fn.argtypes = list(args)
        return fn


class Oracle(_Lib):
    """The CPU restatement (oracle/oracle.cpp)."""

    prefix = "orc_"

    def __init__(self, path=ORACLE_SO):
        super().__init__(path)
        f = self._f
        self._weights = f("error_weights", None, _p)
        self._sweights = f("scaled_weights", None, _dbl, _p)
        self._scale = f("scale_char", C.c_uint, C.c_float)
        self._clamp = f("clamp_pixel", C.c_float, _dbl)
        self._level = f("curve_level", _i32, _i64, _i64)
        self._hp = f("hilbert_point", None, _i64, _i32, _p, _p)
        self._rwalk = f("recursive_walk", _i64, _i32, _p, _p)
        self._seq = f("seq_order", _i64, _i64, _i64, _p)
        self._blk = f("block_order", _i64, _i64, _i64, _p, _p)
        self._build = f("tree_build", _p, _p, _i64, _i64, _i32, _i32, _i32)
        self._free = f("tree_free", None, _p)
        self._count = f("tree_color_count", _i32, _p)
        self._nodes = f("tree_node_count", _i64, _p)
        self._depth = f("tree_depth", _i32, _p)
        self._assoc = f("tree_assoc", _i32, _p)
        self._cmap = f("tree_colormap", None, _p, _p)
        self._nodearr = f("tree_nodes", None, _p, _p, _p, _p, _p, _p, _p, _p, _p)
        self._lookup = f("tree_lookup", _i32, _p, _p)
        self._walk = f("walk_seq", None, _p, _p, _i64, _i64, _dbl, _i32)
        self._walk_idx = f("walk_seq_index", None, _p, _p, _i64, _i64, _dbl, _i32, _p)
        self._blocks = f("walk_blocks", _i32, _p, _p, _i64, _i64, _i32, _i32, _dbl, _p)
        self._blocks16 = f("walk_blocks_u16", _i32, _p, _p, _i64, _i64, _i32, _i32, _dbl, _p)
        self._yuv2f = f("yuv_to_rgbaf", _i32, _p, _i64, _i64, _i32, _i32, _p)
        self._blocksyuv = f("walk_blocks_yuv444", _i32, _p, _p, _i64, _i64, _i32, _i32, _dbl, _p)
        self._screen = f("dither_screen", _i32, _p, _i32, _p, _i64, _i64, _i32, _i32, _dbl)
        self._inv = f("inverse_filter", None, _i32, _dbl, _p)
        self._jac = f("jacobi", _i32, _p, _p, _i64, _i64, _i32, _i32, _i32, _i32, _dbl, _p)
        self._diffuse = f("dither_diffuse", _i32, _p, _i32, _p, _i64, _i64, _i32, _i32, _dbl)
        self._montage = f("video_montage", _i32, _p, _i32, _i32, _i64, _i64, _i32, _i32, _i32,
                          _p, _p, _p)
        self._jinit = f("jacobi_init", _i32, _p, _p, _i64, _i32, _i32, _p)
        self._jsweep = f("jacobi_sweep", _i32, _p, _p, _i64, _p, _p, _p, _i32, _i32, _i32, _dbl,
                         _i32)
        self._many = f("walk_seq_many", _i32, _p, _p, _p, _p, _i32, _dbl, _i32)
        self._misses = f("walk_seq_misses", _i64, _p, _p, _i64, _i64, _dbl)

    def walk_seq_misses(self, tree, img, diffusion=1.0):
        a = _rgba_f32(img).copy()
        h, w = a.shape[:2]
        return int(self._misses(tree.h, _ptr(a), w, h, diffusion))

    # -- numeric -------------------------------------------------------
    def error_weights(self):
        w = np.zeros(16, np.float64)
        self._weights(_ptr(w))
        return w

    def scaled_weights(self, diffusion=1.0):
        w = np.zeros(16, np.float64)
        self._sweights(diffusion, _ptr(w))
        return w

    def scale_char(self, q):
        return int(self._scale(float(q)))

    def clamp_pixel(self, v):
        return float(self._clamp(float(v)))

    # -- curve ---------------------------------------------------------
    def curve_level(self, w, h):
        return int(self._level(w, h))

    def hilbert_point(self, n, level):
        x, y = C.c_int(), C.c_int()
        self._hp(n, level, C.byref(x), C.byref(y))
        return x.value, y.value

    def recursive_walk(self, level):
        cnt = 1 << (2 * level) if level > 0 else 1
        xs = np.zeros(cnt + 1, np.int32)
        ys = np.zeros(cnt + 1, np.int32)
        k = self._rwalk(level, _ptr(xs), _ptr(ys))
        return xs[:k], ys[:k]

    def seq_order(self, w, h):
        out = np.zeros(max(w * h, 1), np.int32)
        k = self._seq(w, h, _ptr(out))
        return out[:k]

    def block_order(self, w, h):
        order = np.zeros(w * h + 1, np.int32)
        owner = np.zeros(max(w * h, 1), np.int32)
        k = self._blk(w, h, _ptr(order), _ptr(owner))
        return order[:k], owner[: w * h]

    # -- octree --------------------------------------------------------
    def tree_build(self, img, colors, assoc, gray=False):
        a = _rgba_f32(img)
        h, w = a.shape[:2]
        return OracleTree(self, self._build(_ptr(a), w, h, colors, int(assoc), int(gray)))

    # -- engines -------------------------------------------------------
    def walk_seq(self, tree, img, diffusion=1.0, use_cache=True, want_index=False):
        a = _rgba_f32(img).copy()
        h, w = a.shape[:2]
        if want_index:
            idx = np.zeros(w * h, np.int32)
            self._walk_idx(tree.h, _ptr(a), w, h, diffusion, int(use_cache), _ptr(idx))
            return a, idx
        self._walk(tree.h, _ptr(a), w, h, diffusion, int(use_cache))
        return a

    def walk_blocks(self, tree, batch, block, diffusion=1.0, want_index=False):
        a = np.ascontiguousarray(batch, dtype=np.float32).copy()
        if a.ndim == 3:
            a = a[None]
        frames, h, w = a.shape[:3]
        n = len(self.block_order(w, h)[0])
        idx = np.zeros(frames * n, np.uint8)
        rc = self._blocks(tree.h, _ptr(a), w, h, frames, block, diffusion, _ptr(idx))
        if rc != 0:
            raise ValueError("refused")
        return (a, idx.reshape(frames, n)) if want_index else a

    def walk_blocks_u16(self, tree, batch, block, diffusion=1.0):
        a = np.ascontiguousarray(batch, dtype=np.float32)
        if a.ndim == 3:
            a = a[None]
        frames, h, w = a.shape[:3]
        out = np.zeros((frames, h, w, 4), np.uint16)
        rc = self._blocks16(tree.h, _ptr(a), w, h, frames, block, diffusion, _ptr(out))
        if rc != 0:
            raise ValueError("refused")
        return out

    def yuv_to_rgbaf(self, planes, w, h, fmt):
        """Planar 8-bit frames (fmt 7 = 4:4:4, 8 = 4:2:0), as a flat uint8
        array of whole frames, widened to (frames, h, w, 4) float32."""
        a = np.ascontiguousarray(planes, dtype=np.uint8).reshape(-1)
        fb = 3 * w * h if fmt == 7 else w * h + 2 * (w // 2) * (h // 2)
        frames = a.size // fb
        out = np.zeros((frames, h, w, 4), np.float32)
        if self._yuv2f(_ptr(a), w, h, frames, fmt, _ptr(out)) != 0:
            raise ValueError("refused")
        return out

    def dither_screen(self, entries, frames, screen, diffusion=1.0):
        """A threshold-screen plugin dither (screen 0..3) of (frames, h, w, 4)
        float32 RgbaF with a (count, 4) float64 palette; returns the result."""
        e = np.ascontiguousarray(entries, dtype=np.float64)
        a = np.ascontiguousarray(frames, dtype=np.float32).copy()
        if a.ndim == 3:
            a = a[None]
        n, h, w = a.shape[:3]
        if self._screen(_ptr(e), len(e), _ptr(a), w, h, n, screen, diffusion) != 0:
            raise ValueError("refused")
        return a

    def dither_diffuse(self, entries, frames, kind, diffusion=1.0):
        """An error-diffusion plugin dither (kind 0 floyd-steinberg, 1 atkinson,
        2 jarvis) of (frames, h, w, 4) float32 RgbaF; returns the result."""
        e = np.ascontiguousarray(entries, dtype=np.float64)
        a = np.ascontiguousarray(frames, dtype=np.float32).copy()
        if a.ndim == 3:
            a = a[None]
        n, h, w = a.shape[:3]
        if self._diffuse(_ptr(e), len(e), _ptr(a), w, h, n, kind, diffusion) != 0:
            raise ValueError("refused")
        return a

    def walk_blocks_yuv444(self, tree, batch, block, diffusion=1.0):
        """The block walk with planar 4:4:4 output: (frames, 3, h, w) uint8."""
        a = np.ascontiguousarray(batch, dtype=np.float32)
        if a.ndim == 3:
            a = a[None]
        frames, h, w = a.shape[:3]
        out = np.zeros((frames, 3, h, w), np.uint8)
        if self._blocksyuv(tree.h, _ptr(a), w, h, frames, block, diffusion, _ptr(out)) != 0:
            raise ValueError("refused")
        return out

    def inverse_filter(self, taps, diffusion=1.0):
        g = np.zeros(taps + 1, np.float64)
        self._inv(taps, diffusion, _ptr(g))
        return g

    def jacobi(self, tree, img, iterations, taps=96, fp64=False, use_tree=True,
               diffusion=1.0, want_q=False):
        a = _rgba_f32(img).copy()
        h, w = a.shape[:2]
        n = len(self.block_order(w, h)[0])
        q = np.zeros(max(n, 1), np.uint8)
        rc = self._jac(tree.h, _ptr(a), w, h, iterations, taps, int(fp64),
                       int(use_tree), diffusion, _ptr(q))
        if rc != 0:
            raise ValueError("refused")
        return (a, q[:n]) if want_q else a

    def video_montage(self, frames, want=None, tile=0, cols=0):
        """The palette montage (RgbaF, mh x mw x 4) of decoder frames: uint8 or
        uint16 arrays of h x w x 3 or 4."""
        frames = [np.ascontiguousarray(f) for f in frames]
        h, w, ch = frames[0].shape
        fmt = (1 if ch == 3 else 2) if frames[0].dtype == np.uint8 else (3 if ch == 3 else 4)
        want = want or len(frames)
        ptrs = (C.c_void_p * len(frames))(*[f.ctypes.data for f in frames])
        mw, mh = C.c_int64(), C.c_int64()
        if self._montage(ptrs, len(frames), fmt, w, h, want, tile, cols, None, C.byref(mw),
                         C.byref(mh)) != 0:
            raise ValueError("refused")
        out = np.zeros((mh.value, mw.value, 4), np.float32)
        self._montage(ptrs, len(frames), fmt, w, h, want, tile, cols, _ptr(out), None, None)
        return out

    def jacobi_gather(self, img, fp64=False):
        """GatherKernelT (src/rd_approx_cuda.cu:110-121): the pixels in block
        curve order, n x 4, as float or double."""
        a = _rgba_f32(img)
        h, w = a.shape[:2]
        order = self.block_order(w, h)[0]
        return np.ascontiguousarray(a.reshape(-1, 4)[order].astype(np.float64 if fp64 else np.float32))

    def jacobi_init(self, tree, cx, use_tree=True):
        cx = np.ascontiguousarray(cx)
        q = np.zeros(max(len(cx), 1), np.uint8)
        if self._jinit(tree.h, _ptr(cx), len(cx), int(cx.dtype == np.float64), int(use_tree),
                       _ptr(q)) != 0:
            raise ValueError("refused")
        return q[:len(cx)]

    def jacobi_sweep(self, tree, cx, q_in, taps=96, use_tree=True, diffusion=1.0, want_t=False,
                     mutant=0):
        """One sweep from q_in; returns q_out (and the unclamped targets, n x 4
        doubles, with want_t)."""
        cx = np.ascontiguousarray(cx)
        n = len(cx)
        q_in = np.ascontiguousarray(q_in, np.uint8)
        assert len(q_in) == n
        q = np.zeros(max(n, 1), np.uint8)
        t = np.zeros((max(n, 1), 4), np.float64) if want_t else None
        rc = self._jsweep(tree.h, _ptr(cx), n, _ptr(q_in), _ptr(q),
                          _ptr(t) if want_t else None, taps, int(cx.dtype == np.float64),
                          int(use_tree), diffusion, mutant)
        if rc != 0:
            raise ValueError("refused")
        return (q[:n], t[:n]) if want_t else q[:n]

    def walk_seq_many(self, trees, images, threads=1, diffusion=1.0):
        """Dithers every image in place (CPU baseline)."""
        n = len(images)
        tp = (C.c_void_p * n)(*[t.h for t in trees])
        ip = (C.c_void_p * n)(*[im.ctypes.data for im in images])
        ws = np.array([im.shape[1] for im in images], np.int64)
        hs = np.array([im.shape[0] for im in images], np.int64)
        self._many(tp, ip, _ptr(ws), _ptr(hs), n, diffusion, threads)


class OracleTree:
    def __init__(self, lib, handle):
        self.lib = lib
        self.h = handle

    def __del__(self):
        try:
            if self.h:
                self.lib._free(self.h)
                self.h = None
        except Exception:
            pass

    @property
    def count(self):
        return int(self.lib._count(self.h))

    @property
    def node_count(self):
        return int(self.lib._nodes(self.h))

    @property
    def depth(self):
        return int(self.lib._depth(self.h))

    @property
    def assoc(self):
        return bool(self.lib._assoc(self.h))

    def colormap(self):
        out = np.zeros((self.count, 4), np.float64)
        self.lib._cmap(self.h, _ptr(out))
        return out

    def data(self):
        n = self.node_count
        child = np.zeros((n, 16), np.int32)
        parent = np.zeros(n, np.int32)
        cn = np.zeros(n, np.uint32)
        nu = np.zeros(n, np.uint64)
        tot = np.zeros((n, 4), np.float64)
        qe = np.zeros(n, np.float64)
        ids = np.zeros(n, np.uint32)
        lev = np.zeros(n, np.uint32)
        self.lib._nodearr(self.h, _ptr(child), _ptr(parent), _ptr(cn), _ptr(nu),
                          _ptr(tot), _ptr(qe), _ptr(ids), _ptr(lev))
        return TreeData(self.colormap(), child, parent, cn, nu, tot, qe, ids, lev,
                        self.depth, self.assoc)

    def lookup(self, target):
        t = np.ascontiguousarray(target, np.float64)
        return int(self.lib._lookup(self.h, _ptr(t)))


class Ref(_Lib):
    """The reference's own CPU sources (oracle/_ref/libref.so)."""

    prefix = "ref_"

    @staticmethod
    def available():
        return os.path.exists(REF_SO)

    def __init__(self, path=REF_SO):
        super().__init__(path)
        f = self._f
        self._build = f("tree_build", _p, _p, _i64, _i64, _i32, _i32, _i32)
        self._free = f("tree_free", None, _p)
        self._count = f("tree_color_count", _i32, _p)
        self._nodes = f("tree_node_count", _i64, _p)
        self._depth = f("tree_depth", _i32, _p)
        self._cmap = f("tree_colormap", None, _p, _p)
        self._nodearr = f("tree_nodes", None, _p, _p, _p, _p, _p, _p, _p, _p, _p)
        self._walk = f("walk_cpu", None, _p, _p, _i64, _i64, _dbl, _i32)
        self._blocks = f("blocks_cpu", _i32, _p, _p, _i64, _i64, _i32, _i32, _dbl)
        self._blocks16 = f("blocks_cpu_u16", _i32, _p, _p, _i64, _i64, _i32, _i32, _dbl, _p)
        self._dither = f("dither", _i32, C.c_char_p, _p, _i32, _p, _i64, _i64, _i32, _dbl)
        self._fused = f("blocks_cpu_fused", _i32, _p, _p, _i64, _i64, _i32, _i32, _dbl, _p, _i32,
                        _p, _i64, _i32, _i32)
        self._bci = f("build_curve_index", _i64, _i32, _i64, _i64, _p, _p)
        self._hp = f("hilbert_point", None, _i64, _i32, _p, _p)
        self._mism = f("curve_mismatches", _i64, _i32, _i32)
        self._level = f("curve_level", _i32, _i64, _i64)
        self._weights = f("error_weights", None, _p)
        self._scale = f("scale_char", C.c_uint, C.c_float)
        self._many = f("walk_cpu_many", _i32, _p, _p, _p, _p, _i32, _dbl, _i32)

    def error_weights(self):
        w = np.zeros(16, np.float64)
        self._weights(_ptr(w))
        return w

    def scale_char(self, q):
        return int(self._scale(float(q)))

    def curve_level(self, w, h):
        return int(self._level(w, h))

    def curve_mismatches(self, level, transform=0):
        return int(self._mism(level, transform))

    def hilbert_point(self, n, level):
        x, y = C.c_int(), C.c_int()
        self._hp(n, level, C.byref(x), C.byref(y))
        return x.value, y.value

    def build_curve_index(self, w, h):
        level = self.curve_level(w, h)
        out = np.zeros(w * h + 1, np.int32)
        owner = np.zeros(max(w * h, 1), np.int32)
        k = self._bci(level, w, h, _ptr(out), _ptr(owner))
        return out[:k], owner[: w * h]

    def tree_build(self, img, colors, assoc, gray=False):
        a = _rgba_f32(img)
        h, w = a.shape[:2]
        return RefTree(self, self._build(_ptr(a), w, h, colors, int(assoc), int(gray)))

    def walk_seq(self, tree, img, diffusion=1.0, use_cache=True):
        a = _rgba_f32(img).copy()
        h, w = a.shape[:2]
        self._walk(tree.h, _ptr(a), w, h, diffusion, int(use_cache))
        return a

    def walk_blocks(self, tree, batch, block, diffusion=1.0):
        a = np.ascontiguousarray(batch, dtype=np.float32).copy()
        if a.ndim == 3:
            a = a[None]
        frames, h, w = a.shape[:3]
        if self._blocks(tree.h, _ptr(a), w, h, frames, block, diffusion) != 0:
            raise ValueError("refused")
        return a

    def walk_blocks_u16(self, tree, batch, block, diffusion=1.0):
        a = np.ascontiguousarray(batch, dtype=np.float32).copy()
        if a.ndim == 3:
            a = a[None]
        frames, h, w = a.shape[:3]
        out = np.zeros((frames, h, w, 4), np.uint16)
        if self._blocks16(tree.h, _ptr(a), w, h, frames, block, diffusion, _ptr(out)) != 0:
            raise ValueError("refused")
        return out

    def video_montage(self, frames, want=None, tile=0, cols=0):
        """The palette montage (RgbaF, mh x mw x 4) of decoder frames: uint8 or
        uint16 arrays of h x w x 3 or 4."""
        frames = [np.ascontiguousarray(f) for f in frames]
        h, w, ch = frames[0].shape
        fmt = (1 if ch == 3 else 2) if frames[0].dtype == np.uint8 else (3 if ch == 3 else 4)
        want = want or len(frames)
        ptrs = (C.c_void_p * len(frames))(*[f.ctypes.data for f in frames])
        mw, mh = C.c_int64(), C.c_int64()
        if self._montage(ptrs, len(frames), fmt, w, h, want, tile, cols, None, C.byref(mw),
                         C.byref(mh)) != 0:
            raise ValueError("refused")
        out = np.zeros((mh.value, mw.value, 4), np.float32)
        self._montage(ptrs, len(frames), fmt, w, h, want, tile, cols, _ptr(out), None, None)
        return out

    def jacobi_gather(self, img, fp64=False):
        """GatherKernelT (src/rd_approx_cuda.cu:110-121): the pixels in block
        curve order, n x 4, as float or double."""
        a = _rgba_f32(img)
        h, w = a.shape[:2]
        order = self.block_order(w, h)[0]
        return np.ascontiguousarray(a.reshape(-1, 4)[order].astype(np.float64 if fp64 else np.float32))

    def jacobi_init(self, tree, cx, use_tree=True):
        cx = np.ascontiguousarray(cx)
        q = np.zeros(max(len(cx), 1), np.uint8)
        if self._jinit(tree.h, _ptr(cx), len(cx), int(cx.dtype == np.float64), int(use_tree),
                       _ptr(q)) != 0:
            raise ValueError("refused")
        return q[:len(cx)]

    def jacobi_sweep(self, tree, cx, q_in, taps=96, use_tree=True, diffusion=1.0, want_t=False,
                     mutant=0):
        """One sweep from q_in; returns q_out (and the unclamped targets, n x 4
        doubles, with want_t)."""
        cx = np.ascontiguousarray(cx)
        n = len(cx)
        q_in = np.ascontiguousarray(q_in, np.uint8)
        assert len(q_in) == n
        q = np.zeros(max(n, 1), np.uint8)
        t = np.zeros((max(n, 1), 4), np.float64) if want_t else None
        rc = self._jsweep(tree.h, _ptr(cx), n, _ptr(q_in), _ptr(q),
                          _ptr(t) if want_t else None, taps, int(cx.dtype == np.float64),
                          int(use_tree), diffusion, mutant)
        if rc != 0:
            raise ValueError("refused")
        return (q[:n], t[:n]) if want_t else q[:n]

    def walk_seq_many(self, trees, images, threads=1, diffusion=1.0):
        n = len(images)
        tp = (C.c_void_p * n)(*[t.h for t in trees])
        ip = (C.c_void_p * n)(*[im.ctypes.data for im in images])
        ws = np.array([im.shape[1] for im in images], np.int64)
        hs = np.array([im.shape[0] for im in images], np.int64)
        self._many(tp, ip, _ptr(ws), _ptr(hs), n, diffusion, threads)


    def dither(self, name, entries, frames, diffusion=1.0):
        """The reference's registered plugin `name` run on (frames, h, w, 4)
        float32 RgbaF with a (count, 4) float64 palette; returns the result."""
        e = np.ascontiguousarray(entries, dtype=np.float64)
        a = np.ascontiguousarray(frames, dtype=np.float32).copy()
        if a.ndim == 3:
            a = a[None]
        n, h, w = a.shape[:3]
        rc = self._dither(name.encode(), _ptr(e), len(e), _ptr(a), w, h, n, diffusion)
        if rc != 0:
            raise ValueError("refused" if rc == -1 else "no such plugin")
        return a

    def walk_blocks_fused(self, tree, block, w, h, frames, batch=None, yuv_in=None,
                          yuv_out=True, diffusion=1.0):
        """RiemersmaBlocksCpu's fused path: reads `batch` (frames, h, w, 4)
        float32 or planar 4:4:4 `yuv_in` (frames * 3 * h * w bytes); returns
        planar 4:4:4 (frames, 3, h, w) uint8 or rgba64le (frames, h, w, 4)."""
        a = None
        if batch is not None:
            a = np.ascontiguousarray(batch, dtype=np.float32).copy()
        g = None
        if yuv_in is not None:
            g = np.ascontiguousarray(yuv_in, dtype=np.uint8).reshape(-1)
        if yuv_out:
            out = np.zeros((frames, 3, h, w), np.uint8)
        else:
            out = np.zeros((frames, h, w, 4), np.uint16)
        rc = self._fused(tree.h, _ptr(a) if a is not None else None, w, h, frames, block,
                         diffusion, _ptr(out), int(yuv_out), _ptr(g) if g is not None else None,
                         3 * w * h, 3 if g is not None else 0, 0)
        if rc != 0:
            raise ValueError("refused")
        return out


class RefTree(OracleTree):
    @property
    def assoc(self):
        raise NotImplementedError

    def data(self, assoc=False):
        n = self.node_count
        child = np.zeros((n, 16), np.int32)
        parent = np.zeros(n, np.int32)
        cn = np.zeros(n, np.uint32)
        nu = np.zeros(n, np.uint64)
        tot = np.zeros((n, 4), np.float64)
        qe = np.zeros(n, np.float64)
        ids = np.zeros(n, np.uint32)
        lev = np.zeros(n, np.uint32)
        self.lib._nodearr(self.h, _ptr(child), _ptr(parent), _ptr(cn), _ptr(nu),
                          _ptr(tot), _ptr(qe), _ptr(ids), _ptr(lev))
        return TreeData(self.colormap(), child, parent, cn, nu, tot, qe, ids, lev,
                        self.depth, assoc)


_oracle = None


def oracle():
    global _oracle
    if _oracle is None:
        _oracle = Oracle()
    return _oracle
