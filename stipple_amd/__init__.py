"""stipple_amd -- MI355X-native Riemersma hot path (host mirror of the C-ABI).

The compute runs in hand-written HIP kernels inside libstipple_amd.so, reached
through the C-ABI declared in include/stipple_amd.h.  This module binds that
ABI with ctypes and mirrors the reference's engine interface
(include/rd_riemersma.h, include/rd_octree.h of the reference) with the same
names, argument meaning and error behaviour: an engine that cannot run raises
EngineError with the message the library returned; nothing here computes a
result on the host.  Device memory and streams are torch's (plumbing only).
"""
from __future__ import annotations

import ctypes as C
import os

from ._lib import (  # noqa: F401
    PIX_INDEX16,
    PIX_INDEX8,
    PIX_RGB16,
    PIX_RGB8,
    PIX_RGBA16,
    PIX_RGBA8,
    PIX_RGBAF32,
    PIX_YUV420,
    PIX_YUV444,
    SA_EDEVICE,
    SA_EINVAL,
    SA_ENOMEM,
    SA_EREFUSED,
    SA_EUNSUPPORTED,
    SA_OK,
    EngineError,
    FrameChain,
    Palette,
    SeqOptions,
    WalkJob,
    available,
    curve_order,
    dither_screen,
    dither_diffuse,
    DIFFUSERS,
    device_name,
    image_spec,
    init,
    inverse_filter,
    lib,
    seq_stats,
    timing_collect,
    timing_enable,
    tune_get,
    tune_set,
    library_path,
    walk_approx,
    walk_blocks,
    walk_blocks_noalloc,
    walk_seq_batch,
    video_montage,
    write_palette_txt,
    read_palette_txt,
)

__all__ = [
    "EngineError",
    "FrameChain",
    "Palette",
    "SeqOptions",
    "available",
    "curve_order",
    "device_name",
    "image_spec",
    "init",
    "lib",
    "seq_stats",
    "timing_collect",
    "timing_enable",
    "library_path",
    "walk_approx",
    "walk_blocks",
    "walk_blocks_noalloc",
    "walk_seq_batch",
    "dither_diffuse",
    "video_montage",
    "write_palette_txt",
    "read_palette_txt",
]
