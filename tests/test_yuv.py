"""The oracle's planar YUV restatement, pinned on the CPU.

* 4:4:4 in (swscale's matrix, include/rd_types.h:75-97) and 4:4:4 out (the
  fused scatter's BT.601 matrix, src/rd_riemersma_cpu.cpp:718-730), and rgba64
  out of 4:4:4 in: equal to the reference CPU engine's fused gather/scatter
  (oracle/_ref RiemersmaBlocksCpu), unvisited pixels included.
* The matrix itself: the grey axis stays grey, video black (16) is 0, and
  grey rises strictly with Y at 65535/219 per step within rounding.
* 4:2:0 in, which only the reference's CUDA reader has: its chroma at luma
  (x, y) is the sample at (x/2, y/2) for even x and y, and otherwise the mean,
  rounded half up, of the 2 or 4 samples that (x, y) lies between, with the
  last row and column repeated; through the 4:4:4 path that makes it equal to
  4:4:4 input with the chroma upsampled that way.
"""
import numpy as np
import pytest

from oracle.oracle import Oracle, Ref

YUV444, YUV420 = 7, 8


@pytest.fixture(scope="module")
def orc():
    return Oracle()


@pytest.fixture(scope="module")
def ref():
    if not Ref.available():
        pytest.skip("oracle/_ref not built (needs the reference sources)")
    return Ref()


def _planes(rng, w, h):
    return rng.integers(0, 256, 3 * w * h * 2, dtype=np.uint8)


@pytest.mark.parametrize("w,h", [(17, 13), (32, 16), (64, 64), (40, 31), (1, 1), (2, 5)])
@pytest.mark.parametrize("colors", [2, 16, 200])
def test_yuv444_paths_match_reference(orc, ref, w, h, colors):
    rng = np.random.default_rng(w * 1000 + h * 10 + colors)
    yuv = _planes(rng, w, h)
    batch = orc.yuv_to_rgbaf(yuv, w, h, YUV444)
    t = orc.tree_build(batch[0], colors, False)
    rt = ref.tree_build(batch[0], colors, False)
    for block in (16, 64):
        got = orc.walk_blocks_yuv444(t, batch, block)
        want = ref.walk_blocks_fused(rt, block, w, h, 2, yuv_in=yuv, yuv_out=True)
        assert np.array_equal(got, want), ("yuv444 -> yuv444", block)
        got = orc.walk_blocks_u16(t, batch, block)
        want = ref.walk_blocks_fused(rt, block, w, h, 2, yuv_in=yuv, yuv_out=False)
        assert np.array_equal(got, want), ("yuv444 -> rgba64", block)
        f = batch.copy()
        f[..., :3] += rng.normal(0, 400, f[..., :3].shape).astype(np.float32)
        got = orc.walk_blocks_yuv444(t, f, block)
        want = ref.walk_blocks_fused(rt, block, w, h, 2, batch=f, yuv_out=True)
        assert np.array_equal(got, want), ("float -> yuv444", block)


def test_sws_matrix_grey_axis(orc):
    ys = np.arange(256, dtype=np.uint8)
    planes = np.concatenate([ys, np.full(256, 128, np.uint8), np.full(256, 128, np.uint8)])
    px = orc.yuv_to_rgbaf(planes, 256, 1, YUV444)[0, 0]
    assert np.all(px[:, 0] == px[:, 1]) and np.all(px[:, 1] == px[:, 2])
    assert np.all(px[:, 3] == 65535.0)
    grey = px[:, 0].astype(np.int64)
    assert np.all(grey[:17] == 0)  # at and below video black
    assert np.all(np.diff(grey[16:236]) > 0)  # strictly rising over video range
    assert np.all(np.diff(grey) >= 0) and grey[255] == 65535  # then clipped at white
    ideal = (np.arange(16, 236) - 16) * 65535.0 / 219.0
    assert np.max(np.abs(grey[16:236] - ideal)) < 0.01 * 65535


def _upsample(plane, w2, h2):
    """Chroma at every luma (x, y): the sample at (x/2, y/2) for even x and y,
    else the rounded-half-up mean of the 2 or 4 samples around it (the last
    row and column repeated)."""
    c = plane.reshape(h2, w2).astype(np.int64)
    right = np.concatenate([c[:, 1:], c[:, -1:]], axis=1)
    down = np.concatenate([c[1:], c[-1:]], axis=0)
    diag = np.concatenate([right[1:], right[-1:]], axis=0)
    out = np.zeros((2 * h2, 2 * w2), np.int64)
    out[0::2, 0::2] = c
    out[0::2, 1::2] = (c + right + 1) // 2
    out[1::2, 0::2] = (c + down + 1) // 2
    out[1::2, 1::2] = (c + right + down + diag + 2) // 4
    return out.astype(np.uint8).reshape(-1)


@pytest.mark.parametrize("w,h", [(2, 2), (4, 2), (16, 10), (34, 18), (64, 64)])
def test_yuv420_is_444_with_bilinear_chroma(orc, w, h):
    rng = np.random.default_rng(w + 100 * h)
    w2, h2 = w // 2, h // 2
    y = rng.integers(0, 256, w * h, dtype=np.uint8)
    u = rng.integers(0, 256, w2 * h2, dtype=np.uint8)
    v = rng.integers(0, 256, w2 * h2, dtype=np.uint8)
    got = orc.yuv_to_rgbaf(np.concatenate([y, u, v]), w, h, YUV420)
    want = orc.yuv_to_rgbaf(np.concatenate([y, _upsample(u, w2, h2), _upsample(v, w2, h2)]), w, h,
                            YUV444)
    assert np.array_equal(got, want)
    flat = np.concatenate([y, np.full(w2 * h2, 77, np.uint8), np.full(w2 * h2, 201, np.uint8)])
    const = np.concatenate([y, np.full(w * h, 77, np.uint8), np.full(w * h, 201, np.uint8)])
    assert np.array_equal(orc.yuv_to_rgbaf(flat, w, h, YUV420), orc.yuv_to_rgbaf(const, w, h, YUV444))


def test_yuv420_refuses_odd_sides(orc):
    with pytest.raises(ValueError):
        orc.yuv_to_rgbaf(np.zeros(3 * 15, np.uint8), 5, 3, YUV420)
