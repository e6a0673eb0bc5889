"""Video palette sampling and palette text files, on the CPU.

* The oracle's montage (VideoBuildPalette's montage path, src/rd_video.cpp:
  2373-2470, :2636-2660) against what the layout fixes, computed here with
  numpy indexing: whole-frame cells are the frames widened x257 (alpha 65535
  without an alpha channel, 16-bit samples as they are), in sample order,
  `cols` = ceil(sqrt(want)) wide; tile cells sample source (tx*W/tile,
  ty*H/tile); every cell past the last sample repeats it.  A transposed cell
  grid, a row-major/column-major swap of the nearest-neighbour index or a
  repeat of the first sample instead of the last each fail one of these.
* The palette .txt form (src/rd_im.cpp:994-1024, :1175-1330) written and read
  by the library's host code (sa_palette_write_txt / sa_palette_read_txt, no
  device needed): the exact bytes of the reference's printf format, round
  trips, the legacy rows, the narrowing and canonical order, and every
  refusal.  The reference's file functions live in an ImageMagick translation
  unit that cannot be built here, so no file written by the reference itself
  is available; the format is fixed by the reference's format strings.
"""
import math
import os

import numpy as np
import pytest

import stipple_amd as sa


def _widen(f):
    f = np.asarray(f)
    q = f.astype(np.float32) * (np.float32(257.0) if f.dtype == np.uint8 else np.float32(1.0))
    if f.shape[-1] == 3:
        q = np.concatenate([q, np.full(f.shape[:2] + (1,), 65535.0, np.float32)], -1)
    return q


def _frames(rng, n, w, h, ch, dtype):
    top = 256 if dtype == np.uint8 else 65536
    return [rng.integers(0, top, (h, w, ch)).astype(dtype) for _ in range(n)]


@pytest.mark.parametrize("n,want,ch,dtype", [(1, 1, 3, np.uint8), (3, 5, 4, np.uint8),
                                             (4, 4, 3, np.uint16), (2, 7, 4, np.uint16),
                                             (9, 9, 3, np.uint8)])
def test_montage_whole_frames(oracle, n, want, ch, dtype):
    rng = np.random.default_rng(n * 10 + want)
    w, h = 13, 7
    fr = _frames(rng, n, w, h, ch, dtype)
    m = oracle.video_montage(fr, want=want)
    cols = math.ceil(math.sqrt(want))
    rows = (want + cols - 1) // cols
    assert m.shape == (rows * h, cols * w, 4)
    for cell in range(cols * rows):
        s = min(cell, n - 1)
        cy, cx = divmod(cell, cols)
        got = m[cy * h:(cy + 1) * h, cx * w:(cx + 1) * w]
        assert np.array_equal(got, _widen(fr[s])), cell


@pytest.mark.parametrize("w,h,tile", [(40, 30, 16), (30, 40, 16), (8, 6, 13), (64, 64, 64),
                                      (37, 11, 5)])
def test_montage_tiles_nearest_neighbour(oracle, w, h, tile):
    rng = np.random.default_rng(w + h + tile)
    fr = _frames(rng, 3, w, h, 4, np.uint8)
    m = oracle.video_montage(fr, want=3, tile=tile)
    cols = 2
    assert m.shape == (2 * tile, cols * tile, 4)
    sy = (np.arange(tile, dtype=np.int64) * h) // tile
    sx = (np.arange(tile, dtype=np.int64) * w) // tile
    for cell in range(4):
        s = min(cell, 2)
        cy, cx = divmod(cell, cols)
        exp = _widen(fr[s])[sy[:, None], sx[None, :]]
        assert np.array_equal(m[cy * tile:(cy + 1) * tile, cx * tile:(cx + 1) * tile], exp)


def test_montage_columns_given(oracle):
    """cols = want (the reference's one-row montage of --im-palette)."""
    rng = np.random.default_rng(3)
    fr = _frames(rng, 4, 5, 3, 3, np.uint8)
    m = oracle.video_montage(fr, want=4, cols=4)
    assert m.shape == (3, 20, 4)
    assert np.array_equal(m[:, 15:20], _widen(fr[3]))


# ---- palette .txt -----------------------------------------------------------------------
def _ref_text(ent):
    """The reference's rows (src/rd_im.cpp:1008-1018 format strings)."""
    out = f"# rdither palette, {len(ent)} colours, Q16 (0..65535) and 8-bit hex\n"
    out += "# count\tr\tg\tb\ta\thex\n"
    for i, (r, g, b, a) in enumerate(ent):
        out += "%d\t%.6f\t%.6f\t%.6f\t%.6f\t%02X%02X%02X\n" % (
            i, r, g, b, a, int(r / 257.0 + 0.5), int(g / 257.0 + 0.5), int(b / 257.0 + 0.5))
    return out


def _random_palette(rng, n, alpha=False):
    ent = np.zeros((n, 4))
    ent[:, :3] = rng.uniform(0, 65535, (n, 3))
    ent[:, 3] = rng.uniform(0, 65535, n) if alpha else 65535.0
    return ent


def test_palette_txt_bytes(tmp_path):
    rng = np.random.default_rng(1)
    ent = _random_palette(rng, 37, alpha=True)
    ent[0] = [0.0, 65535.0, 128.5, 0.0]
    ent[1] = [7844.4814, 6801.6, 21875.5, 65535.0]
    p = tmp_path / "pal.txt"
    sa.write_palette_txt(ent, str(p))
    assert p.read_text() == _ref_text(ent)


def test_palette_txt_round_trip(tmp_path):
    """Values a float holds survive exactly; the entries come back in canonical
    (r, g, b, a) order; alpha comes back as written."""
    rng = np.random.default_rng(2)
    ent = _random_palette(rng, 64, alpha=True)
    ent[:, :3] = np.round(ent[:, :3] * 64) / 64          # exactly representable, 6 decimals
    ent[:, 3] = np.round(ent[:, 3])
    p = tmp_path / "pal.text"
    sa.write_palette_txt(ent, str(p))
    got, assoc = sa.read_palette_txt(str(p))
    order = np.lexsort((ent[:, 3], ent[:, 2], ent[:, 1], ent[:, 0]))
    assert np.array_equal(got, ent[order])
    assert assoc


def test_palette_txt_read_narrows_and_sorts(tmp_path):
    rng = np.random.default_rng(3)
    ent = _random_palette(rng, 50)
    p = tmp_path / "pal.txt"
    sa.write_palette_txt(ent, str(p))
    got, assoc = sa.read_palette_txt(str(p), expect_colors=50)
    parsed = np.array([[float("%.6f" % v) for v in row] for row in ent])
    parsed[:, :3] = parsed[:, :3].astype(np.float32).astype(np.float64)
    order = np.lexsort((parsed[:, 3], parsed[:, 2], parsed[:, 1], parsed[:, 0]))
    assert np.array_equal(got, parsed[order])
    assert not assoc


def test_palette_txt_legacy_rows(tmp_path):
    """Four- and five-token rows (no alpha column) read as opaque; comments and
    blank lines are skipped; rows may come in any order."""
    p = tmp_path / "old.txt"
    p.write_text("# old\n\n2\t300.5\t10\t20\tFFFFFF\n0 1000 2000 3000\r\n1 5.25 6 7\n")
    got, assoc = sa.read_palette_txt(str(p))
    assert np.array_equal(got, [[5.25, 6, 7, 65535], [300.5, 10, 20, 65535],
                                [1000, 2000, 3000, 65535]])
    assert not assoc


@pytest.mark.parametrize("text,code,needle", [
    ("0 1 2 3 4 5\n2 1 2 3 4 5\n", sa.SA_EINVAL, "no row declares index 1"),
    ("# nothing\n", sa.SA_EINVAL, "contains no palette rows"),
    ("0 1 2\n", sa.SA_EINVAL, "is not a palette row"),
    ("0 1 2 3 70000 AABBCC\n", sa.SA_EINVAL, "alpha '70000'"),
    ("0 1 2 3 x9 AABBCC\n", sa.SA_EINVAL, "alpha 'x9'"),
    ("65536 1 2 3\n", sa.SA_EINVAL, "index 65536 is out of range"),
    ("-1 1 2 3\n", sa.SA_EINVAL, "index -1 is out of range"),
])
def test_palette_txt_refusals(tmp_path, text, code, needle):
    p = tmp_path / "bad.txt"
    p.write_text(text)
    with pytest.raises(sa.EngineError) as ei:
        sa.read_palette_txt(str(p))
    assert ei.value.code == code and needle in str(ei.value)


def test_palette_txt_count_and_suffix(tmp_path):
    p = tmp_path / "pal.txt"
    sa.write_palette_txt(_random_palette(np.random.default_rng(4), 5), str(p))
    with pytest.raises(sa.EngineError) as ei:
        sa.read_palette_txt(str(p), expect_colors=6)
    assert ei.value.code == sa.SA_EREFUSED and "holds 5 colours but --colors is 6" in str(ei.value)
    with pytest.raises(sa.EngineError) as ei:
        sa.write_palette_txt(np.zeros((2, 4)), str(tmp_path / "pal.png"))
    assert ei.value.code == sa.SA_EUNSUPPORTED
    with pytest.raises(sa.EngineError) as ei:
        sa.write_palette_txt(np.zeros((0, 4)), str(tmp_path / "empty.txt"))
    assert "palette is empty" in str(ei.value)
    assert not os.path.exists(tmp_path / "empty.txt")
