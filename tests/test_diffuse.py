"""The oracle's error-diffusion dithers pinned against the reference's own
plugins (oracle/_ref: src/rd_dithers.cc run through FindDither):
floyd-steinberg, atkinson and jarvis, on seeded frames and palettes --
degenerate widths and heights (1, 2, 3; jarvis's column taps wrap across rows
at x = 1 and x = w-2), a one-colour palette, duplicate entries, diffusion 0,
out-of-range and NaN pixels."""
import numpy as np
import pytest

from oracle.oracle import Oracle, Ref

NAMES = ["floyd-steinberg", "atkinson", "jarvis"]


@pytest.fixture(scope="module")
def orc():
    return Oracle()


@pytest.fixture(scope="module")
def ref():
    if not Ref.available():
        pytest.skip("oracle/_ref not built (needs the reference sources)")
    return Ref()


def case(seed, count, w, h, frames=2, nan=False):
    rng = np.random.default_rng(seed)
    ent = np.zeros((count, 4))
    ent[:, :3] = rng.uniform(0, 65535, (count, 3))
    ent[:, 3] = 65535.0
    if count > 3:
        ent[2] = ent[1]  # a duplicate entry
    px = rng.uniform(-500, 66000, (frames, h, w, 4)).astype(np.float32)
    px[..., 3] = rng.uniform(0, 65535, (frames, h, w)).astype(np.float32)
    if nan:
        px[0, h // 2, w // 2, 1] = np.nan
    return ent, px


SIZES = [(1, 1), (1, 7), (7, 1), (2, 2), (3, 2), (2, 5), (5, 3), (9, 7), (33, 17), (70, 65),
         (130, 3)]


@pytest.mark.parametrize("kind", range(3))
@pytest.mark.parametrize("w,h", SIZES)
def test_diffuse_matches_reference(orc, ref, kind, w, h):
    for count, diffusion in ((1, 1.0), (16, 1.0), (5, 0.75), (200, 0.0)):
        ent, px = case(kind * 1000 + w * 31 + h + count, count, w, h, nan=(count == 5))
        got = orc.dither_diffuse(ent, px, kind, diffusion)
        want = ref.dither(NAMES[kind], ent, px, diffusion)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), (count, diffusion)


def test_diffuse_refusals(orc, ref):
    ent, px = case(1, 4, 5, 5)
    for kind in range(3):
        with pytest.raises(ValueError):
            orc.dither_diffuse(ent[:0], px, kind)
        with pytest.raises(ValueError):
            ref.dither(NAMES[kind], ent[:0], px)
