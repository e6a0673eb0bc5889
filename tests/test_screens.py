"""The oracle's threshold-screen dithers pinned against the reference's own
plugins (oracle/_ref: src/rd_dithers_ordered.cc and src/rd_dithers.cc run
through FindDither): bayer-ordered, void-and-cluster, void-and-cluster-fast
and clustered-dot, on seeded frames and palettes, including a one-colour
palette, duplicate entries and fractional Q16 entries."""
import numpy as np
import pytest

from oracle.oracle import Oracle, Ref

NAMES = ["bayer-ordered", "void-and-cluster", "void-and-cluster-fast", "clustered-dot"]


@pytest.fixture(scope="module")
def orc():
    return Oracle()


@pytest.fixture(scope="module")
def ref():
    if not Ref.available():
        pytest.skip("oracle/_ref not built (needs the reference sources)")
    return Ref()


def _case(seed, count, w, h, frames=2):
    rng = np.random.default_rng(seed)
    ent = np.zeros((count, 4))
    ent[:, :3] = rng.uniform(0, 65535, (count, 3))
    ent[:, 3] = 65535.0
    if count > 3:
        ent[2] = ent[1]  # a duplicate entry
    px = rng.uniform(-500, 66000, (frames, h, w, 4)).astype(np.float32)
    px[..., 3] = rng.uniform(0, 65535, (frames, h, w)).astype(np.float32)
    return ent, px


@pytest.mark.parametrize("screen", range(4))
@pytest.mark.parametrize("count,w,h,diffusion", [
    (1, 9, 7, 1.0), (2, 33, 17, 1.0), (16, 70, 65, 0.75), (200, 40, 30, 1.0), (7, 130, 3, 0.0)])
def test_screen_matches_reference(orc, ref, screen, count, w, h, diffusion):
    ent, px = _case(screen * 100 + count, count, w, h)
    got = orc.dither_screen(ent, px, screen, diffusion)
    want = ref.dither(NAMES[screen], ent, px, diffusion)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_screen_refusals(orc, ref):
    ent, px = _case(1, 4, 5, 5)
    for screen in range(4):
        with pytest.raises(ValueError):
            orc.dither_screen(ent[:0], px, screen)
        with pytest.raises(ValueError):
            ref.dither(NAMES[screen], ent[:0], px)
