"""Tuning knobs and the autotuner's cache (no device needed: setting a knob
only records a launch shape)."""
import pytest

import stipple_amd as sa
from stipple_amd import tune


def test_knobs_validate_and_round_trip():
    for knob, values in tune.KNOBS.items():
        before = sa.tune_get(knob)
        for v in values:
            sa.tune_set(knob, v)
            assert sa.tune_get(knob) == v
        sa.tune_set(knob, before)
    with pytest.raises(sa.EngineError):
        sa.tune_set("blocks.threads", 100)
    with pytest.raises(sa.EngineError):
        sa.tune_set("approx.per", 3)
    with pytest.raises(sa.EngineError):
        sa.tune_set("seq.memo_bits", 11)
    with pytest.raises(sa.EngineError):
        sa.tune_set("no.such.knob", 1)


def test_cache_round_trip(tmp_path):
    p = str(tmp_path / "tuned.json")
    sig = tune.signature("blocks", 1920, 1080, 16, 16, block=64)
    c = tune.TuneCache(p)
    c.put(sig, "blocks.threads", 128)
    c2 = tune.TuneCache(p)
    assert c2.get(sig) == {"blocks.threads": 128}
    before = sa.tune_get("blocks.threads")
    assert c2.apply(sig) == {"blocks.threads": 128}
    assert sa.tune_get("blocks.threads") == 128
    sa.tune_set("blocks.threads", before)
