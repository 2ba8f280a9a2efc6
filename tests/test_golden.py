"""The C1 golden digests (tests/golden/c1_digests.json, written by the
oracle's run in tests/golden/c1_golden.py, sprouting oracle included) on the
CPU: the oracle's day phase re-run for all 1,000 ticks, with the three
Nights' sprouting done by the engine's host pass (axc_sprout_pass). Every
tick's raster digest and spike count and every state digest must match, so
the oracle, the generator and the host sprouting at C1 scale are all held to
the committed fixture."""
import os
import sys

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))

import c1_golden as G  # noqa: E402


def _first_mismatch(got, want):
    for k, (g, w) in enumerate(zip(got, want)):
        if g != w:
            return k, g, w
    return None


def test_c1_golden_fixture_shape():
    rec = G.load()
    assert rec["ticks"] == G.TICKS and len(rec["raster"]) == len(rec["spikes"]) == G.TICKS
    kinds = [s[0] for s in rec["states"]]
    assert kinds == ["start"] + ["before_night", "after_night"] * 3 + ["end"]
    assert all(s[3] > 0 for s in rec["states"] if s[0] == "after_night")  # every Night sprouts
    assert 0.01 < sum(rec["spikes"]) / (G.TICKS * 4096) < 0.2  # the brain fires


def test_c1_oracle_with_host_sprouting_matches_the_golden_digests():
    want = G.load()
    got = G.run_oracle(native_sprouting=True)
    assert _first_mismatch(got["spikes"], want["spikes"]) is None, _first_mismatch(got["spikes"], want["spikes"])
    assert _first_mismatch(got["raster"], want["raster"]) is None, _first_mismatch(got["raster"], want["raster"])
    for g, w in zip(got["states"], want["states"]):
        assert g == w, (g, w)
