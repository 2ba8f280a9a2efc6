"""The staged dendrite scan (per-block presynaptic entries in LDS) against the
CPU oracle and against the direct scan, on the cases that take its other
paths: blocks whose distinct presynaptic axons overflow the LDS capacity,
segments whose segment + propagation length passes 255 (the 16-byte
fallback), filled slots after an empty one (ignored, as the reference's
`break`), and targets changed between steps (the plan is rebuilt)."""
import ctypes as C

import numpy as np
import pytest

from axicor_amd import brain as B
from axicor_amd._ffi import day_lib
from axicor_amd.layout import StateView
from helpers import Pair

pytestmark = pytest.mark.gpu


@pytest.fixture
def direct_scan():
    """Runs the test with the direct scan, restoring the staged default."""
    day_lib().axc_staging_enable(0)
    yield
    day_lib().axc_staging_enable(1)


def test_overflow_blocks_far_targets(gpu):
    """Targets drawn far from their somas (mean distance 60 voxels): most
    256-soma blocks have more distinct presynaptic axons than the LDS holds
    and read the reference targets; the rest are staged."""
    spec = B.spec_small(12000, 120, 600, seed=21, axon_len=(1, 64))
    spec.mean_dist = 60.0
    p = Pair(B.generate(spec), n_out=32)
    for k in range(4):
        out_ref, _ = p.step(15, dopamine=(120 if k % 2 else -80))
        np.testing.assert_array_equal(p.outputs(15), out_ref)
        p.check(f"far targets batch {k}")


def test_wide_segments_use_the_compact_snapshot(gpu):
    """Axons 200-255 segments long and propagation length 60: a large share
    of slots has segment + length > 255."""
    lut = B.lut_from([B.template_variant(False, signal_propagation_length=60),
                      B.template_variant(True, signal_propagation_length=60)])
    spec = B.spec_small(6000, 64, 400, seed=23, axon_len=(200, 255), lut=lut)
    p = Pair(B.generate(spec), n_out=0)
    for k in range(5):
        p.step(40, dopamine=200)
        p.check(f"wide segments batch {k}")


def test_holes_after_the_first_empty_slot_are_ignored(gpu):
    """Somas with filled slots after an empty one (as sprouting leaves them
    until the next sort): the scan and GSOP stop at the first empty target."""
    b = B.generate(B.spec_small(4000, 80, 200, seed=25))
    v = StateView(b.state, b.padded_n)
    rng = np.random.default_rng(3)
    somas = rng.choice(4000, 600, replace=False)
    hole = rng.integers(5, 70, somas.size)
    v.targets[hole, somas] = 0
    p = Pair(b, n_out=0)
    for k in range(3):
        p.step(20, dopamine=300)
        p.check(f"holes batch {k}")


def test_targets_uploaded_between_steps_rebuild_the_plan(gpu):
    """A host upload of the targets (as the Night's sprouting upload does)
    between two batches: the next batch sees the new targets."""
    p = Pair(B.generate(B.spec_small(5000, 64, 250, seed=27)), n_out=0)
    p.step(20)
    p.check("before upload")
    v = StateView(p.ref_state, p.b.padded_n)
    rng = np.random.default_rng(5)
    cols = rng.choice(p.b.padded_n, 1500, replace=False)
    rows = rng.integers(0, 64, cols.size)
    v.targets[rows, cols] = (rng.integers(0, 200, cols.size).astype(np.uint32) << 24) | \
        (rng.integers(0, p.b.total_axons, cols.size).astype(np.uint32) + 1)
    tg = np.ascontiguousarray(v.targets)
    assert day_lib().gpu_memcpy_host_to_device(p.vram.ptrs.dendrite_targets, tg.ctypes.data, tg.nbytes)
    for k in range(3):
        p.step(20)
        p.check(f"after upload batch {k}")


def test_staged_and_direct_scans_agree(gpu):
    """The same brain stepped with the staged scan and with the direct one
    ends in the same state (and both equal the oracle)."""
    spec = B.spec_small(9000, 100, 450, seed=29, axon_len=(1, 255), lut=B.lut_c2(),
                        type_ppm=[400_000, 400_000, 100_000, 100_000])
    b = B.generate(spec)
    a = Pair(b.copy(), n_out=24)
    for k in range(3):
        a.step(30, dopamine=(200 - 150 * k))
        a.check(f"staged batch {k}")
    lib = day_lib()
    lib.axc_staging_enable(0)
    try:
        d = Pair(b.copy(), n_out=24)
        for k in range(3):
            d.step(30, dopamine=(200 - 150 * k))
            d.check(f"direct batch {k}")
    finally:
        lib.axc_staging_enable(1)
    assert np.array_equal(a.vram.download_state(), d.vram.download_state())
    assert np.array_equal(a.vram.download_axons(), d.vram.download_axons())


def test_direct_scan_per_tick(gpu, direct_scan):
    """The direct scan still passes the per-tick path against the oracle."""
    p = Pair(B.generate(B.spec_small(3000, 48, 150, seed=31)), n_out=16)
    for t in range(10):
        out_ref, _ = p.step(1, dopamine=50)
        np.testing.assert_array_equal(p.outputs(1), out_ref)
        p.check(f"direct tick {t}")
