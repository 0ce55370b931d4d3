"""GPU: the fused direct Fleet (correlation and FP64 scoring in one launch per
member group, no C' map) against the unfused form and the one-shot direct
Match.

Both forms compute the same exact integers and evaluate the same binary64
score op for op, so every member's six-tuple must be identical; the one-shot
direct Match (itself equal bit for bit to the FP64 oracle rounded to float32,
tests/test_gpu_direct.py) is the reference.  The cases cover gray (one byte
a pixel) and RGBA with cn 3 and 4, template widths on both sides of the
scoring strip's limit (4 or 8 columns per lane), 64-bit window square sums
(cn * 65025 * area >= 2^32), result maps whose sides are not multiples of the
32 x 128 block, and every waves-per-workgroup choice.
"""
import sys

import numpy as np
import pytest

from cvmatch_amd import synth

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def scene(seed, w, h, kind):
    rng = np.random.default_rng(seed)
    if kind == "gray":
        f = synth.flat_rect_scene(w, h, rng, n_rects=30, channels=1)
        f = synth.add_noise(f, 3.0, rng)
        return np.ascontiguousarray(f), rng
    f = synth.flat_rect_scene(w, h, rng, n_rects=30, channels=4)
    f = synth.add_noise(f, 3.0, rng, channels=[0, 1, 2, 3] if kind == "rgba4" else [0, 1, 2])
    if kind == "rgba3":
        f[..., 3] = 255
    return np.ascontiguousarray(f), rng


CASES = {
    # kind: (frame w, h, template sizes (tw, th))
    "gray": (640, 360, [(16, 16), (40, 33), (96, 17), (130, 20), (200, 61), (257, 30), (5, 250)]),
    "rgba3": (512, 300, [(16, 16), (64, 64), (100, 120), (120, 200), (140, 100), (33, 7)]),
    "rgba4": (520, 320, [(16, 16), (65, 255), (130, 100), (48, 90), (7, 9)]),
}


def crops(frame, rng, sizes):
    h, w = frame.shape[:2]
    out = []
    for tw, th in sizes:
        x, y = int(rng.integers(0, w - tw + 1)), int(rng.integers(0, h - th + 1))
        out.append(np.ascontiguousarray(frame[y:y + th, x:x + tw]))
    return out


def fleet_results(cvm, subs, frame, reps=2):
    fl = cvm.NewFleet(*[(cvm.NewGrayMatcher if s.ndim == 2 else cvm.NewMatcher)(s) for s in subs])
    got = [tuple(r) for r in fl.find_all(frame, direct=True)]
    for _ in range(reps - 1):  # later calls replay the cached tables (and the captured graph)
        assert [tuple(r) for r in fl.find_all(frame, direct=True)] == got
    return got


def fused_choices(cvm):
    return {k: d for k, d in cvm.tune_report().items() if k.startswith("fleet.fused/")}


@pytest.mark.parametrize("kind", list(CASES))
def test_fused_equals_unfused_and_direct_match(cvm, monkeypatch, kind):
    w, h, sizes = CASES[kind]
    frame, rng = scene(300 + len(kind), w, h, kind)
    subs = crops(frame, rng, sizes)
    say(f"[fused] {kind} {w}x{h}: templates {sizes}")
    want = [tuple(cvm.match_direct(frame, s)[1]) for s in subs]
    monkeypatch.setenv("CVMATCH_FLEET_FUSED", "0")
    unfused = fleet_results(cvm, subs, frame)
    assert unfused == want
    monkeypatch.setenv("CVMATCH_FLEET_FUSED", "1")
    fused = fleet_results(cvm, subs, frame)
    ch = fused_choices(cvm)
    assert any(d["choice"] == 1 and d["source"] == "override" for d in ch.values()), ch
    for i, (a, b) in enumerate(zip(fused, want)):
        assert a == b, (kind, sizes[i], a, b)


@pytest.mark.parametrize("waves,merge", [("1", "1"), ("2", "1"), ("4", "1"), ("2", "0")])
def test_fused_waves_change_no_bits(cvm, monkeypatch, waves, merge):
    """Every waves-per-workgroup choice, and member groups merged into one
    launch per (K step, square width) or kept one launch per chunk count."""
    frames, tpls = synth.workload_fleet(seed=31, frames=2, members=24, w=640, h=384)
    frame = np.ascontiguousarray(frames[1])
    say(f"[fused] 640x384 RGBA fleet of {len(tpls)}, waves {waves}, merge {merge}")
    want = [tuple(cvm.match_direct(frame, t)[1]) if t.std() > 0 else None for t in tpls]
    monkeypatch.setenv("CVMATCH_FLEET_FUSED", "1")
    monkeypatch.setenv("CVMATCH_FLEET_MERGE", merge)
    monkeypatch.setenv("CVMATCH_DIRECT_WAVES", waves)
    got = fleet_results(cvm, tpls, frame)
    for i, (g, wv) in enumerate(zip(got, want)):
        if wv is not None:
            assert g == wv, (i, tpls[i].shape, g, wv)


def test_unaligned_frame_takes_the_unfused_form(cvm, monkeypatch):
    """A frame whose rows are not 16-byte multiples cannot be staged by the
    fused kernel's vector loads: FindAll falls back to the unfused launches
    and its results are unchanged."""
    frame, rng = scene(77, 322, 150, "rgba3")
    subs = crops(frame, rng, [(16, 16), (40, 30), (70, 50)])
    say("[fused] unaligned 322x150 RGBA")
    monkeypatch.setenv("CVMATCH_FLEET_FUSED", "1")
    got = fleet_results(cvm, subs, frame)
    assert got == [tuple(cvm.match_direct(frame, s)[1]) for s in subs]


@pytest.mark.parametrize("fused", ["0", "1"])
def test_tables_follow_the_frame_shape(cvm, monkeypatch, fused):
    """Either form, forced: after a larger frame (plain and direct FindAlls,
    which grow and move the shared scratch and rebuild the job tables), the
    next direct FindAll on the small frame rebuilds its own tables and gives
    the same results as the first."""
    monkeypatch.setenv("CVMATCH_FLEET_FUSED", fused)
    rng = np.random.default_rng(78)
    small = synth.flat_rect_scene(400, 300, rng, n_rects=20, channels=4)
    small[..., 3] = 255
    big = synth.flat_rect_scene(1600, 1000, rng, n_rects=60, channels=4)
    big[..., 3] = 255
    subs = [small[50:82, 60:100].copy(), small[200:224, 300:330].copy(), small[10:110, 10:140].copy()]
    say(f"[fused] tables across frame shapes, fused={fused}")
    fl = cvm.NewFleet(*[cvm.NewMatcher(s) for s in subs])
    first = [tuple(r) for r in fl.find_all(small, direct=True)]
    assert first == [tuple(cvm.match_direct(small, s)[1]) for s in subs]
    fl.find_all(big)
    assert [tuple(r) for r in fl.find_all(small, direct=True)] == first
    big_direct = [tuple(r) for r in fl.find_all(big, direct=True)]
    assert big_direct == [tuple(cvm.match_direct(big, s)[1]) for s in subs]
    for _ in range(3):  # eager, captured, replayed
        assert [tuple(r) for r in fl.find_all(small, direct=True)] == first
