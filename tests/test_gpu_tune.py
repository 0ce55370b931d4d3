"""GPU: the autotuner and the mixed Fleet form.

The tuner's choices (waves per direct workgroup, scoring strip width, the
mixed form's crossover) change no output bit: forced choices give the same
results as measured ones.  The mixed FindAll follows its plan member by
member: a member the gate routed to the FFT form returns exactly the plain
Find's six-tuple, every other member exactly the direct Find's, and the plan
itself is consistent (routed members measured within 1e-5 of the exact
scores, candidates above it kept direct)."""
import sys

import numpy as np
import pytest

from cvmatch_amd import synth

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def scene(seed, w, h, textured):
    rng = np.random.default_rng(seed)
    if textured:
        f = synth.box_blur5(rng.integers(0, 256, (h, w, 4), dtype=np.uint8))
    else:
        f = synth.add_noise(synth.flat_rect_scene(w, h, rng, n_rects=40, channels=4), 2.0, rng, channels=[0, 1, 2])
    f[..., 3] = 255
    subs = []
    for k in range(10):
        tw, th = int(rng.integers(20, 129)), int(rng.integers(20, 129))
        x, y = int(rng.integers(0, w - tw + 1)), int(rng.integers(0, h - th + 1))
        subs.append(np.ascontiguousarray(f[y:y + th, x:x + tw]))
    subs.append(np.full((12, 12, 4), 90, np.uint8))  # flat
    return np.ascontiguousarray(f), subs


@pytest.mark.parametrize("textured", [True, False])
def test_mixed_follows_its_plan(cvm, monkeypatch, textured):
    monkeypatch.setenv("CVMATCH_CROSSOVER", "24")  # every member at least 24 px on a side is a candidate
    frame, subs = scene(70 + textured, 720, 400, textured)
    ms = [cvm.NewMatcher(t) for t in subs]
    fl = cvm.NewFleet(*ms)
    say(f"[tune] mixed FindAll, textured={textured}: {len(ms)} members")
    got = [tuple(r) for r in fl.find_all(frame, mixed=True)]
    routed, gate, xo = fl.mixed_plan()
    assert xo == 24
    say("[tune] routed", sum(routed), "gate", [round(g, 8) for g in gate])
    for i, m in enumerate(ms):
        if subs[i].std() == 0:
            assert got[i] == (1.0, 0, 0, 1.0, 0, 0) and not routed[i]
            continue
        if routed[i]:
            assert 0.0 <= gate[i] <= 1e-5 or gate[i] == -2.0, (i, gate[i])
            assert got[i] == m.find(frame), i
        else:
            assert gate[i] > 1e-5 or gate[i] == -1.0, (i, gate[i])
            assert got[i] == m.find(frame, direct=True), i
    # replays (captured graph) give the same
    for _ in range(3):
        assert [tuple(r) for r in fl.find_all(frame, mixed=True)] == got


def test_mixed_routed_members_take_the_fft_form(cvm, monkeypatch):
    """With the gate opened wide every candidate is routed: those members
    return the plain Find's six-tuple bit for bit (the form the gate
    measured), the others the direct Find's."""
    monkeypatch.setenv("CVMATCH_CROSSOVER", "64")
    monkeypatch.setenv("CVMATCH_MIXED_GATE", "1.0")
    frame, subs = scene(90, 720, 400, False)
    ms = [cvm.NewMatcher(t) for t in subs]
    fl = cvm.NewFleet(*ms)
    for rep in range(3):
        got = [tuple(r) for r in fl.find_all(frame, mixed=True)]
        routed, gate, xo = fl.mixed_plan()
        assert xo == 64
        assert any(routed) and not all(routed)
        for i, m in enumerate(ms):
            big = max(subs[i].shape[:2]) >= 64 and subs[i].std() > 0
            assert routed[i] == big, (i, subs[i].shape)
            assert got[i] == (m.find(frame) if routed[i] else m.find(frame, direct=True)), (rep, i)


def test_mixed_without_crossover_is_direct(cvm, monkeypatch):
    monkeypatch.setenv("CVMATCH_CROSSOVER", "0")
    frame, subs = scene(80, 640, 360, True)
    fl = cvm.NewFleet(*[cvm.NewMatcher(t) for t in subs])
    assert [tuple(r) for r in fl.find_all(frame, mixed=True)] == [tuple(r) for r in fl.find_all(frame, direct=True)]
    routed, _, xo = fl.mixed_plan()
    assert xo == 0 and not any(routed)


@pytest.mark.parametrize("waves,q", [("1", "4"), ("2", "2"), ("4", "4")])
def test_forced_choices_change_no_bits(cvm, monkeypatch, waves, q):
    frames, tpls = synth.workload_fleet(seed=12, frames=3, members=20, w=800, h=480)
    frame = np.ascontiguousarray(frames[1])
    want = None
    for force in (False, True):
        if force:
            monkeypatch.setenv("CVMATCH_DIRECT_WAVES", waves)
            monkeypatch.setenv("CVMATCH_NORM_Q", q)
        fl = cvm.NewFleet(*[cvm.NewMatcher(t) for t in tpls])
        got = [tuple(r) for r in fl.find_all(frame, direct=True)]
        one = [cvm.NewMatcher(t).find(frame, direct=True) for t in tpls[:4]]
        if want is None:
            want = (got, one)
        else:
            assert (got, one) == want


def test_tuner_report(cvm):
    frames, tpls = synth.workload_fleet(seed=13, frames=2, members=12, w=640, h=360)
    fl = cvm.NewFleet(*[cvm.NewMatcher(t) for t in tpls])
    fl.find_all(np.ascontiguousarray(frames[0]), direct=True)
    rep = cvm.tune_report()
    keys = list(rep)
    say("[tune] report keys", len(keys))
    assert any(k.startswith("fleet.direct.waves/") for k in keys)
    assert any(k.startswith("fleet.norm.q/") for k in keys)
    for k, d in rep.items():
        assert d["source"] in ("measured", "cache", "override", "default"), (k, d)
        if k.startswith("fleet.direct.waves/") and d["source"] == "measured":
            assert d["choice"] in (1, 2, 4) and len(d["ms"]) == 3


@pytest.mark.parametrize("q", ["1", "2", "4", "9", "10", "12"])
def test_forced_strip_forms_change_no_bits(cvm, orc, monkeypatch, q):
    """Every strip form of a one-shot call's normalisation forced, full bands
    (q) and half bands (q + 8), in the compact and the forked call forms:
    the plain map is the oracle's bit for bit, the direct map the unforced
    direct call's.  Result heights that are not multiples of 8 or 16 and
    template heights that are not multiples of 16 (window rows read directly
    at the window's end) are among the shapes."""
    rng = np.random.default_rng(500 + int(q))
    cases = [(200, 150, 4, 4, 24, 19), (333, 97, 1, 1, 40, 33), (150, 90, 3, 4, 17, 20)]
    for iw, ih, cn, step, tw, th in cases:
        img = rng.integers(0, 256, ih * iw * step, dtype=np.uint8)
        x, y = int(rng.integers(0, iw - tw + 1)), int(rng.integers(0, ih - th + 1))
        tpl = np.ascontiguousarray(img.reshape(ih, iw * step)[y:y + th, x * step:(x + tw) * step]).ravel()
        args = (img, iw * step, iw, ih, tpl, tw * step, tw, th, cn, step)
        want = orc.match_u8(*args, threads=8)
        monkeypatch.delenv("CVMATCH_NORM_Q", raising=False)
        want_direct = cvm.match_u8(*args, want_map=True, direct=True)
        for compact in ("2500000", "0"):
            say(f"[tune] strip form {q}, {iw}x{ih} cn {cn} tpl {tw}x{th}, compact bound {compact}")
            monkeypatch.setenv("CVMATCH_NORM_Q", q)
            monkeypatch.setenv("CVMATCH_COMPACT", compact)
            got = cvm.match_u8(*args, want_map=True)
            assert got[1] == want[1], (got[1], want[1])
            assert np.array_equal(got[0].view(np.uint32), want[0].view(np.uint32))
            gd = cvm.match_u8(*args, want_map=True, direct=True)
            assert gd[1] == want_direct[1]
            assert np.array_equal(gd[0].view(np.uint32), want_direct[0].view(np.uint32))
