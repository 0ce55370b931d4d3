"""GPU: every public call is safe for concurrent use (README.md:419-431;
TestConcurrentMatch concurrency_test.go:13-45, TestMatcherConcurrent
matcher_test.go:69-95, TestFleetConcurrent fleet_test.go:155-180): eight
host threads calling the library at once (ctypes releases the GIL during
each call) get exactly the single-threaded results."""
import sys
import threading

import numpy as np

import pytest

from cvmatch_amd import synth

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


def run_threads(fn, workers=8):
    errors = []

    def body(w):
        try:
            fn(w)
        except BaseException as e:  # reported below, never lost in the thread
            errors.append((w, repr(e)))

    ts = [threading.Thread(target=body, args=(w,)) for w in range(workers)]
    for t in ts:
        t.start()
    for t in ts:
        t.join(timeout=120)
    assert not any(t.is_alive() for t in ts), "worker did not finish"
    assert not errors, errors


def test_concurrent_match_gray_map_peaks(cvm):
    rng = np.random.default_rng(31)
    parent = rng.integers(0, 256, (240, 320, 4), dtype=np.uint8)
    sub = parent[80:104, 100:132].copy()
    gp, gs = cvm.rgba_to_gray(parent), cvm.rgba_to_gray(sub)
    want = cvm.match(parent, sub)
    want_g = cvm.match_gray(gp, gs)
    want_map = bits(cvm.match_map(parent, sub)[0])
    want_pk = cvm.match_peaks(parent, sub, 0.5)
    want_d = cvm.match_direct(parent, sub)[1]
    say("[concurrency] match / gray / map / peaks / direct, 8 threads x 5")

    def work(w):
        for _ in range(5):
            assert cvm.match(parent, sub) == want
            assert cvm.match_gray(gp, gs) == want_g
            assert np.array_equal(bits(cvm.match_map(parent, sub)[0]), want_map)
            assert cvm.match_peaks(parent, sub, 0.5) == want_pk
            assert tuple(cvm.match_direct(parent, sub)[1]) == tuple(want_d)

    run_threads(work)


def test_concurrent_matcher_fleet_engine(cvm):
    frame = synth.flat_rect_scene(640, 400, np.random.default_rng(7), n_rects=30, channels=4)
    frame[..., 3] = 255
    subs = [frame[y:y + h, x:x + w].copy() for (x, y, w, h) in
            [(10, 20, 32, 24), (300, 100, 48, 48), (500, 300, 20, 40), (100, 250, 64, 32)]]
    ms = [cvm.NewMatcher(s) for s in subs]
    fl = cvm.NewFleet(*ms)
    want_find = [m.find(frame) for m in ms]
    want_all = [tuple(r) for r in fl.find_all(frame)]
    want_direct = [tuple(r) for r in fl.find_all(frame, direct=True)]
    from cvmatch_amd import windowrules as W
    base = synth.window_frame()
    eng = W.NewEngine([W.Rule(1, "ok", Threshold=0.95, Templates=[W.Template(1, synth.crop(base, synth.WINDOW_RECTS["ok"]))]),
                       W.Rule(2, "tray", ROI=(1150, 0, 1280, 48), Gray=True, Mode=W.ModeAll, Threshold=0.95,
                              Templates=[W.Template(2, synth.crop(base, synth.WINDOW_RECTS["wifi"]))])])
    want_eng = eng.run(base)
    say("[concurrency] Find / FindAll / direct FindAll / engine Run, 8 threads x 4")

    def work(w):
        for _ in range(4):
            assert [ms[i].find(frame) for i in range(len(ms))] == want_find
            assert [tuple(r) for r in fl.find_all(frame)] == want_all
            assert [tuple(r) for r in fl.find_all(frame, direct=True)] == want_direct
            assert eng.run(base) == want_eng

    run_threads(work)
