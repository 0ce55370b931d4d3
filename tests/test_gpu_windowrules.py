"""GPU: the window-rules engine on the reference's demo window
(engine_test.go:14-219 scenarios) and its parity with FindAll on host views:
every group's device-view FindAll equals FindAll on the same region passed
from the host, bit for bit, gray groups included."""
import sys

import numpy as np
import pytest

from cvmatch_amd import synth
from cvmatch_amd import windowrules as W

pytestmark = pytest.mark.gpu
R = synth.WINDOW_RECTS


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def tl(rect):
    return rect[:2]


def demo_rules(base):
    """demoRules main.go:107-127: whole window / ROI, color / gray, both modes,
    template reuse across rules, an ROI too small for its template."""
    def tpl(i, r):
        return W.Template(i, synth.crop(base, r))
    panel = tpl(105, R["panel"])
    return [
        W.Rule(ID=1, Name="confirm-button", Mode=W.ModeFirst, Threshold=0.95,
               Templates=[W.Template(100, synth.decoy_widget(120, 36, 99)), tpl(101, R["ok"])]),
        W.Rule(ID=2, Name="tray-status", ROI=(1150, 0, 1280, 48), Gray=True, Mode=W.ModeAll, Threshold=0.95,
               Templates=[tpl(102, R["wifi"]), tpl(103, R["mute"]), W.Template(104, synth.decoy_widget(24, 24, 77))]),
        W.Rule(ID=3, Name="main-panel", ROI=(180, 180, 520, 420), Mode=W.ModeFirst, Threshold=0.95,
               Templates=[panel]),
        W.Rule(ID=4, Name="oversized-roi", ROI=(0, 0, 40, 40), Mode=W.ModeFirst, Threshold=0.95, Templates=[panel]),
    ]


def test_demo_scene(cvm):
    base = synth.window_frame()
    eng = W.NewEngine(demo_rules(base))
    say("[windowrules] demo scene")
    first = eng.run(base)
    r = first[0]
    assert r.Matched and r.Best.TemplateID == 101 and tl(r.Best.Rect) == tl(R["ok"]), r
    r = first[1]
    assert r.Matched and len(r.All) == 3 and r.Best.TemplateID in (102, 103), r
    want = {102: tl(R["wifi"]), 103: tl(R["mute"])}
    for h in r.All:
        if h.TemplateID in want:
            assert tl(h.Rect) == want[h.TemplateID], h
    assert first[2].Matched and tl(first[2].Best.Rect) == tl(R["panel"])
    assert not first[3].Matched
    assert eng.run(base) == first  # warm re-run: fleet caches hit
    moved = eng.run(synth.window_frame(move_button=True))
    assert moved[0].Matched and tl(moved[0].Best.Rect) == tl(R["ok_moved"])


def test_input_isolation(cvm):
    base = synth.window_frame()
    rules = demo_rules(base)
    eng = W.NewEngine(rules)
    want = eng.run(base)
    rules[0].Threshold = 2
    rules[0].Mode = W.ModeAll
    rules[0].Templates[1] = W.Template(999, np.zeros((8, 8, 4), np.uint8))
    rules[1].ROI = (0, 0, 10, 10)
    rules[1].Templates = []
    assert eng.run(base) == want


def test_reversed_roi_and_partial_fit(cvm):
    base = synth.window_frame()
    r = demo_rules(base)[1]
    r.ROI = (1280, 48, 1150, 0)
    res = W.NewEngine([r]).run(base)
    assert res[0].Matched and tl(res[0].Best.Rect) in (tl(R["wifi"]), tl(R["mute"]))
    eng = W.NewEngine([W.Rule(ID=10, Name="partial", ROI=(1150, 0, 1280, 48), Mode=W.ModeAll, Threshold=0.95,
                              Templates=[W.Template(1, synth.crop(base, R["wifi"])),
                                         W.Template(2, synth.crop(base, R["panel"]))])])
    for rnd in range(2):  # the second round hits the memoised sub-fleet
        r = eng.run(base)[0]
        assert r.Matched and len(r.All) == 1 and r.All[0].TemplateID == 1 and tl(r.Best.Rect) == tl(R["wifi"]), rnd


def test_groups_equal_host_view_findall(cvm):
    """Each group's results equal FindAll on the host view of its region
    (gray groups: of the gray frame), bit for bit: the device views, the
    shared upload and the one-pass gray conversion change nothing."""
    base = synth.window_frame(move_button=True)
    eng = W.NewEngine(demo_rules(synth.window_frame()))
    eng.run(base)
    gray = cvm.rgba_to_gray(base)
    for g in eng._groups:
        v = eng._group_view(g)
        if v is None:
            continue
        ptr, vstride, (x0, y0, x1, y1) = v
        fleet, idx = g.fit(x1 - x0, y1 - y0)
        if fleet is None:
            continue
        want = fleet.find_all((gray if g.gray else base)[y0:y1, x0:x1])
        got = fleet.find_all_device(ptr, vstride, x1 - x0, y1 - y0)
        say(f"[windowrules] group roi {g.roi} gray {g.gray}: {len(want)} members")
        for a, b in zip(got, want):
            assert tuple(a) == tuple(b) and np.float32(a.MaxV).tobytes() == np.float32(b.MaxV).tobytes()
