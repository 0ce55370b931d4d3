"""Window-rules engine: load-time validation (CPU only -- every check runs
before any template reaches the GPU), after engine_test.go:126-186."""
import numpy as np
import pytest

from cvmatch_amd import synth
from cvmatch_amd import windowrules as W


def test_roi_canonicalised_and_degenerate_rejected(cvm):
    assert W.resolve_roi((1280, 48, 1150, 0)) == (1150, 0, 1280, 48)
    assert W.resolve_roi((0, 0, 0, 0)) == (0, 0, 0, 0)
    tpl = synth.crop(synth.window_frame(), synth.WINDOW_RECTS["wifi"])
    with pytest.raises(W.EngineError, match="rule 2: degenerate ROI"):
        W.NewEngine([W.Rule(ID=2, ROI=(10, 10, 10, 50), Threshold=0.9, Templates=[W.Template(1, tpl)])])


def test_template_size_conflict(cvm):
    base = synth.window_frame()
    small = synth.crop(base, synth.WINDOW_RECTS["wifi"])  # 24x24
    big = synth.crop(base, synth.WINDOW_RECTS["panel"])  # 300x200, same id
    with pytest.raises(W.EngineError) as e:
        W.NewEngine([W.Rule(ID=1, ROI=(0, 0, 100, 100), Threshold=0.9, Templates=[W.Template(5, small)]),
                     W.Rule(ID=2, Threshold=0.9, Templates=[W.Template(5, big)])])
    assert "template 5" in str(e.value) and "300x200" in str(e.value)


def test_oversized_template_and_bad_rules(cvm):
    huge = np.zeros((2442, 2442, 4), np.uint8)  # 5,963,364 px > the cn=4 bound
    with pytest.raises(W.EngineError) as e:
        W.NewEngine([W.Rule(ID=7, Threshold=0.9, Templates=[W.Template(3, huge)])])
    assert "rule 7" in str(e.value) and "template 3" in str(e.value) and "exact-statistics" in str(e.value)
    tpl = np.zeros((4, 4, 4), np.uint8)
    with pytest.raises(W.EngineError, match="unknown mode 5"):
        W.NewEngine([W.Rule(ID=8, Mode=5, Templates=[W.Template(1, tpl)])])
    with pytest.raises(W.EngineError, match="template 4 has no image"):
        W.NewEngine([W.Rule(ID=9, Templates=[W.Template(4, None)])])
    with pytest.raises(W.EngineError, match="template 6 is empty"):
        W.NewEngine([W.Rule(ID=9, Templates=[W.Template(6, np.zeros((0, 3, 4), np.uint8))])])


def test_window_scene_generator():
    """The scene formulas: gradient corners, widget border, LCG interior."""
    f = synth.window_frame()
    assert tuple(f[0, 0]) == (30, 40, 64, 255) and tuple(f[799, 1279]) == (30 + 1279 * 50 // 1280, 89, 64, 255)
    assert tuple(f[200, 200]) == (20, 20, 24, 255)
    s = 33
    for _ in range(302):  # pixel (1, 1) of the panel is its 302nd in raster order
        s = (s * 1664525 + 1013904223) & 0xFFFFFFFF
    assert tuple(f[201, 201]) == (140 + ((s >> 8) & 63), 120 + ((s >> 16) & 63), 90 + ((s >> 24) & 63), 255)
    m = synth.window_frame(move_button=True)
    assert np.array_equal(f[:690], m[:690]) and not np.array_equal(f, m)
