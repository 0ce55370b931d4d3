"""GPU: every occurrence (threshold + local-maximum suppression) on the device
against the oracle's restatement of TestFindAllOccurrences' loop
(cvmatch_test.go:346-365): same count, same peaks in raster order, same
score bits.  Quantised maps make ties (tied maxima are all peaks); windows
range from 0 x 0 to wider than the map; thresholds from dense (every point a
candidate) to sparse; caps below the count keep the raster-order prefix."""
import ctypes as C
import sys

import numpy as np
import pytest

from cvmatch_amd import synth

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


class Dev:
    def __init__(self, L, n):
        self.L = L
        self.p = C.c_void_p()
        assert L.cvm_device_alloc(max(n, 4), C.byref(self.p)) == 0

    def free(self):
        self.L.cvm_device_free(self.p)


def device_peaks(cvm, resp, thresh, rx, ry, cap):
    L = cvm.lib()
    h, w = resp.shape
    resp = np.ascontiguousarray(resp, dtype=np.float32)
    dm, dxy, dv = Dev(L, resp.nbytes), Dev(L, 8 * cap), Dev(L, 4 * cap)
    try:
        assert L.cvm_copy_h2d(dm.p, C.c_void_p(resp.ctypes.data), resp.nbytes) == 0
        n = cvm.find_peaks_device(dm.p, w, h, thresh, rx, ry, dxy.p, dv.p, cap)
        k = min(n, cap)
        xy = np.zeros((max(k, 1), 2), np.int32)
        vals = np.zeros(max(k, 1), np.float32)
        if k:
            assert L.cvm_copy_d2h(C.c_void_p(xy.ctypes.data), dxy.p, 8 * k) == 0
            assert L.cvm_copy_d2h(C.c_void_p(vals.ctypes.data), dv.p, 4 * k) == 0
        return n, xy[:k], vals[:k]
    finally:
        for d in (dm, dxy, dv):
            d.free()


def same(got, want, label):
    assert got[0] == want[0], (label, got[0], want[0])
    assert np.array_equal(got[1], want[1]), label
    assert np.array_equal(bits(got[2]), bits(want[2])), label


@pytest.mark.parametrize("h,w,levels", [(200, 300, 7), (3, 1000, 4), (700, 5, 4), (1, 1, 2), (257, 513, 1000)])
def test_device_peaks_equal_oracle(cvm, orc, h, w, levels):
    rng = np.random.default_rng(h * 7 + w)
    resp = (rng.integers(0, levels, (h, w)).astype(np.float32) / max(levels - 1, 1) * 2 - 1).astype(np.float32)
    windows = [(0, 0), (1, 2), (5, 3), (16, 12), (40, 70), (384, 4), (385, 4)]
    if h * w <= 5000:
        windows.append((3000, 900))  # wider than the map on both axes
    for rx, ry in windows:
        for thresh in (-5.0, 0.0, 0.75):
            say(f"[peaks] map {w}x{h} levels {levels} rx {rx} ry {ry} thresh {thresh}")
            # the oracle scans its window literally: give it the same window
            # clipped to the map (the same elements); the GPU gets rx, ry
            want = orc.find_peaks(resp, thresh, min(rx, w - 1), min(ry, h - 1))
            same(device_peaks(cvm, resp, thresh, rx, ry, max(want[0], 1)), want, (rx, ry, thresh))
            if want[0] > 3:
                cap = want[0] // 3
                got = device_peaks(cvm, resp, thresh, rx, ry, cap)
                assert got[0] == want[0]
                assert np.array_equal(got[1], want[1][:cap]) and np.array_equal(bits(got[2]), bits(want[2][:cap]))
            if thresh == 0.0 and rx == 1:
                assert device_peaks(cvm, resp, thresh, rx, ry, 0)[0] == want[0]  # count only


def test_device_peaks_wide_windows(cvm, orc):
    """Row windows wide enough to leave the doubling form for the linear scan
    of the staged row (more than 8192 staged elements), and just below."""
    rng = np.random.default_rng(9000)
    x = np.arange(9000)
    # a slow quantised wave (period 2300, decaying): the peak set depends on
    # the window width (2 to 30 peaks for the windows below)
    resp = np.stack([0.9 * np.sin(2 * np.pi * x / 2300.0 + 1.3 * y) * (1 - x / 20000.0) for y in range(3)])
    resp = (np.round((resp + rng.normal(0, 0.01, resp.shape)) * 100) / 100).astype(np.float32)
    for rx, ry in [(5000, 1), (3969, 0), (3967, 2), (8064, 2), (2000, 0)]:
        for thresh in (-5.0, 0.5):
            say(f"[peaks] wide map 9000x3 rx {rx} ry {ry} thresh {thresh}")
            want = orc.find_peaks(resp, thresh, rx, ry)
            same(device_peaks(cvm, resp, thresh, rx, ry, max(want[0], 1)), want, (rx, ry, thresh))
    # past the staged row's limit (64 KiB of LDS): refused, not truncated
    with pytest.raises(cvm.CvmatchError) as e:
        device_peaks(cvm, resp, 0.5, 8065, 0, 4)
    assert e.value.code == cvm.CVM_E_UNSUPPORTED


def test_find_all_occurrences(cvm, orc):
    """TestFindAllOccurrences cvmatch_test.go:320-381 with this suite's RNG:
    five stamped copies of a 32x24 template on weak noise are recovered
    exactly, in raster order, and the peaks equal the oracle's over the map."""
    rng = np.random.default_rng(101)
    parent = rng.integers(0, 64, (300, 400, 4), dtype=np.uint8)
    parent[..., 3] = 255
    tpl = rng.integers(0, 256, (24, 32, 4), dtype=np.uint8)
    tpl[..., 3] = 255
    spots = [(10, 12), (200, 40), (350, 260), (60, 240), (180, 150)]
    for x, y in spots:
        parent[y:y + 24, x:x + 32] = tpl
    peaks, total, t6 = cvm.match_peaks(parent, tpl, 0.999, 16, 12)
    assert total == 5 and [(x, y) for x, y, _ in peaks] == sorted(spots, key=lambda p: (p[1], p[0]))
    assert t6 == cvm.match(parent, tpl)
    resp, _, _ = cvm.match_map(parent, tpl)
    n, xy, vals = orc.find_peaks(resp, 0.999, 16, 12)
    assert n == total and [tuple(p) for p in xy] == [(x, y) for x, y, _ in peaks]
    assert np.array_equal(bits(vals), bits(np.array([v for _, _, v in peaks], np.float32)))


def test_match_peaks_4k_and_gray(cvm, orc):
    """The 4K RGBA map (3713 x 2033) and a 1080p gray map: peaks on the GPU
    equal the oracle's peaks of the (bit-exact) map, for sparse thresholds and
    a dense one (every point a candidate); the six-tuple equals Match's."""
    parent, sub, (cx, cy) = synth.workload_map_4k(seed=3, alpha="random")
    resp, _, _ = cvm.match_map(parent, sub)
    for thresh, rx, ry in [(0.5, 64, 64), (0.05, 8, 3), (-1.0, 2, 2)]:
        say(f"[peaks] 4K RGBA thresh {thresh} rx {rx} ry {ry}")
        want = orc.find_peaks(resp, thresh, rx, ry, cap=1 << 20)
        peaks, total, t6 = cvm.match_peaks(parent, sub, thresh, rx, ry, cap=1 << 20)
        assert total == want[0]
        assert [(x, y) for x, y, _ in peaks] == [tuple(p) for p in want[1]]
        assert np.array_equal(bits(np.array([v for _, _, v in peaks], np.float32)), bits(want[2]))
        assert t6 == cvm.match(parent, sub)
    assert (cx, cy) in [(x, y) for x, y, _ in cvm.match_peaks(parent, sub, 0.5)[0]]
    gray, gsub = synth.workload_gray_ui_1080p(seed=2)[:2]
    gh, gw = gray.shape
    sh, sw = gsub.shape
    gm, _ = cvm.match_u8(gray, gw, gw, gh, gsub, sw, sw, sh, 1, 1, want_map=True)
    want = orc.find_peaks(gm, 0.9, sw // 2, sh // 2, cap=5000)
    peaks, total, _ = cvm.match_peaks(gray, gsub, 0.9, cap=5000)
    assert total == want[0] and [(x, y) for x, y, _ in peaks] == [tuple(p) for p in want[1]]


def test_flat_template_every_point(cvm):
    """A flat template gives a map of ones (impl.go:1137): every point ties,
    so every point is a peak; the cap keeps the first ones in raster order."""
    rng = np.random.default_rng(4)
    parent = rng.integers(0, 256, (60, 70, 4), dtype=np.uint8)
    flat = np.full((9, 11, 4), 40, np.uint8)
    peaks, total, _ = cvm.match_peaks(parent, flat, 0.999, 3, 3, cap=100)
    rw, rh = 70 - 11 + 1, 60 - 9 + 1
    assert total == rw * rh
    assert [(x, y) for x, y, _ in peaks] == [(i % rw, i // rw) for i in range(100)]
    assert all(v == 1.0 for _, _, v in peaks)
