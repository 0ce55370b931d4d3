"""GPU: the direct (mixed-precision) form against the FP64 oracle.

The direct form computes the window x template sum exactly on int8 MFMA and
every score in binary64 from exact integers, so each score must equal the
FP64 oracle's score rounded once to float32 -- bit for bit, not merely within
1e-5 -- and the six-tuple must be the first-occurrence extrema of that map.
Shapes cover gray and RGBA, cn = 1 / 3 / 4 (the RGBA alpha rule both ways),
odd sizes, 1x1 and whole-image templates, strided sub-image views (the
unaligned load paths) and the shape limits.
"""
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


def first_extrema(m):
    """(min_v, min_x, min_y, max_v, max_x, max_y), first in raster order."""
    rw = m.shape[1]
    i0, i1 = int(np.argmin(m)), int(np.argmax(m))
    return (float(m.flat[i0]), i0 % rw, i0 // rw, float(m.flat[i1]), i1 % rw, i1 // rw)


def oracle_map(orc, parent, sub, cn):
    ph, pw = parent.shape[:2]
    sh, sw = sub.shape[:2]
    step = 1 if parent.ndim == 2 else 4
    p = np.ascontiguousarray(parent)
    s = np.ascontiguousarray(sub)
    return orc.fp64_map(p, pw * step, pw, ph, s, sw * step, sw, sh, cn, step).astype(np.float32)


def rgba_cn(orc, parent, sub):
    return 3 if orc.alpha_const(np.ascontiguousarray(parent)) and orc.alpha_const(np.ascontiguousarray(sub)) else 4


def check(cvm, orc, parent, sub, label):
    say(f"[direct] {label}: parent {parent.shape} template {sub.shape}")
    cn = 1 if parent.ndim == 2 else rgba_cn(orc, parent, sub)
    gm, gt = cvm.match_direct(parent, sub, want_map=True)
    om = oracle_map(orc, parent, sub, cn)
    diff = np.nonzero(bits(gm) != bits(om))
    assert diff[0].size == 0, (f"{label}: {diff[0].size} scores differ from the FP64 oracle in float32, first at "
                               f"{diff[0][0]},{diff[1][0]}: {gm[diff[0][0], diff[1][0]]} vs {om[diff[0][0], diff[1][0]]}")
    assert tuple(gt) == first_extrema(om), (label, gt, first_extrema(om))
    return gm, gt


GRAY_SHAPES = [  # (iw, ih, tw, th)
    (97, 61, 17, 13),
    (300, 200, 31, 29),
    (160, 120, 48, 48),
    (64, 64, 64, 64),
    (50, 40, 1, 1),
    (40, 300, 33, 257),
    (257, 33, 255, 7),
    (129, 77, 16, 16),
]


@pytest.mark.parametrize("shape", GRAY_SHAPES)
def test_direct_gray_equals_fp64(cvm, orc, shape):
    iw, ih, tw, th = shape
    rng = np.random.default_rng(iw * 1000 + th)
    parent = rng.integers(0, 256, size=(ih, iw), dtype=np.uint8)
    x, y = int(rng.integers(0, iw - tw + 1)), int(rng.integers(0, ih - th + 1))
    sub = parent[y:y + th, x:x + tw].copy()
    _, gt = check(cvm, orc, parent, sub, f"gray {shape}")
    assert gt[4:] == (x, y) or gt[3] == 1.0


RGBA_SHAPES = [  # (iw, ih, tw, th, alpha)
    (97, 61, 17, 13, "random"),
    (120, 90, 24, 18, "const"),
    (320, 200, 96, 32, "const"),
    (200, 180, 128, 128, "random"),
    (50, 40, 1, 1, "random"),
    (131, 75, 40, 21, "template-const"),
]


@pytest.mark.parametrize("shape", RGBA_SHAPES)
def test_direct_rgba_equals_fp64(cvm, orc, shape):
    iw, ih, tw, th, alpha = shape
    rng = np.random.default_rng(iw * 7 + th)
    parent = rng.integers(0, 256, size=(ih, iw, 4), dtype=np.uint8)
    if alpha == "const":
        parent[..., 3] = 255
    x, y = int(rng.integers(0, iw - tw + 1)), int(rng.integers(0, ih - th + 1))
    sub = parent[y:y + th, x:x + tw].copy()
    if alpha == "template-const":
        sub[..., 3] = 200  # constant template alpha, non-constant parent: cn = 4
    check(cvm, orc, parent, sub, f"rgba {shape}")


def test_direct_views_and_noise(cvm, orc):
    """Strided sub-image views (unaligned rows) and a noisy, blurred scene."""
    rng = np.random.default_rng(11)
    big = synth.box_blur5(rng.integers(0, 256, size=(230, 333, 4), dtype=np.uint8))
    parent = big[3:, 5:]  # row pitch 333*4, origin not 16-byte aligned
    sub = synth.add_noise(parent[40:72, 60:100].copy(), 2.0, rng, channels=[0, 1, 2])
    check(cvm, orc, parent, sub, "rgba view")
    g = big[..., 1]
    check(cvm, orc, g[7:, 1:], g[20:45, 30:61], "gray view")


def test_direct_flat_and_limits(cvm, orc):
    rng = np.random.default_rng(3)
    parent = rng.integers(0, 256, size=(80, 90), dtype=np.uint8)
    flat = np.full((9, 11), 77, dtype=np.uint8)
    gm, gt = cvm.match_direct(parent, flat, want_map=True)
    assert gt == (1.0, 0, 0, 1.0, 0, 0) and np.all(gm == 1.0)
    big = rng.integers(0, 256, size=(400, 400, 4), dtype=np.uint8)
    with pytest.raises(cvm.CvmatchError):
        cvm.match_direct(big, big[:200, :200].copy())  # area*cn = 160000
    tall = rng.integers(0, 256, size=(320, 60), dtype=np.uint8)
    with pytest.raises(cvm.CvmatchError):
        cvm.match_direct(tall, tall[:300, :10].copy())  # th > 257


@pytest.mark.parametrize("fw,fh", [(320, 240), (317, 201)])
def test_direct_matcher_and_fleet(cvm, orc, fw, fh):
    """Find and FindAll in the direct form equal the one-shot direct Match.
    The Fleet stages constant-alpha members as RGB (three bytes a pixel) and
    the others as RGBA, the one-shot call always RGBA; 317 columns take the
    unaligned staging loads."""
    rng = np.random.default_rng(21 + fw)
    frame = synth.flat_rect_scene(fw, fh, rng, n_rects=40, channels=4)
    frame[..., 3] = 255
    frame = synth.add_noise(frame, 3.0, rng, channels=[0, 1, 2])
    subs = []
    for (tw, th) in [(16, 16), (24, 40), (57, 13), (64, 64), (9, 31), (100, 70)]:
        x, y = int(rng.integers(0, fw - tw + 1)), int(rng.integers(0, fh - th + 1))
        subs.append(frame[y:y + th, x:x + tw].copy())
    subs[2][..., 3] = 9  # constant alpha still
    subs[4][0, 0, 3] = 0  # non-constant template alpha: cn = 4
    ms = [cvm.NewMatcher(s) for s in subs]
    want = []
    for s in subs:
        _, t = cvm.match_direct(frame, s)
        want.append(tuple(t))
    got_find = [tuple(m.find(frame, direct=True)) for m in ms]
    assert got_find == want
    fl = cvm.NewFleet(*ms)
    for rep in range(2):  # second call reuses the cached states and fragments
        got = [tuple(r) for r in fl.find_all(frame, direct=True)]
        assert got == want, rep
    # against the FP64 oracle directly
    for s, w in zip(subs, want):
        om = oracle_map(orc, frame, s, rgba_cn(orc, frame, s))
        assert w == first_extrema(om)


def test_direct_1440p_gray(cvm, orc):
    """The mixed-precision configuration: 2560x1440 gray, 16/24/32/48 crops;
    FP64 checked at 4096 seeded windows plus the extrema, crop found."""
    img, subs = synth.workload_direct_1440p()
    rng = np.random.default_rng(99)
    for sub, (x, y) in subs:
        th, tw = sub.shape
        say(f"[direct] 1440p gray {tw}x{th}")
        gm, gt = cvm.match_direct(img, sub, want_map=True)
        assert gt[4:] == (x, y) and gt[3] == 1.0, gt
        rh, rw = gm.shape
        xs = np.concatenate([rng.integers(0, rw, 4096), [gt[1], gt[4]]]).astype(np.int32)
        ys = np.concatenate([rng.integers(0, rh, 4096), [gt[2], gt[5]]]).astype(np.int32)
        ref = orc.fp64_points(img, img.shape[1], img.shape[1], img.shape[0], sub, tw, tw, th, 1, 1, xs, ys)
        assert np.array_equal(bits(gm[ys, xs]), bits(ref.astype(np.float32)))


@pytest.mark.parametrize("form", ["plain", "direct"])
def test_delta_vs_fp64_4k_and_ui(cvm, orc, form):
    """Deviation from the FP64 oracle on the 4K RGBA MatchMap and the 1080p UI
    gray workloads: 8192 seeded windows plus the extrema.  The direct form is
    exact to float32 rounding; the plain (reference float32 FFT) form is
    reported, with a loose bound."""
    rng = np.random.default_rng(5)
    parent, sub, _ = synth.workload_map_4k(seed=3, alpha="random")
    gray, gsub, _ = synth.workload_gray_ui_1080p(seed=2)[:3]
    cases = [("4k-rgba", parent, sub, 4, 4), ("1080p-ui-gray", gray, gsub, 1, 1)]
    for label, p, s, cn, step in cases:
        ph, pw = p.shape[:2]
        sh, sw = s.shape[:2]
        if form == "direct":
            m, t = cvm.match_direct(p, s, want_map=True)
        elif step == 4:
            m, _, _ = cvm.match_map(p, s)
            t = cvm.match(p, s)
        else:
            m, t = cvm.match_u8(p, pw, pw, ph, s, sw, sw, sh, 1, 1, want_map=True)
        rh, rw = m.shape
        xs = np.concatenate([rng.integers(0, rw, 8192), [t[1], t[4]]]).astype(np.int32)
        ys = np.concatenate([rng.integers(0, rh, 8192), [t[2], t[5]]]).astype(np.int32)
        ref = orc.fp64_points(np.ascontiguousarray(p), pw * step, pw, ph, np.ascontiguousarray(s), sw * step, sw, sh,
                              cn, step, xs, ys)
        d = np.abs(m[ys, xs].astype(np.float64) - ref)
        say(f"[delta] {form} {label}: max |d| = {d.max():.3e}, mean |d| = {d.mean():.3e} over {d.size} windows")
        if form == "direct":
            assert np.array_equal(bits(m[ys, xs]), bits(ref.astype(np.float32)))
        else:
            assert d.max() < 1e-3


@pytest.mark.parametrize("size", [16, 24, 32, 48])
def test_direct_1440p_full_map_fp64(cvm, orc, size):
    """The mixed-precision configuration over the whole map: 2560x1440 gray,
    the seeded size x size crop; every one of the ~3.5 M scores equals the
    FP64 oracle's rounded to float32 (max |d| is float32 rounding alone)."""
    img, subs = synth.workload_direct_1440p(sizes=(size,))
    sub, (x, y) = subs[0]
    say(f"[direct] 1440p gray {size}x{size}: full map vs FP64")
    gm, gt = cvm.match_direct(img, sub, want_map=True)
    om = orc.fp64_map(img, 2560, 2560, 1440, sub, size, size, size, 1, 1)
    d = np.abs(gm.astype(np.float64) - om)
    say(f"[direct] {size}x{size}: max |d| = {d.max():.3e} over {d.size} windows")
    assert np.array_equal(bits(gm), bits(om.astype(np.float32)))
    assert gt[4:] == (x, y) and tuple(gt) == first_extrema(om.astype(np.float32))


def test_direct_fleet_tables_survive_scratch_growth(cvm, orc):
    """The direct FindAll's cached job tables point into the fleet's
    normalisation scratch, which a plain FindAll on a larger frame grows (and
    so moves): the next direct FindAll on the small frame must rebuild its
    tables and give the same results."""
    rng = np.random.default_rng(77)
    small = synth.flat_rect_scene(400, 300, rng, n_rects=20, channels=4)
    small[..., 3] = 255
    big = synth.flat_rect_scene(1600, 1000, rng, n_rects=60, channels=4)
    big[..., 3] = 255
    subs = [small[50:82, 60:100].copy(), small[200:224, 300:330].copy()]
    fl = cvm.NewFleet(*[cvm.NewMatcher(s) for s in subs])
    first = [tuple(r) for r in fl.find_all(small, direct=True)]
    fl.find_all(big)  # the plain form grows the shared scratch
    assert [tuple(r) for r in fl.find_all(small, direct=True)] == first
    fl.find_all(big, direct=True)  # tables for the large frame, then back
    assert [tuple(r) for r in fl.find_all(small, direct=True)] == first
    for s, w in zip(subs, first):
        assert w == first_extrema(oracle_map(orc, small, s, 3))
