"""GPU parity: the HIP path against the CPU oracle, bit for bit.

Every map and six-tuple produced on the MI355X through the C ABI must equal
the oracle's restatement of the reference's scalar op sequence exactly
(integer/index work and float bits alike; the tolerance is zero).  The golden
hash #110 is checked straight from the GPU map, tying the device path to the
reference's own recorded bits.
"""
import ctypes as C
import sys

import numpy as np
import pytest

from cvmatch_amd import synth

pytestmark = pytest.mark.gpu

GOLDEN_110 = 0x99534488138E0635

# goldenCases impl_golden_test.go:60-71 (shapes; inputs from PCG64 here)
GOLDEN_SHAPES = [
    (101, 97, 61, 17, 13, 1, 1, 0, 0),
    (102, 97, 61, 17, 13, 4, 4, 0, 0),
    (103, 120, 90, 24, 18, 3, 4, 0, 0),
    (104, 300, 200, 31, 29, 1, 1, 7, 3),
    (105, 64, 64, 64, 64, 1, 1, 0, 0),
    (106, 50, 40, 1, 1, 4, 4, 0, 0),
    (107, 640, 400, 96, 32, 3, 4, 0, 0),
    (108, 128, 96, 5, 4, 3, 3, 0, 0),
    (109, 40, 300, 33, 257, 1, 1, 0, 0),
]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


def same(gpu, ora):
    gm, gt = gpu
    om, ot = ora
    assert gt == ot, (gt, ot)
    if om is not None:
        diff = np.nonzero(bits(gm) != bits(om))
        assert diff[0].size == 0, f"{diff[0].size} elements differ, first at {diff[0][0]},{diff[1][0]}"


def test_fp_selftest(cvm):
    """The device's binary32 sqrt and quotient (the functions normOne uses)
    and int64 -> float32 conversion against numpy's correctly rounded ones,
    on a million seeded inputs: random floats of several ranges, random bit
    patterns over every finite positive float, subnormals."""
    rng = np.random.default_rng(1)
    n = 1 << 20
    a = np.concatenate([
        rng.random(n // 4, dtype=np.float32) * 4.0,
        (rng.random(n // 4) * 1e6).astype(np.float32),
        np.frombuffer(rng.integers(1, 0x7F7FFFFF, n // 4, dtype=np.uint32).tobytes(), np.float32),
        np.frombuffer(rng.integers(1, 0x00800000, n // 4, dtype=np.uint32).tobytes(), np.float32),  # subnormals
    ]).astype(np.float32)
    b = np.abs(a[::-1]).copy() + np.float32(1e-30)
    a = np.abs(a)
    ints = rng.integers(-(1 << 62), 1 << 62, n, dtype=np.int64)
    ints[: n // 8] >>= 20
    osq, odv, ocv = (np.empty(n, np.float32) for _ in range(3))
    p = lambda x: x.ctypes.data_as(C.c_void_p)
    assert cvm.lib().cvm_selftest_fp(p(a), p(b), p(ints), n, p(osq), p(odv), p(ocv)) == 0
    assert np.array_equal(bits(osq), bits(np.sqrt(a.astype(np.float64)).astype(np.float32)))
    assert np.array_equal(bits(odv), bits((a.astype(np.float64) / b.astype(np.float64)).astype(np.float32)))
    assert np.array_equal(bits(ocv), bits(ints.astype(np.float64).astype(np.float32)))


@pytest.mark.parametrize("lg", list(range(0, 14)))
def test_twiddles_bitwise(cvm, orc, lg):
    n = 1 << lg
    want_tri = orc.tri_twiddles(n)
    want_brev = orc.bitrev(n)
    tri = np.zeros(max(want_tri.size, 1) * 2, np.float32)
    brev = np.zeros(n, np.int32)
    ln = C.c_int32()
    assert cvm.lib().cvm_twiddles(n, tri.ctypes.data_as(C.c_void_p), brev.ctypes.data_as(C.c_void_p), C.byref(ln)) == 0
    assert ln.value == want_tri.size
    assert np.array_equal(tri[: 2 * ln.value].view(np.uint32), want_tri.view(np.float32).view(np.uint32))
    assert np.array_equal(brev, want_brev)


def test_golden_110_from_gpu(cvm, orc):
    iw, ih, tw, th = 200, 150, 32, 32
    img = np.full(iw * ih, 128, np.uint8)
    tpl = np.full(tw * th, 128, np.uint8)
    img[75 * iw + 100] = 129
    tpl[5 * tw + 7] = 129
    m, t = cvm.match_u8(img, iw, iw, ih, tpl, tw, tw, th, 1, 1, want_map=True)
    assert orc.golden_hash(m, t) == GOLDEN_110


@pytest.mark.parametrize("case", GOLDEN_SHAPES, ids=[str(c[0]) for c in GOLDEN_SHAPES])
def test_golden_shapes_bitwise(cvm, orc, case):
    seed, iw, ih, tw, th, cn, step, ipad, tpad = case
    say("golden shape", case)
    rng = np.random.default_rng(seed)
    istride, tstride = (iw + ipad) * step, (tw + tpad) * step
    img = rng.integers(0, 256, istride * ih, dtype=np.uint8)
    tpl = rng.integers(0, 256, tstride * th, dtype=np.uint8)
    same(cvm.match_u8(img, istride, iw, ih, tpl, tstride, tw, th, cn, step, want_map=True),
         orc.match_u8(img, istride, iw, ih, tpl, tstride, tw, th, cn, step, threads=8))
    # extrema-only call gives the same six-tuple
    _, t = cvm.match_u8(img, istride, iw, ih, tpl, tstride, tw, th, cn, step, want_map=False)
    _, ot = orc.match_u8(img, istride, iw, ih, tpl, tstride, tw, th, cn, step, want_map=False)
    assert t == ot


def _random_cases(n=24, seed=99):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        cn, step = [(1, 1), (3, 4), (4, 4), (3, 3), (2, 2), (1, 4)][i % 6]
        iw, ih = (int(v) for v in rng.integers(20, 900, 2))
        tw = int(rng.integers(1, min(iw, 200) + 1))
        th = int(rng.integers(1, min(ih, 200) + 1))
        out.append((i, iw, ih, tw, th, cn, step, int(rng.integers(0, 5)), int(rng.integers(0, 3))))
    return out


@pytest.mark.parametrize("case", _random_cases(), ids=lambda c: f"r{c[0]}")
def test_random_shapes_bitwise(cvm, orc, case):
    i, iw, ih, tw, th, cn, step, ipad, tpad = case
    say("random shape", case)
    rng = np.random.default_rng(1000 + i)
    istride, tstride = (iw + ipad) * step, (tw + tpad) * step
    img = rng.integers(0, 256, istride * ih, dtype=np.uint8)
    tpl = rng.integers(0, 256, tstride * th, dtype=np.uint8)
    same(cvm.match_u8(img, istride, iw, ih, tpl, tstride, tw, th, cn, step, want_map=True),
         orc.match_u8(img, istride, iw, ih, tpl, tstride, tw, th, cn, step, threads=8))


@pytest.mark.parametrize("alpha", ["random", "opaque"])
def test_workload_match_640(cvm, orc, alpha):
    parent, sub, (x, y) = synth.workload_match_640(seed=1, alpha=alpha)
    say("640x480 RGBA", alpha)
    t = cvm.match(parent, sub)
    om, ot = orc.match_rgba(parent, sub, threads=8)
    assert t == ot
    assert (t[4], t[5]) == (x, y) and t[3] >= 0.999
    m, w, h = cvm.match_map(parent, sub)
    assert (w, h) == (593, 433)
    assert np.array_equal(bits(m), bits(om))


def test_workload_gray_ui_1080p(cvm, orc):
    img, sub, crop, copies = synth.workload_gray_ui_1080p(seed=2)
    say("1080p gray UI")
    t = cvm.match_gray(img, sub)
    om, ot = orc.match_gray(img, sub, threads=8)
    assert t == ot
    assert (t[4], t[5]) in [crop] + copies


@pytest.mark.parametrize("alpha", ["random", "opaque"])
def test_workload_map_4k(cvm, orc, alpha):
    parent, sub, (x, y) = synth.workload_map_4k(seed=3, alpha=alpha)
    say("4K RGBA map", alpha)
    m, w, h = cvm.match_map(parent, sub)
    assert (w, h) == (3713, 2033)
    om, ot = orc.match_rgba(parent, sub, threads=16)
    assert np.array_equal(bits(m), bits(om))
    t = cvm.match(parent, sub)
    assert t == ot and (t[4], t[5]) == (x, y)


def test_match_gray_rgba_inputs(cvm, orc):
    parent, sub, (x, y) = synth.workload_match_640(seed=6, alpha="opaque")
    g = cvm.rgba_to_gray(parent)
    assert np.array_equal(g, orc.rgba_to_gray(parent))
    t = cvm.match_gray(parent, sub)
    _, ot = orc.match_gray(orc.rgba_to_gray(parent), orc.rgba_to_gray(sub), threads=8)
    assert t == ot and (t[4], t[5]) == (x, y)


def test_sub_image_view(cvm, orc):
    rng = np.random.default_rng(9)
    parent = rng.integers(0, 256, (240, 320, 4), dtype=np.uint8)
    parent[..., 3] = 255
    region = parent[60:140, 50:130]  # row-strided view, stride 1280
    t = cvm.match(parent, region)
    assert (t[4], t[5]) == (50, 60) and t[3] >= 0.999
    _, ot = orc.match_rgba(parent, np.ascontiguousarray(region))
    assert t == ot


def test_alpha_const_and_rules(cvm, orc):
    rng = np.random.default_rng(92)
    parent = rng.integers(0, 256, (90, 120, 4), dtype=np.uint8)
    sub = parent[40:58, 60:84].copy()
    sub[..., 3] = 128  # template constant, parent varying: full 4-channel path
    assert not cvm.alpha_const(parent) and cvm.alpha_const(sub)
    t = cvm.match(parent, sub)
    _, ot = orc.match_u8(parent.ravel(), 480, 120, 90, sub.ravel(), 96, 24, 18, 4, 4)
    assert t == ot


def test_flat_template(cvm):
    parent = np.zeros((100, 100, 4), np.uint8)
    sub = np.zeros((10, 10, 4), np.uint8)
    assert cvm.match(parent, sub) == (1.0, 0, 0, 1.0, 0, 0)
    m, w, h = cvm.match_map(parent, sub)
    assert np.all(m == 1.0)


def test_errors(cvm):
    with pytest.raises(cvm.CvmatchError) as e:
        cvm.match(np.zeros((10, 10, 4), np.uint8), np.zeros((20, 20, 4), np.uint8))
    assert e.value.code == cvm.CVM_E_BAD_ARGS
    assert "cvmatch: bad match arguments (20x20 in 10x10, cn=4 step=4)" in str(e.value)
    with pytest.raises(cvm.CvmatchError) as e:
        cvm.match_u8(np.zeros(16, np.uint8), 2500 * 4, 2500, 2500, np.zeros(16, np.uint8), 2500 * 4, 2500, 2500, 4, 4)
    assert e.value.code == cvm.CVM_E_STATS_CAP
    assert "exceeds the exact-statistics bound for cn=4" in str(e.value)
    with pytest.raises(cvm.CvmatchError) as e:
        cvm.Fleet()
    assert e.value.code == cvm.CVM_E_EMPTY_FLEET


def test_matcher_bitwise(cvm, orc):
    parent, sub, (x, y) = synth.workload_match_640(seed=7, alpha="opaque")
    m = cvm.Matcher(sub)
    want = cvm.match(parent, sub)
    for _ in range(3):
        assert m.find(parent) == want
    _, ot = orc.matcher_find(parent, sub, threads=8)
    assert want == ot
    # geometry change rebuilds the cached spectrum
    small = parent[:300, :400]
    assert m.find(small) == cvm.match(small, sub)
    assert m.find(parent) == want
    g = cvm.Matcher(sub, gray=True)
    assert g.find(parent) == cvm.match_gray(parent, sub)


def test_matcher_varying_parent_alpha(cvm):
    parent, sub, _ = synth.workload_match_640(seed=8, alpha="opaque")
    m = cvm.Matcher(sub)
    a = m.find(parent)
    parent2 = parent.copy()
    parent2[5, 5, 3] = 7  # the parent's alpha now varies: cn must become 4
    assert m.find(parent2) == cvm.match(parent2, sub)
    assert m.find(parent) == a


def test_fleet_bitwise(cvm, orc):
    rng = np.random.default_rng(11)
    desk = synth.flat_rect_scene(800, 500, rng, channels=4)
    desk[..., 3] = 255
    desk = synth.add_noise(desk, 3.0, rng, channels=[0, 1, 2])
    sizes = [(24, 24), (96, 32), (48, 48), (300, 200), (16, 16), (64, 128), (33, 17)]
    subs = []
    for i in range(10):
        w, h = sizes[i % len(sizes)]
        x, y = int(rng.integers(0, 800 - w)), int(rng.integers(0, 500 - h))
        subs.append(desk[y:y + h, x:x + w].copy())
    subs.append(np.full((12, 12, 4), 9, np.uint8))  # flat member
    f = cvm.Fleet(*[cvm.Matcher(s) for s in subs])
    got = [tuple(r) for r in f.find_all(desk)]
    want = orc.fleet_find_all(desk, subs)
    assert got == want
    assert [tuple(r) for r in f.find_all(desk)] == want
    # gray fleet
    g = cvm.rgba_to_gray(desk)
    gsubs = [cvm.rgba_to_gray(s) for s in subs]
    fg = cvm.Fleet(*[cvm.Matcher(s, gray=True) for s in gsubs])
    assert [tuple(r) for r in fg.find_all(g)] == orc.fleet_find_all(g, gsubs)


def test_fleet_mixed_normalisation_forms(cvm, orc):
    """A FindAll whose members take both normalisation forms: most are
    scored in the batched wave-form launches, a 400-wide and a 300-tall
    member take the int64 form on their own; varying-alpha members (cn = 4)
    sit beside cn = 3 ones.  Two frame sizes in turn (the batch plan is
    rebuilt for the second), each bit for bit against the oracle's FindAll."""
    rng = np.random.default_rng(31)
    big = synth.add_noise(synth.flat_rect_scene(960, 720, rng, channels=4), 3.0, rng, channels=[0, 1, 2])
    big[..., 3] = 255
    sizes = [(400, 90), (60, 300), (40, 40), (120, 64), (17, 23), (200, 150), (90, 90)]
    subs = []
    for k, (w, h) in enumerate(sizes):
        x, y = int(rng.integers(0, 960 - w)), int(rng.integers(0, 720 - h))
        s = big[y:y + h, x:x + w].copy()
        if k in (2, 5):
            s[..., 3] = rng.integers(0, 256, (h, w), dtype=np.uint8)  # varying alpha: cn = 4
        subs.append(s)
    f = cvm.Fleet(*[cvm.Matcher(s) for s in subs])
    for fw, fh in [(960, 720), (700, 520), (960, 720)]:
        frame = np.ascontiguousarray(big[:fh, :fw])
        say(f"[fleet] mixed forms on {fw}x{fh}")
        assert [tuple(r) for r in f.find_all(frame)] == orc.fleet_find_all(frame, subs)


def test_fleet_single_equals_find(cvm):
    parent, sub, _ = synth.workload_match_640(seed=12, alpha="opaque")
    m = cvm.Matcher(sub)
    assert tuple(cvm.Fleet(m).find_all(parent)[0]) == m.find(parent)


def test_device_resident_path(cvm):
    parent, sub, _ = synth.workload_match_640(seed=13)
    L = cvm.lib()
    ph, pw = parent.shape[:2]
    th, tw = sub.shape[:2]
    dp, dt, dm = C.c_void_p(), C.c_void_p(), C.c_void_p()
    rw, rh = pw - tw + 1, ph - th + 1
    assert L.cvm_device_alloc(parent.nbytes, C.byref(dp)) == 0
    assert L.cvm_device_alloc(sub.nbytes, C.byref(dt)) == 0
    assert L.cvm_device_alloc(rw * rh * 4, C.byref(dm)) == 0
    try:
        L.cvm_copy_h2d(dp, parent.ctypes.data_as(C.c_void_p), parent.nbytes)
        L.cvm_copy_h2d(dt, sub.ctypes.data_as(C.c_void_p), sub.nbytes)
        out = cvm.cvm_result()
        assert L.cvm_match_u8_device(dp, pw * 4, pw, ph, dt, tw * 4, tw, th, 0, 4, dm, C.byref(out), None) == 0
        m = np.empty((rh, rw), np.float32)
        L.cvm_copy_d2h(m.ctypes.data_as(C.c_void_p), dm, m.nbytes)
        hm, _, _ = cvm.match_map(parent, sub)
        assert out.tuple() == cvm.match(parent, sub)
        assert np.array_equal(bits(m), bits(hm))
    finally:
        for p in (dp, dt, dm):
            L.cvm_device_free(p)


@pytest.mark.parametrize("direct", [False, True])
def test_device_resident_replays(cvm, direct):
    """Repeated device-resident calls with the same arguments replay a
    captured call graph: every replay (interleaved with other calls, host
    work and synchronisation) gives the eager call's map and tuple bitwise."""
    L = cvm.lib()
    fn = L.cvm_match_u8_direct_device if direct else L.cvm_match_u8_device
    rng = np.random.default_rng(77)
    parent = synth.box_blur5(rng.integers(0, 256, size=(360, 512, 4), dtype=np.uint8))
    subs = [parent[40:88, 100:164].copy(), parent[200:231, 300:317].copy()]
    ph, pw = parent.shape[:2]
    bufs = []

    def dalloc(n):
        p = C.c_void_p()
        assert L.cvm_device_alloc(n, C.byref(p)) == 0
        bufs.append(p)
        return p

    try:
        dp = dalloc(parent.nbytes)
        L.cvm_copy_h2d(dp, parent.ctypes.data_as(C.c_void_p), parent.nbytes)
        cases = []
        for sub in subs:
            th, tw = sub.shape[:2]
            dt = dalloc(sub.nbytes)
            L.cvm_copy_h2d(dt, sub.ctypes.data_as(C.c_void_p), sub.nbytes)
            rw, rh = pw - tw + 1, ph - th + 1
            dm = dalloc(rw * rh * 4)
            if direct:
                want_m, want_t = cvm.match_direct(parent, sub, want_map=True)
            else:
                want_m, _, _ = cvm.match_map(parent, sub)
                want_t = cvm.match(parent, sub)
            cases.append((dt, dm, tw, th, rw, rh, want_m, tuple(want_t)))
        # runs of repeats: the eager call, the capture, then replays; then
        # the other template; then back
        order = [0] * 7 + [1] * 6 + [0] * 6 + [1] * 2 + [0]
        for k, ci in enumerate(order):
            dt, dm, tw, th, rw, rh, want_m, want_t = cases[ci]
            out = cvm.cvm_result()
            assert fn(dp, pw * 4, pw, ph, dt, tw * 4, tw, th, 0, 4, dm, C.byref(out), None) == 0
            m = np.empty((rh, rw), np.float32)
            L.cvm_copy_d2h(m.ctypes.data_as(C.c_void_p), dm, m.nbytes)
            assert out.tuple() == want_t, (k, ci)
            assert np.array_equal(bits(m), bits(want_m)), (k, ci)
            if k in (5, 16):
                cvm.match(parent, subs[1])  # other eager work between replays
    finally:
        for p in bufs:
            L.cvm_device_free(p)


def test_repeat_determinism(cvm):
    parent, sub, _ = synth.workload_match_640(seed=14)
    a, _, _ = cvm.match_map(parent, sub)
    for _ in range(3):
        b, _, _ = cvm.match_map(parent, sub)
        assert np.array_equal(bits(a), bits(b))


@pytest.mark.parametrize("w,h", [(37, 21), (64, 40), (1001, 13)])
def test_alpha_scan_positions(cvm, orc, w, h):
    base = np.zeros((h, w, 4), np.uint8)
    base[..., 3] = 77
    assert cvm.alpha_const(base) and orc.alpha_const(base)
    for (y, x) in [(0, 0), (0, w - 1), (7, 3), (8, 0), (h - 1, w - 1), (h // 2, w // 2)]:
        img = base.copy()
        img[y, x, 3] = 78
        assert cvm.alpha_const(img) is False and orc.alpha_const(img) is False, (y, x)
        # unaligned row stride: a view with one extra pixel per row
        padded = np.zeros((h, w + 1, 4), np.uint8)
        padded[:, :w] = img
        assert cvm.alpha_const(padded[:, :w]) is False


def _random_cases_large(n=12, seed=7):
    """Templates past the wave-form normalisation (th > 257, or wide enough to
    take the int64 form), odd strides, thin images."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        cn, step = [(1, 1), (4, 4), (3, 4), (2, 2)][i % 4]
        iw, ih = int(rng.integers(300, 1400)), int(rng.integers(270, 900))
        tw = int(rng.integers(150, min(iw, 700) + 1))
        th = int(rng.integers(100, min(ih, 420) + 1))
        if i % 3 == 0:
            iw, tw = int(rng.integers(2, 40)), 1  # thin column images
        out.append((100 + i, iw, ih, tw, th, cn, step, int(rng.integers(0, 7)), int(rng.integers(0, 5))))
    return out


@pytest.mark.parametrize("case", _random_cases_large(), ids=lambda c: f"L{c[0]}")
def test_random_large_templates_bitwise(cvm, orc, case):
    i, iw, ih, tw, th, cn, step, ipad, tpad = case
    if tw * th > orc.stats_cap(cn):
        pytest.skip("template above the exact-statistics bound (rejected by both)")
    say("random large template", case)
    rng = np.random.default_rng(2000 + i)
    istride, tstride = (iw + ipad) * step, (tw + tpad) * step
    img = rng.integers(0, 256, istride * ih, dtype=np.uint8)
    tpl = rng.integers(0, 256, tstride * th, dtype=np.uint8)
    same(cvm.match_u8(img, istride, iw, ih, tpl, tstride, tw, th, cn, step, want_map=True),
         orc.match_u8(img, istride, iw, ih, tpl, tstride, tw, th, cn, step, threads=8))


@pytest.mark.parametrize("seed", [21, 22, 23, 24, 25, 26, 27, 28])
def test_random_fleets_bitwise(cvm, orc, seed):
    """Random fleets against the oracle's FindAll: 6-12 members of random
    sizes (1 to 220 per side), some with varying alpha (their own cn = 4 on a
    constant-alpha frame), some flat; the same members as a gray fleet."""
    rng = np.random.default_rng(seed)
    fw, fh = int(rng.integers(300, 900)), int(rng.integers(200, 600))
    frame = synth.add_noise(synth.flat_rect_scene(fw, fh, rng, channels=4), 4.0, rng, channels=[0, 1, 2])
    frame[..., 3] = 255
    subs = []
    for k in range(int(rng.integers(6, 13))):
        w, h = int(rng.integers(1, min(fw, 220) + 1)), int(rng.integers(1, min(fh, 220) + 1))
        x, y = int(rng.integers(0, fw - w + 1)), int(rng.integers(0, fh - h + 1))
        s = frame[y:y + h, x:x + w].copy()
        if k % 4 == 1:
            s[..., 3] = rng.integers(0, 256, (h, w), dtype=np.uint8)  # varying alpha
        if k % 5 == 4:
            s[...] = 77  # flat
        subs.append(s)
    say(f"[fleet] seed {seed}: frame {fw}x{fh}, {len(subs)} members")
    got = [tuple(r) for r in cvm.Fleet(*[cvm.Matcher(s) for s in subs]).find_all(frame)]
    assert got == orc.fleet_find_all(frame, subs)
    g = cvm.rgba_to_gray(frame)
    gsubs = [cvm.rgba_to_gray(s) for s in subs]
    assert [tuple(r) for r in cvm.Fleet(*[cvm.Matcher(s, gray=True) for s in gsubs]).find_all(g)] == \
        orc.fleet_find_all(g, gsubs)


def test_8k_map_bitwise(cvm, orc):
    """A 7680x4320 RGBA MatchMap (4x the headline's pixels, 3x8+ tiles, a
    133 MB parent): the whole map bit-identical to the oracle's."""
    rng = np.random.default_rng(88)
    parent = synth.box_blur5(rng.integers(0, 256, size=(4320, 7680, 4), dtype=np.uint8))
    sub = synth.add_noise(parent[2000:2048, 5000:5064].copy(), 2.0, rng, channels=[0, 1, 2])
    say("[8k] MatchMap 7680x4320 RGBA vs 64x48")
    m, _, _ = cvm.match_map(parent, sub)
    om, ot = orc.match_rgba(parent, sub, threads=16)
    assert np.array_equal(m.view(np.uint32), om.view(np.uint32))
    assert cvm.match(parent, sub) == ot
