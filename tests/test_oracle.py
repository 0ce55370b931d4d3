"""Pins of the CPU oracle (oracle/cvm_oracle.c) -- CPU only.

What fixes each part of the oracle:
  planner          the reference's 15 planner pins (plan_test.go:12-29), exact,
                   in test_golden.py from the committed fixture
  whole pipeline   golden hash #110 (impl_golden_test.go:70-86 input,
                   golden_hashes_test.go:21 value), exact bits of the reference
  stats caps       closed form (statscap_test.go:16-35)
  twiddles         special values + <= 1 ulp of float64 cos/sin, exact w3 fold
  FFT engines      naive/f64 DFT tolerance (radix4_test.go:40-62), round trip
                   (:64-81), columns == rows bitwise (radix4cols_test.go:16-68)
  normalisation    an independent float64 numpy transliteration of refMatch
                   (cvmatch_test.go:13-81), <= 1e-4 on its 7 shapes
  plumbing         gray formula, alpha predicate, cn3 == cn4 on constant alpha
  Matcher / Fleet  Find == Match bitwise, single-member fleet == Find bitwise,
                   bucketize cases (fleet_internal_test.go:27-51)
  FP64 FFT route   fp64_map_fft == the brute-force FP64 map bit for bit
Inputs come from numpy's PCG64, not Go's math/rand, so golden hashes
#101-#109 cannot be reproduced here (their inputs need Go's generator).
"""
import numpy as np
import pytest
from numpy.lib.stride_tricks import sliding_window_view

GOLDEN_110 = 0x99534488138E0635


def canary_110():
    iw, ih, tw, th = 200, 150, 32, 32
    img = np.full(iw * ih, 128, np.uint8)
    tpl = np.full(tw * th, 128, np.uint8)
    img[75 * iw + 100] = 129
    tpl[5 * tw + 7] = 129
    return img, tpl, iw, ih, tw, th


@pytest.mark.parametrize("threads", [1, 4])
def test_golden_110(orc, threads):
    img, tpl, iw, ih, tw, th = canary_110()
    m, t = orc.match_u8(img, iw, iw, ih, tpl, tw, tw, th, 1, 1, threads)
    assert orc.golden_hash(m, t) == GOLDEN_110


def test_golden_110_detects_a_changed_bit(orc):
    img, tpl, iw, ih, tw, th = canary_110()
    m, t = orc.match_u8(img, iw, iw, ih, tpl, tw, tw, th, 1, 1, 1)
    m = m.copy()
    m.view(np.uint32)[3, 3] ^= 0x80000000  # +0 -> -0: one flipped bit must move the hash
    assert orc.golden_hash(m, t) != GOLDEN_110


def test_stats_cap_closed_form(orc):
    limit = (1 << 63) - 1
    for cn in range(1, 5):
        a = orc.stats_cap(cn)
        v = cn * 65025
        assert a * a <= limit // v
        assert (a + 1) * (a + 1) > limit // v
    assert 3000 * 2000 <= orc.stats_cap(1)


def test_stats_cap_rejects(orc):
    with pytest.raises(orc.OracleError) as e:
        orc.match_u8(np.zeros(16, np.uint8), 2500 * 4, 2500, 2500, np.zeros(16, np.uint8), 2500 * 4, 2500, 2500, 4, 4)
    assert e.value.code == 2


def test_bitrev_involution(orc):
    for lg in range(0, 13):
        n = 1 << lg
        b = orc.bitrev(n)
        assert np.array_equal(b[b], np.arange(n))
        want = [int(format(i, f"0{lg}b")[::-1], 2) if lg else 0 for i in range(n)]
        assert b.tolist() == want


def test_twiddle_special_values(orc):
    c, s = orc.sincospi(0, 8)
    assert (c, s) == (1.0, 0.0) and not np.signbit(s)
    c, s = orc.sincospi(4, 8)  # pi/2: the quarter rotation yields cos = -0 exactly
    assert c == 0.0 and np.signbit(c) and s == 1.0
    c, s = orc.sincospi(2, 8)  # pi/4
    assert c == s == np.float32(np.sqrt(0.5))


def test_twiddles_within_one_ulp(orc):
    worst = 0
    for half in [1, 2, 4, 8, 64, 512, 4096]:
        for j in range(half):
            c, s = orc.sincospi(j, half)
            ang = np.pi * j / half
            for got, want in ((c, np.cos(ang)), (s, np.sin(ang))):
                if abs(want) < 1e-12:  # float64 rounding of an exact zero (cos(pi/2))
                    assert got == 0.0
                    continue
                w32 = np.float32(want)
                if got != w32:
                    ulp = np.spacing(np.float32(abs(want)))
                    worst = max(worst, abs(float(got) - want) / ulp)
    assert worst <= 1.0


def test_tri_twiddles_layout(orc):
    for n in [4, 8, 16, 32, 1024, 2048]:
        tri = orc.tri_twiddles(n)
        h = 2 if (n.bit_length() % 2 == 0) else 1
        at = 0
        while 4 * h <= n:
            j = np.arange(h)
            for blk, ang in enumerate((np.pi * j / (2 * h), np.pi * j / h, 3 * np.pi * j / (2 * h))):
                got = tri[at + blk * h:at + (blk + 1) * h]
                want = np.exp(-1j * ang)
                assert np.max(np.abs(got - want)) < 2e-7, (n, h, blk)
            at += 3 * h
            h *= 4
        assert at == len(tri)


def _dft64(x, inverse):
    n = x.size
    if n <= 512:
        k = np.arange(n)
        idx = np.outer(k, k) % n
        sign = 1 if inverse else -1
        return (np.exp(sign * 2j * np.pi * idx / n) * x.astype(np.complex128)[None, :]).sum(axis=1)
    xx = x.astype(np.complex128)
    return np.fft.ifft(xx) * n if inverse else np.fft.fft(xx)


@pytest.mark.parametrize("inverse", [False, True])
def test_fft_against_dft(orc, inverse):
    for lg in range(0, 13):
        n = 1 << lg
        rng = np.random.default_rng(n + 7)
        x = (rng.random(n, dtype=np.float32) * 2 - 1) + 1j * (rng.random(n, dtype=np.float32) * 2 - 1)
        x = x.astype(np.complex64)
        got = orc.fft_r4(x, inverse)
        want = _dft64(x, inverse)
        tol = 3e-7 * np.sqrt(np.sum(np.abs(want) ** 2)) * np.sqrt(np.log2(n) + 1) * 4
        assert np.max(np.abs(got.astype(np.complex128) - want)) <= tol, n


def test_fft_round_trip(orc):
    for n in [2, 4, 8, 64, 512, 1024]:
        rng = np.random.default_rng(n + 3)
        x = ((rng.random(n) * 2 - 1) + 1j * (rng.random(n) * 2 - 1)).astype(np.complex64)
        a = orc.fft_r4(orc.fft_r4(x, False), True)
        a = (a.real * np.float32(1 / n)).astype(np.float32) + 1j * (a.imag * np.float32(1 / n)).astype(np.float32)
        assert np.max(np.abs(a - x)) <= 2e-5


@pytest.mark.parametrize("n", [2, 4, 8, 16, 64, 128, 512])
def test_cols_equal_rows_bitwise(orc, n):
    brev = orc.bitrev(n)
    for width in [1, 7, 24, 33]:
        rng = np.random.default_rng(n * width)
        src = ((rng.random((n, width)) * 2 - 1) + 1j * (rng.random((n, width)) * 2 - 1)).astype(np.complex64)
        for inverse in (False, True):
            want = np.stack([orc.fft_r4(src[:, c], inverse) for c in range(width)], axis=1)
            placed = np.empty_like(src)
            placed[brev] = src
            got = orc.cols_r4(placed, inverse, use_rmap=False)
            assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
            got2 = orc.cols_r4(src, inverse, use_rmap=True)
            assert np.array_equal(got2[brev].view(np.uint32), want.view(np.uint32))


# ------------------------------------------------------------ whole path


def ref_match_np(img, istride, iw, ih, tpl, tstride, tw, th, cn, step):
    """cvmatch_test.go:13-81 transliterated to float64 numpy."""
    I = np.lib.stride_tricks.as_strided(img, shape=(ih, iw, step), strides=(istride, step, 1))[..., :cn]
    T = np.lib.stride_tricks.as_strided(tpl, shape=(th, tw, step), strides=(tstride, step, 1))[..., :cn]
    I = I.astype(np.float64)
    T = T.astype(np.float64)
    area = float(tw * th)
    mean = T.sum(axis=(0, 1)) / area
    templ = float(np.sum((T * T).sum(axis=(0, 1)) / area - mean * mean))
    rw, rh = iw - tw + 1, ih - th + 1
    if templ < 2.220446049250313e-16:
        return np.ones((rh, rw))
    templ = np.sqrt(templ * area)
    win = sliding_window_view(I, (th, tw), axis=(0, 1))  # rh, rw, cn, th, tw
    corr = np.einsum("yxkij,ijk->yx", win, T)
    wsum = win.sum(axis=(3, 4))  # rh, rw, cn
    wsum2 = (win * win).sum(axis=(2, 3, 4))
    num = corr - (wsum * mean).sum(axis=2)
    wmean2 = (wsum * wsum).sum(axis=2) / area
    diff2 = np.maximum(wsum2 - wmean2, 0)
    lim = np.minimum(0.5, 10 * 1.1920929e-07 * wsum2)
    den = np.where(diff2 > lim, np.sqrt(diff2) * templ, 0.0)
    an = np.abs(num)
    return np.where(an < den, num / np.where(den > 0, den, 1), np.where(an < den * 1.125, np.sign(num) + (num == 0) * -1, 0.0))


REF_SHAPES = [  # cvmatch_test.go:93-101
    (97, 61, 17, 13, 1, 0, 0),
    (97, 61, 17, 13, 4, 0, 0),
    (64, 64, 64, 64, 1, 0, 0),
    (50, 40, 1, 1, 4, 0, 0),
    (300, 200, 31, 29, 1, 7, 3),
    (40, 300, 33, 257, 1, 0, 0),
    (128, 96, 5, 4, 3, 0, 0),
]


@pytest.mark.parametrize("case", REF_SHAPES)
def test_against_float64_reference(orc, case):
    iw, ih, tw, th, cn, ipad, tpad = case
    rng = np.random.default_rng(42 + iw + th)
    istride, tstride = (iw + ipad) * cn, (tw + tpad) * cn
    img = rng.integers(0, 256, istride * ih, dtype=np.uint8)
    tpl = rng.integers(0, 256, tstride * th, dtype=np.uint8)
    got, tup = orc.match_u8(img, istride, iw, ih, tpl, tstride, tw, th, cn, cn, 4)
    want = ref_match_np(img, istride, iw, ih, tpl, tstride, tw, th, cn, cn)
    assert np.max(np.abs(got.astype(np.float64) - want)) <= 1e-4
    # the exact-integer FP64 oracle agrees with the transliteration
    fp = orc.fp64_map(img, istride, iw, ih, tpl, tstride, tw, th, cn, cn)
    assert np.max(np.abs(fp - want)) <= 1e-9
    # first-occurrence scan of the map itself
    flat = got.ravel()
    imin, imax = int(np.argmin(flat)), int(np.argmax(flat))
    rw = iw - tw + 1
    assert (tup[0], tup[1] + tup[2] * rw) == (flat[imin], imin)
    assert (tup[3], tup[4] + tup[5] * rw) == (flat[imax], imax)


def test_threads_bit_identical(orc):
    rng = np.random.default_rng(77)
    iw, ih, tw, th = 700, 500, 96, 64
    img = rng.integers(0, 256, iw * ih * 4, dtype=np.uint8)
    tpl = rng.integers(0, 256, tw * th * 4, dtype=np.uint8)
    ref, rt = orc.match_u8(img, iw * 4, iw, ih, tpl, tw * 4, tw, th, 3, 4, 1)
    for n in (2, 3, 8):
        got, t = orc.match_u8(img, iw * 4, iw, ih, tpl, tw * 4, tw, th, 3, 4, n)
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)) and t == rt


def test_gray_formula(orc):
    rng = np.random.default_rng(5)
    pix = rng.integers(0, 256, (37, 53, 4), dtype=np.uint8)
    r, g, b = (pix[..., k].astype(np.uint32) for k in range(3))
    want = ((4899 * r + 9617 * g + 1868 * b + 8192) >> 14).astype(np.uint8)
    assert np.array_equal(orc.rgba_to_gray(pix), want)
    # the weights sum to 2^14: white stays 255, black stays 0
    assert orc.rgba_to_gray(np.full((1, 1, 4), 255, np.uint8))[0, 0] == 255


def test_alpha_const(orc):
    pix = np.zeros((9, 11, 4), np.uint8)
    pix[..., 3] = 200
    assert orc.alpha_const(pix)
    pix[8, 10, 3] = 199
    assert not orc.alpha_const(pix)


def test_constant_alpha_skip(orc):
    rng = np.random.default_rng(21)
    iw, ih, tw, th = 120, 90, 24, 18
    img = rng.integers(0, 256, iw * ih * 4, dtype=np.uint8)
    tpl = rng.integers(0, 256, tw * th * 4, dtype=np.uint8)
    for a, b in [(200, 55), (0, 255), (255, 0), (128, 128)]:
        img[3::4] = a
        tpl[3::4] = b
        g3, _ = orc.match_u8(img, iw * 4, iw, ih, tpl, tw * 4, tw, th, 3, 4, 4)
        g4, _ = orc.match_u8(img, iw * 4, iw, ih, tpl, tw * 4, tw, th, 4, 4, 4)
        assert np.max(np.abs(g3 - g4)) <= 1e-5


def test_alpha_matters(orc):
    rng = np.random.default_rng(91)
    iw, ih, tw, th = 160, 120, 32, 24
    img = rng.integers(0, 256, iw * ih * 4, dtype=np.uint8)
    tpl = rng.integers(0, 256, tw * th * 4, dtype=np.uint8)
    g3, _ = orc.match_u8(img, iw * 4, iw, ih, tpl, tw * 4, tw, th, 3, 4, 2)
    g4, _ = orc.match_u8(img, iw * 4, iw, ih, tpl, tw * 4, tw, th, 4, 4, 2)
    assert np.max(np.abs(g3 - g4)) >= 1e-3


def test_planted_and_flat(orc):
    rng = np.random.default_rng(7)
    parent = rng.integers(0, 256, (400, 640, 4), dtype=np.uint8)
    parent[..., 3] = 255
    sub = parent[77:77 + 48, 123:123 + 64].copy()
    _, t = orc.match_rgba(parent, sub)
    assert (t[4], t[5]) == (123, 77) and t[3] >= 0.999
    _, t = orc.match_gray(orc.rgba_to_gray(parent), orc.rgba_to_gray(sub))
    assert (t[4], t[5]) == (123, 77) and t[3] >= 0.999
    m, t = orc.match_rgba(np.zeros((100, 100, 4), np.uint8), np.zeros((10, 10, 4), np.uint8))
    assert t == (1.0, 0, 0, 1.0, 0, 0) and np.all(m == 1)


def test_find_all_occurrences(orc):
    rng = np.random.default_rng(101)
    parent = rng.integers(0, 64, (300, 400, 4), dtype=np.uint8)
    parent[..., 3] = 255
    tpl = rng.integers(0, 256, (24, 32, 4), dtype=np.uint8)
    tpl[..., 3] = 255
    spots = [(10, 12), (200, 40), (350, 260), (60, 240), (180, 150)]
    for x, y in spots:
        parent[y:y + 24, x:x + 32] = tpl
    resp, _ = orc.match_rgba(parent, tpl)
    hits = []
    for y, x in zip(*np.nonzero(resp >= 0.999)):
        nb = resp[max(0, y - 12):y + 13, max(0, x - 16):x + 17]
        if not np.any(nb > resp[y, x]):
            hits.append((int(x), int(y)))
    assert sorted(hits) == sorted(spots)
    # the oracle's restatement of the same loop: raster order, same hits
    n, xy, vals = orc.find_peaks(resp, 0.999, 16, 12)
    assert n == 5 and [tuple(p) for p in xy] == sorted(spots, key=lambda p: (p[1], p[0]))
    assert np.array_equal(vals, resp[xy[:, 1], xy[:, 0]])


@pytest.mark.parametrize("rx,ry", [(0, 0), (1, 0), (0, 2), (3, 2), (7, 11), (40, 3), (80, 64)])
def test_find_peaks_against_max_filter(orc, rx, ry):
    """orc_find_peaks against an independent formulation: a peak is a point
    >= thresh that equals the maximum of its clipped window (scipy's
    maximum_filter with -inf outside the map).  Quantised values make ties
    (both tied points stay peaks: suppression needs a strictly greater
    neighbour); the truncated call keeps the raster-order prefix."""
    from scipy.ndimage import maximum_filter

    rng = np.random.default_rng(rx * 31 + ry)
    for h, w, levels in [(37, 53, 5), (1, 70, 3), (64, 1, 4), (90, 120, 50)]:
        resp = (rng.integers(0, levels, (h, w)).astype(np.float32) / (levels - 1) * 2 - 1).astype(np.float32)
        mf = maximum_filter(resp, size=(2 * ry + 1, 2 * rx + 1), mode="constant", cval=-np.inf)
        for thresh in (-2.0, 0.0, 0.5, 1.0):
            want = np.argwhere((resp >= thresh) & (resp >= mf))  # (y, x), raster order
            n, xy, vals = orc.find_peaks(resp, thresh, rx, ry)
            assert n == len(want), (h, w, thresh)
            assert np.array_equal(xy[:, ::-1], want), (h, w, thresh)
            assert np.array_equal(vals, resp[want[:, 0], want[:, 1]])
            if n > 2:
                k, xy2, _ = orc.find_peaks(resp, thresh, rx, ry, cap=n // 2)
                assert k == n and np.array_equal(xy2, xy[:n // 2])


def test_matcher_find_equals_match(orc):
    rng = np.random.default_rng(12)
    parent = rng.integers(0, 256, (300, 500, 4), dtype=np.uint8)
    parent[..., 3] = 255
    sub = parent[40:100, 70:150].copy()
    want_map, want = orc.match_rgba(parent, sub, want_map=True)
    got_map, got = orc.matcher_find(parent, sub, want_map=True)
    assert got == want and np.array_equal(got_map.view(np.uint32), want_map.view(np.uint32))


def test_bucketize(orc):
    bs = orc.bucketize([(300, 8), (8, 300)])
    assert len(bs) == 2
    for tw, th, _ in bs:
        assert tw <= 300 and th <= 300 and not (tw == 300 and th == 300)
    bs = orc.bucketize([(200, 100), (24, 24)])
    assert len(bs) == 1 and bs[0][:2] == (200, 100)
    bs = orc.bucketize([(400, 300), (24, 24)], flat=[1, 0])
    assert len(bs) == 1 and bs[0][:2] == (24, 24)


def _fleet_scene(seed=11, n=12):
    rng = np.random.default_rng(seed)
    from cvmatch_amd import synth

    desk = synth.flat_rect_scene(800, 500, rng, channels=4)
    desk[..., 3] = 255
    desk = synth.add_noise(desk, 3.0, rng, channels=[0, 1, 2])
    sizes = [(24, 24), (96, 32), (48, 48), (300, 200), (16, 16), (64, 128)]
    subs, at = [], []
    for i in range(n):
        w, h = sizes[i % len(sizes)]
        x, y = int(rng.integers(0, 800 - w)), int(rng.integers(0, 500 - h))
        subs.append(desk[y:y + h, x:x + w].copy())
        at.append((x, y))
    return desk, subs, at


def test_fleet_single_member_bit_identical(orc):
    desk, subs, _ = _fleet_scene(n=3)
    for s in subs:
        _, want = orc.matcher_find(desk, s)
        assert orc.fleet_find_all(desk, [s])[0] == want


def test_fleet_matches_solo(orc):
    desk, subs, at = _fleet_scene()
    got = orc.fleet_find_all(desk, subs)
    hits = 0
    for i, r in enumerate(got):
        _, solo = orc.matcher_find(desk, subs[i])
        assert (r[4], r[5]) == (solo[4], solo[5])
        assert abs(r[3] - solo[3]) <= 1e-3
        hits += (r[4], r[5]) == at[i]
    assert hits >= len(got) * 2 // 3
    # order independence: reversing the members reverses the answers
    rev = orc.fleet_find_all(desk, subs[::-1])
    assert rev[::-1] == got


@pytest.mark.parametrize("case", [
    # (iw, ih, tw, th, cn, planted)
    (97, 61, 17, 13, 1, False),
    (80, 50, 9, 31, 3, True),
    (64, 48, 16, 16, 4, True),
    (40, 30, 40, 30, 4, False),
    (33, 20, 1, 1, 1, False),
    (70, 45, 12, 10, 3, False),
])
def test_fp64_map_fft_equals_brute_force(orc, case):
    """The FFT-route FP64 map equals the brute-force FP64 map bit for bit:
    random bytes, a planted exact crop (the +-1 band), flat regions in the
    parent (den = 0) and a flat template (all ones)."""
    iw, ih, tw, th, cn, planted = case
    rng = np.random.default_rng(iw * 3 + th)
    ch = 1 if cn == 1 else 4
    parent = rng.integers(0, 256, (ih, iw, ch), dtype=np.uint8)
    parent[: ih // 3, : iw // 2] = 77  # flat block: windows inside it have den = 0
    if ch == 4 and cn == 3:
        parent[..., 3] = 255
    sub = rng.integers(0, 256, (th, tw, ch), dtype=np.uint8)
    if planted:
        sub = parent[ih - th:, iw - tw:].copy()
    if ch == 4 and cn == 3:
        sub[..., 3] = 255
    p2 = parent[..., 0] if ch == 1 else parent
    s2 = sub[..., 0] if ch == 1 else sub
    want = orc.fp64_map(np.ascontiguousarray(p2), iw * ch, iw, ih, np.ascontiguousarray(s2), tw * ch, tw, th, cn, ch)
    got = orc.fp64_map_fft(p2, s2, cn)
    assert got.shape == want.shape
    assert np.array_equal(got.view(np.uint64), want.view(np.uint64))
    flat = np.full_like(s2, 9)
    assert np.all(orc.fp64_map_fft(p2, flat, cn) == 1.0)


def test_fp64_map_fft_at_frame_scale(orc):
    """A 2560x1440 RGBA frame against a 128x128 template (products near
    2^31): the FFT route still rounds to the exact integers (its residue
    assertion) and matches the brute-force oracle at seeded windows."""
    rng = np.random.default_rng(77)
    parent = rng.integers(200, 256, (1440, 2560, 4), dtype=np.uint8)
    sub = rng.integers(200, 256, (128, 128, 4), dtype=np.uint8)
    got = orc.fp64_map_fft(parent, sub, 4)
    xs = rng.integers(0, got.shape[1], 64).astype(np.int32)
    ys = rng.integers(0, got.shape[0], 64).astype(np.int32)
    want = orc.fp64_points(parent, 2560 * 4, 2560, 1440, sub, 128 * 4, 128, 128, 4, 4, xs, ys)
    assert np.array_equal(got[ys, xs].view(np.uint64), want.view(np.uint64))


def test_guard_shortcut_identity():
    """The FP64 scoring skips the guard's division fl(idiff / area) > lim
    (lim = min(0.5, 10 * 2^-23 * s2)) when idiff > area (always passes) or
    idiff == 0 (always fails); the identity, on exact integers in float64."""
    rng = np.random.default_rng(17)
    area = rng.integers(1, 131072, 200000).astype(np.float64)
    s2 = rng.integers(0, 1 << 40, 200000).astype(np.float64)
    pick = rng.integers(0, 4, 200000)
    idiff = np.where(pick == 0, 0.0, np.where(pick == 1, area, np.where(
        pick == 2, np.floor(area * rng.random(200000)), rng.integers(0, 1 << 44, 200000).astype(np.float64))))
    s2[:1000] = 0.0
    lim = np.minimum(0.5, 1.1920928955078125e-06 * s2)
    full = idiff / area > lim
    short = (idiff > area) | ((idiff != 0) & (idiff / area > lim))
    assert np.array_equal(full, short)
