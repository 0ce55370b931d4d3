"""GPU parity at the edges of the shapes the reference accepts: templates
wider than any strip of 2048 columns (the plain path takes every template
up to the widest row transform, 4096 columns), and the statsCap(3) bound
that Match applies once both alpha planes are constant (cvmatch.go:82-86
then impl.go:1131), bit for bit against the oracle, and the errors the
reference raises when the alpha rule leaves cn = 4 or the area exceeds even
statsCap(3)."""
import ctypes as C
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


def rgb_const_alpha(rng, h, w):
    a = rng.integers(0, 256, (h, w, 4), dtype=np.uint8)
    a[..., 3] = 255
    return a


def test_wide_template_rgba_bitwise(cvm, orc):
    """1900x120 template on a 3840x400 RGBA parent (random alpha: cn = 4)."""
    rng = np.random.default_rng(501)
    parent = rng.integers(0, 256, (400, 3840, 4), dtype=np.uint8)
    sub = parent[150:270, 1000:2900].copy()
    say("[shapes] 1900x120 RGBA template in 3840x400")
    m, w, h = cvm.match_map(parent, sub)
    om, ot = orc.match_rgba(parent, sub, threads=8)
    assert (w, h) == (1941, 281)
    assert np.array_equal(bits(m), bits(om))
    assert cvm.match(parent, sub) == ot
    assert ot[4:] == (1000, 150)


def test_wide_template_gray_bitwise(cvm, orc):
    """3000x40 gray template on 3600x160 (a 4096-point row transform)."""
    rng = np.random.default_rng(502)
    parent = rng.integers(0, 256, (160, 3600), dtype=np.uint8)
    sub = parent[70:110, 300:3300].copy()
    say("[shapes] 3000x40 gray template in 3600x160")
    gm, gt = cvm.match_u8(parent, 3600, 3600, 160, sub, 3000, 3000, 40, 1, 1, want_map=True)
    om, ot = orc.match_u8(parent, 3600, 3600, 160, sub, 3000, 3000, 40, 1, 1, threads=8)
    assert gt == ot
    assert np.array_equal(bits(gm), bits(om))
    assert cvm.match_gray(parent, sub) == ot


def test_alpha_rule_statscap3_bitwise(cvm, orc):
    """A constant-alpha 1790x3400 template (6.09 M pixels: above statsCap(4),
    within statsCap(3)) on a constant-alpha 1900x3500 parent: Match and
    MatchMap take cn = 3 and answer as the reference does; the device-resident
    call with the alpha rule (cn = 0) too."""
    rng = np.random.default_rng(503)
    parent = rgb_const_alpha(rng, 3500, 1900)
    sub = parent[40:3440, 60:1850].copy()
    assert 1790 * 3400 > cvm.stats_cap(4) and 1790 * 3400 <= cvm.stats_cap(3)
    say("[shapes] constant-alpha 1790x3400 template in 1900x3500")
    m, w, h = cvm.match_map(parent, sub)
    om, ot = orc.match_rgba(parent, sub, threads=8)
    assert (w, h) == (111, 101)
    assert np.array_equal(bits(m), bits(om))
    assert cvm.match(parent, sub) == ot
    assert ot[4:] == (60, 40)
    L = cvm.lib()
    bufs = []

    def dev(a):
        p = C.c_void_p()
        assert L.cvm_device_alloc(max(a.nbytes, 4), C.byref(p)) == 0
        assert L.cvm_copy_h2d(p, C.c_void_p(a.ctypes.data), a.nbytes) == 0
        bufs.append(p)
        return p

    try:
        dp, ds, dm = dev(parent), dev(sub), dev(np.zeros(w * h, np.float32))
        out = cvm.cvm_result()
        for k in range(3):  # eager, capture, replay
            rc = L.cvm_match_u8_device(dp, 1900 * 4, 1900, 3500, ds, 1790 * 4, 1790, 3400, 0, 4, dm, C.byref(out),
                                       None)
            assert rc == 0, L.cvm_last_error_detail()
            got = np.empty(w * h, np.float32)
            assert L.cvm_copy_d2h(C.c_void_p(got.ctypes.data), dm, got.nbytes) == 0
            assert np.array_equal(got.view(np.uint32), bits(om).ravel()), k
            assert out.tuple() == ot, k
    finally:
        for p in bufs:
            L.cvm_device_free(p)


def test_alpha_rule_statscap_errors(cvm):
    """The same area with one varying alpha byte: the rule picks cn = 4 and
    the reference panics with its cn = 4 bound; above statsCap(3) a
    constant-alpha pair fails with the cn = 3 bound."""
    rng = np.random.default_rng(504)
    parent = rgb_const_alpha(rng, 3500, 1900)
    sub = parent[40:3440, 60:1850].copy()
    parent[3000, 1000, 3] = 7
    with pytest.raises(cvm.CvmatchError) as e:
        cvm.match(parent, sub)
    assert e.value.code == cvm.CVM_E_STATS_CAP
    assert "cvmatch: template area 1790x3400 exceeds the exact-statistics bound for cn=4" in str(e.value)
    big = rgb_const_alpha(rng, 2650, 2700)
    bsub = big[:2600, :2700].copy()
    assert 2700 * 2600 > cvm.stats_cap(3)
    with pytest.raises(cvm.CvmatchError) as e:
        cvm.match_map(big, bsub)
    assert e.value.code == cvm.CVM_E_STATS_CAP
    assert "exceeds the exact-statistics bound for cn=3" in str(e.value)
