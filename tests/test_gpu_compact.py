"""GPU: the compact form of a small one-shot call (one stream, two
independent jobs per launch) against the forked form (three streams) and the
oracle.  Both forms run the same device code per job, so maps and six-tuples
must be identical bit for bit; the cases cover gray and RGBA, the alpha rule
with a varying and a constant alpha plane (cn 4 and 3), two- and
three-channel buffers, a flat template, row strides that defeat the vector
loads, template heights that are not multiples of 16 and maps that take
several tiles."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


# iw, ih, cn, step, tw, th, extra stride bytes, what
CASES = [
    (640, 480, 0, 4, 48, 48, 0, "alpha rule, varying alpha"),
    (300, 200, 0, 4, 33, 21, 0, "alpha rule, constant alpha"),
    (333, 97, 1, 1, 40, 33, 3, "gray, odd stride"),
    (257, 190, 2, 2, 30, 17, 0, "two channels"),
    (210, 150, 3, 3, 19, 40, 5, "three-byte pixels, odd stride"),
    (500, 300, 4, 4, 8, 8, 0, "flat template"),
    (1100, 700, 4, 4, 64, 50, 16, "several tiles"),
]


def buffers(rng, iw, ih, cn, step, tw, th, pad, what):
    istride = iw * step + pad
    img = rng.integers(0, 256, ih * istride, dtype=np.uint8)
    v = img.reshape(ih, istride)
    if "constant alpha" in what:
        v[:, 3:iw * 4:4] = 200
    x, y = int(rng.integers(0, iw - tw + 1)), int(rng.integers(0, ih - th + 1))
    tpl = np.ascontiguousarray(v[y:y + th, x * step:(x + tw) * step]).ravel()
    if what == "flat template":
        tpl[:] = 91
    return img, istride, tpl, tw * step


@pytest.mark.parametrize("case", range(len(CASES)))
def test_compact_equals_forked_and_oracle(cvm, orc, monkeypatch, case):
    iw, ih, cn, step, tw, th, pad, what = CASES[case]
    rng = np.random.default_rng(900 + case)
    img, istride, tpl, tstride = buffers(rng, iw, ih, cn, step, tw, th, pad, what)
    args = (img, istride, iw, ih, tpl, tstride, tw, th, cn, step)
    if cn == 0:  # Match's alpha rule: the RGBA entry on a strided view
        parent = img.reshape(ih, istride)[:, :iw * 4].reshape(ih, iw, 4)
        sub = tpl.reshape(th, tw, 4)
    got = {}
    for bound in ("2500000", "0"):
        monkeypatch.setenv("CVMATCH_COMPACT", bound)
        say(f"[compact] {what}: {iw}x{ih} cn {cn} step {step} tpl {tw}x{th}, compact bound {bound}")
        if cn == 0:
            got[bound] = (cvm.match_map(parent, sub)[0], tuple(cvm.match(parent, sub)))
        else:
            got[bound] = cvm.match_u8(*args, want_map=True)
    a, b = got["2500000"], got["0"]
    assert a[1] == b[1], (a[1], b[1])
    assert np.array_equal(a[0].view(np.uint32), b[0].view(np.uint32))
    if cn == 0:
        om, ot = orc.match_rgba(parent, sub, threads=8)
    else:
        om, ot = orc.match_u8(*args, threads=8)
    assert tuple(a[1]) == tuple(ot), (a[1], ot)
    assert np.array_equal(a[0].view(np.uint32), np.asarray(om, np.float32).reshape(a[0].shape).view(np.uint32))


def test_state_left_clear_between_compact_calls(cvm, orc, monkeypatch):
    """The compact form accumulates each call's template statistics and alpha
    flags into a state that its final reduction clears for the next call (no
    reset launch).  Alternating a varying-alpha call (cn 4) with a
    constant-alpha one (cn 3, which a leftover flag would turn into cn 4) and
    a second template (whose sums would add to leftover ones), through the
    eager, captured and replayed forms, must give the oracle's bits every
    time."""
    monkeypatch.setenv("CVMATCH_COMPACT", "2500000")
    rng = np.random.default_rng(77)
    cases = []
    for what, (tw, th) in (("alpha rule, varying alpha", (48, 48)), ("alpha rule, constant alpha", (33, 21)),
                           ("alpha rule, constant alpha", (40, 29))):
        img, istride, tpl, _ = buffers(rng, 320, 240, 0, 4, tw, th, 0, what)
        parent = img.reshape(240, istride)[:, :320 * 4].reshape(240, 320, 4)
        sub = tpl.reshape(th, tw, 4)
        om, ot = orc.match_rgba(parent, sub, threads=8)
        cases.append((parent, sub, np.asarray(om, np.float32), tuple(ot)))
    say("[compact] state cleared between calls: three alpha-rule shapes alternated 4 times")
    for rep in range(4):
        for i, (parent, sub, om, ot) in enumerate(cases):
            m, _, _ = cvm.match_map(parent, sub)
            assert tuple(cvm.match(parent, sub)) == ot, (rep, i)
            assert np.array_equal(m.view(np.uint32), om.reshape(m.shape).view(np.uint32)), (rep, i)


@pytest.mark.parametrize("case", [0, 6])
def test_column_slab_widths_change_no_bits(cvm, monkeypatch, case):
    """512-row compact column passes run in slabs of 8 columns by default;
    CVMATCH_COMPACT_CG=16 keeps 16-column slabs, CVMATCH_COMPACT_WIDE=0 the
    256-thread form.  Every form gives the same bits."""
    iw, ih, cn, step, tw, th, pad, what = CASES[case]
    rng = np.random.default_rng(950 + case)
    img, istride, tpl, tstride = buffers(rng, iw, ih, cn, step, tw, th, pad, what)
    args = (img, istride, iw, ih, tpl, tstride, tw, th, cn if cn else 4, step)
    monkeypatch.setenv("CVMATCH_COMPACT", "2500000")
    got = []
    for cg, wide in (("8", "1"), ("16", "1"), ("8", "0")):
        monkeypatch.setenv("CVMATCH_COMPACT_CG", cg)
        monkeypatch.setenv("CVMATCH_COMPACT_WIDE", wide)
        say(f"[compact] {what}: column slabs of {cg}, wide {wide}")
        got.append(cvm.match_u8(*args, want_map=True))
    for m, r in got[1:]:
        assert tuple(r) == tuple(got[0][1])
        assert np.array_equal(m.view(np.uint32), got[0][0].view(np.uint32))
