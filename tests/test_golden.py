"""The CPU oracle against the committed fixtures in tests/golden/ -- CPU only.

  plan pins      the reference's 15 (plan_test.go:12-29), exact
  golden #110    the reference's recorded hash, hashed here in numpy, so the
                 fixture's hash function is pinned as well as the oracle
  case hashes    recorded from the oracle (1 and 4 threads): any moved bit of
                 any of the nine golden-shape maps fails them
  FP64 scores    refMatch's float64 steps (cvmatch_test.go:13-81) at sampled
                 windows, computed without the oracle: the oracle's exact
                 FP64 scores within 1e-9, its float32 map within 1e-4 (the
                 reference's own tolerance, cvmatch_test.go:103-123)
"""
import numpy as np
import pytest

from golden import fixtures as F

G = F.load()
CASES = G["cases"]


@pytest.mark.parametrize("pin", G["plan_pins"])
def test_fixture_plan_pins(orc, pin):
    got, _ = orc.plan(*pin[:4])
    assert got == tuple(pin[4:])


def test_fixture_golden_110(orc):
    spec = G["golden_110"]
    img, iw, ih, tpl, tw, th = F.canary_110(spec)
    for threads in (1, 4):
        m, t = orc.match_u8(img, iw, iw, ih, tpl, tw, tw, th, spec["cn"], spec["step"], threads)
        assert F.fnv_hash(m, t) == int(spec["hash"], 16)


def test_fixture_hash_moves_with_one_bit(orc):
    spec = G["golden_110"]
    img, iw, ih, tpl, tw, th = F.canary_110(spec)
    m, t = orc.match_u8(img, iw, iw, ih, tpl, tw, tw, th, 1, 1, 1)
    m = m.copy()
    m.view(np.uint32)[0, 0] ^= 1
    assert F.fnv_hash(m, t) != int(spec["hash"], 16)
    assert F.fnv_hash(m, t) == orc.golden_hash(m, t)


@pytest.mark.parametrize("case", CASES, ids=[str(c["id"]) for c in CASES])
def test_fixture_case_hashes(orc, case):
    img, istride, tpl, tstride = F.case_inputs(case)
    for threads in (1, 4):
        m, t = orc.match_u8(img, istride, case["iw"], case["ih"], tpl, tstride, case["tw"], case["th"],
                            case["cn"], case["step"], threads)
        assert t == F.tuple_of(case["tuple"])
        assert F.fnv_hash(m, t) == int(case["hash"], 16)


@pytest.mark.parametrize("case", CASES, ids=[str(c["id"]) for c in CASES])
def test_fixture_fp64_scores(orc, case):
    img, istride, tpl, tstride = F.case_inputs(case)
    iw, ih, tw, th, cn, step = (case[k] for k in ("iw", "ih", "tw", "th", "cn", "step"))
    fp = case["fp64"]
    xs, ys = np.array(fp["xs"], np.int32), np.array(fp["ys"], np.int32)
    want = np.array([float.fromhex(v) for v in fp["scores"]])
    got = orc.fp64_points(img, istride, iw, ih, tpl, tstride, tw, th, cn, step, xs, ys)
    assert np.max(np.abs(got - want)) <= 1e-9
    m, t = orc.match_u8(img, istride, iw, ih, tpl, tstride, tw, th, cn, step, 1)
    assert np.max(np.abs(m[ys, xs].astype(np.float64) - want)) <= 1e-4
    # the last two sampled windows are the six-tuple's locations
    assert (xs[-2], ys[-2], xs[-1], ys[-1]) == (t[1], t[2], t[4], t[5])
