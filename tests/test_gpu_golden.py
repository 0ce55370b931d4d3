"""GPU: the HIP path against the committed fixtures in tests/golden/, with no
oracle call at run time.

  plain form   every golden-shape map (impl_golden_test.go:60-69 shapes,
               PCG64 inputs) hashed as goldenHash does
               (impl_golden_test.go:27-44): hash and six-tuple equal the
               recorded ones, so the device map is bit-identical to the
               recorded oracle map; the extrema-only call gives the same
               six-tuple; #110 gives the reference's own hash
  direct form  scores at the sampled windows within 1e-7 of refMatch's
               float64 brute force (cvmatch_test.go:13-81): float32
               rounding of the FP64 score (<= 6e-8 on [-1, 1]) plus the brute
               force's own float64 error (< 1e-9); packed 3-byte pixels are
               refused
"""
import sys

import numpy as np
import pytest

from golden import fixtures as F

pytestmark = pytest.mark.gpu

G = F.load()
CASES = G["cases"]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def args_of(case):
    img, istride, tpl, tstride = F.case_inputs(case)
    return (img, istride, case["iw"], case["ih"], tpl, tstride, case["tw"], case["th"], case["cn"], case["step"])


def test_gpu_fixture_golden_110(cvm):
    spec = G["golden_110"]
    img, iw, ih, tpl, tw, th = F.canary_110(spec)
    m, t = cvm.match_u8(img, iw, iw, ih, tpl, tw, tw, th, spec["cn"], spec["step"], want_map=True)
    assert F.fnv_hash(m, t) == int(spec["hash"], 16)


@pytest.mark.parametrize("case", CASES, ids=[str(c["id"]) for c in CASES])
def test_gpu_fixture_hash(cvm, case):
    say(f"[golden] plain case {case['id']}")
    a = args_of(case)
    m, t = cvm.match_u8(*a, want_map=True)
    assert t == F.tuple_of(case["tuple"])
    assert F.fnv_hash(m, t) == int(case["hash"], 16)
    _, t2 = cvm.match_u8(*a, want_map=False)
    assert t2 == t


@pytest.mark.parametrize("case", CASES, ids=[str(c["id"]) for c in CASES])
def test_gpu_fixture_direct_scores(cvm, case):
    a = args_of(case)
    cn, step = case["cn"], case["step"]
    if step not in (1, 4):
        say(f"[golden] direct case {case['id']}: packed {step}-byte pixels must be refused")
        with pytest.raises(cvm.CvmatchError) as e:
            cvm.match_u8(*a, want_map=True, direct=True)
        assert e.value.code == cvm.CVM_E_BAD_ARGS
        return
    say(f"[golden] direct case {case['id']}")
    m, t = cvm.match_u8(*a, want_map=True, direct=True)
    fp = case["fp64"]
    xs, ys = np.array(fp["xs"]), np.array(fp["ys"])
    want = np.array([float.fromhex(v) for v in fp["scores"]])
    assert np.max(np.abs(m[ys, xs].astype(np.float64) - want)) <= 1e-7
    # first-occurrence extrema of the direct map itself
    i0, i1 = int(np.argmin(m)), int(np.argmax(m))
    rw = m.shape[1]
    assert t == (float(m.flat[i0]), i0 % rw, i0 // rw, float(m.flat[i1]), i1 % rw, i1 // rw)
