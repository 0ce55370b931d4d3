"""GPU: the direct FindAll on the Fleet workload itself (2560x1440 RGBA frames
of the seeded flat-rectangle stream, alpha 255, 256 templates U[16,128]^2,
half with noise) against the FP64 oracle.

Coverage per checked frame:
  * every member's six-tuple: both reported locations scored by the FP64
    oracle, values equal to those scores rounded to float32;
  * every member: 4096 seeded windows of its direct map equal the FP64
    oracle's scores rounded to float32, bit for bit, and FindAll's six-tuple
    equals the one-shot direct Match's (whose whole-map extrema are checked
    against FP64 maps elsewhere);
  * 16 members (every 16th, all sizes): the full map equals the FP64 map
    rounded to float32 bit for bit, and the six-tuple is its first-occurrence
    extrema.
The plain float32 path (the reference's op sequence) is reported beside it:
max |delta| of its scores against FP64 at the same windows.  The plain
FindAll itself is checked bit for bit against the oracle's FindAll on the same
stream with 30 members (several normalisation batches).
"""
import os
import sys
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from cvmatch_amd import synth

pytestmark = pytest.mark.gpu

FRAMES = (0, 41, 63)
SAMPLES = 4096


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


def first_extrema(m):
    rw = m.shape[1]
    i0, i1 = int(np.argmin(m)), int(np.argmax(m))
    return (float(m.flat[i0]), i0 % rw, i0 // rw, float(m.flat[i1]), i1 % rw, i1 // rw)


@pytest.fixture(scope="module")
def workload(cvm):
    # the benchmarked workload itself (default arguments: 64 frames, whose
    # count also seeds the template choice)
    frames, tpls = synth.workload_fleet()
    ms = [cvm.NewMatcher(t) for t in tpls]
    return frames, tpls, cvm.NewFleet(*ms)


@pytest.mark.parametrize("fi", FRAMES)
def test_direct_fleet_equals_fp64(cvm, orc, workload, fi):
    frames, tpls, fl = workload
    frame = np.ascontiguousarray(frames[fi])
    fh, fw = frame.shape[:2]
    say(f"[fleet-fp64] frame {fi}: direct FindAll of {len(tpls)} members")
    got = [tuple(r) for r in fl.find_all(frame, direct=True)]
    rng = np.random.default_rng(1000 + fi)
    jobs = []
    plain_delta = 0.0
    for i, t in enumerate(tpls):
        th, tw = t.shape[:2]
        rw, rh = fw - tw + 1, fh - th + 1
        g = got[i]
        xs = np.concatenate([rng.integers(0, rw, SAMPLES), [g[1], g[4]]]).astype(np.int32)
        ys = np.concatenate([rng.integers(0, rh, SAMPLES), [g[2], g[5]]]).astype(np.int32)
        gm, gt = cvm.match_direct(frame, t, want_map=True)
        assert tuple(gt) == g, (i, gt, g)
        pm, _, _ = cvm.match_map(frame, t)
        jobs.append((i, t, xs, ys, gm[ys, xs].copy(), pm[ys, xs].copy()))
        if i % 16 == 0:
            say(f"[fleet-fp64] frame {fi} member {i} ({tw}x{th}): full FP64 map")
            om = orc.fp64_map_fft(frame, t, 3).astype(np.float32)
            nd = int(np.count_nonzero(bits(gm) != bits(om)))
            assert nd == 0, (i, nd)
            assert g == first_extrema(om), (i, g, first_extrema(om))

    def score(job):
        i, t, xs, ys, gv, pv = job
        th, tw = t.shape[:2]
        ov = orc.fp64_points(frame, fw * 4, fw, fh, np.ascontiguousarray(t), tw * 4, tw, th, 3, 4, xs, ys)
        return i, ov, gv, pv

    say(f"[fleet-fp64] frame {fi}: FP64 scores at {SAMPLES} + 2 windows per member")
    workers = max(1, min(16, len(os.sched_getaffinity(0))))
    with ThreadPoolExecutor(workers) as ex:
        for i, ov, gv, pv in ex.map(score, jobs):
            o32 = ov.astype(np.float32)
            assert np.array_equal(bits(gv), bits(o32)), (i, int(np.count_nonzero(bits(gv) != bits(o32))))
            g = got[i]
            assert (g[0], g[3]) == (float(o32[-2]), float(o32[-1])), i
            plain_delta = max(plain_delta, float(np.max(np.abs(pv.astype(np.float64) - ov))))
    say(f"[fleet-fp64] frame {fi}: direct == FP64 (float32) at {len(tpls) * (SAMPLES + 2)} windows; "
        f"plain float32 path max |delta| vs FP64 {plain_delta:.3g}")


def test_plain_fleet_multi_batch_bitwise(cvm, orc):
    """The plain FindAll on the 2560x1440 stream with 30 members: their maps
    (~103 M floats) span three normalisation batches of at most 192 MiB, and
    the bucket members go through the paired column launch; every six-tuple
    bit for bit against the oracle's FindAll, on two frames."""
    frames, tpls = synth.workload_fleet(frames=24, members=30)
    fl = cvm.NewFleet(*[cvm.NewMatcher(t) for t in tpls])
    for fi in (0, 23):
        frame = np.ascontiguousarray(frames[fi])
        say(f"[fleet] plain FindAll, 30 members, frame {fi}")
        got = [tuple(r) for r in fl.find_all(frame)]
        assert got == orc.fleet_find_all(frame, tpls), fi
