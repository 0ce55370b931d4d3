"""GPU: frame streams (cvm_frames_*): Find over a sequence of host frames
with uploads overlapping the matching of earlier frames gives, frame by
frame and in submission order, exactly the synchronous Find's six-tuple
(matcher.go:83-130), plain and direct, for several depths, two matchers
interleaved, a flat template, a synchronous call in between and frames
whose buffers are overwritten right after submission."""
import sys

import numpy as np
import pytest

from cvmatch_amd import synth

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def frames(n, seed, w=1280, h=720):
    rng = np.random.default_rng(seed)
    base = synth.add_noise(synth.flat_rect_scene(w + 64, h + 64, rng, n_rects=60, channels=4), 3.0, rng,
                           channels=[0, 1, 2])
    base[..., 3] = 255
    out = []
    for i in range(n):
        dx, dy = int(rng.integers(0, 64)), int(rng.integers(0, 64))
        out.append(np.ascontiguousarray(base[dy:dy + h, dx:dx + w]))
    return base, out


@pytest.mark.parametrize("depth", [2, 3])
def test_stream_equals_sync_find(cvm, depth):
    base, fr = frames(9, 40 + depth)
    m1 = cvm.NewMatcher(base[200:248, 300:364].copy())
    m2 = cvm.NewMatcher(base[500:530, 900:940].copy())
    flat = cvm.NewMatcher(np.full((8, 8, 4), 50, np.uint8))
    plan = [(m1, False), (m2, False), (m1, True), (flat, False), (m2, True)]
    want = []
    for i, f in enumerate(fr):
        for j, (m, d) in enumerate(plan):
            want.append(((i << 8) | j, m.find(f, direct=d)))
    q = cvm.FrameStream(depth)
    got = []
    say(f"[frames] depth {depth}: {len(fr)} frames x {len(plan)} finds")
    scratch = np.empty_like(fr[0])
    for i, f in enumerate(fr):
        for j, (m, d) in enumerate(plan):
            scratch[...] = f  # the stream copies the frame out before returning
            q.find(m, scratch, tag=(i << 8) | j, direct=d)
            if q.pending() > depth + 1:
                got.append(q.result())
        if i == 4:
            assert m1.find(fr[0]) == want[0][1]  # a synchronous call in between
    while q.pending():
        got.append(q.result())
    assert got == want
    with pytest.raises(cvm.CvmatchError):
        q.result()


def test_stream_4k_matches_sync(cvm):
    parent, sub, (cx, cy) = synth.workload_map_4k(seed=3, alpha="random")
    m = cvm.NewMatcher(sub)
    want = m.find(parent)
    assert want[4:] == (cx, cy)
    q = cvm.FrameStream(2)
    for k in range(6):
        q.find(m, parent, tag=k)
    for k in range(6):
        assert q.result() == (k, want)
