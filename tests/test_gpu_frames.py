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


@pytest.mark.parametrize("pinned,size", [(True, (1280, 720)), (False, (320, 200))])
def test_frame_buffer_reused_after_submit(cvm, pinned, size):
    """The caller's frame buffer may be reused as soon as find returns, also
    when the upload reads it directly (a cvm_host_alloc frame, or a frame
    below the staging size): it is overwritten right after each submission
    and every result still equals the synchronous Find of the frame it held."""
    w, h = size
    base, fr = frames(6, 77, w=w, h=h)
    m = cvm.NewMatcher(base[50:90, 60:110].copy())
    want = [(i, m.find(f)) for i, f in enumerate(fr)]
    buf = cvm.host_array(fr[0].shape) if pinned else np.empty_like(fr[0])
    q = cvm.FrameStream(2)
    got = []
    say(f"[frames] {'pinned' if pinned else 'pageable'} {w}x{h} buffer overwritten after each find")
    for i, f in enumerate(fr):
        buf[...] = f
        q.find(m, buf, tag=i)
        buf[...] = 0
        if q.pending() >= 2:
            got.append(q.result())
    while q.pending():
        got.append(q.result())
    assert got == want


def test_frames_in_flight_beside_caller_stream_calls(cvm):
    """4K frames in flight on a frame stream while device-resident calls run
    on a caller's HIP stream (Match with its map, FindAll plain and direct,
    Find): the two share no device scratch, so every result equals the
    synchronous call's bit for bit (README.md:419-431)."""
    import ctypes as C

    L = cvm.lib()
    hip = C.CDLL("libamdhip64.so.7")  # the runtime the library has loaded
    stream = C.c_void_p()
    assert hip.hipStreamCreateWithFlags(C.byref(stream), C.c_uint(1)) == 0
    bufs = []

    def dev(a):
        p = C.c_void_p()
        assert L.cvm_device_alloc(max(a.nbytes, 4), C.byref(p)) == 0
        assert L.cvm_copy_h2d(p, C.c_void_p(a.ctypes.data), a.nbytes) == 0
        bufs.append(p)
        return p

    try:
        big, sub4k, _ = synth.workload_map_4k(seed=3, alpha="random")
        shifted = np.ascontiguousarray(np.roll(big, 5, axis=1))
        m4 = cvm.NewMatcher(sub4k)
        want_frames = [m4.find(big), m4.find(shifted)]
        parent, sub, _ = synth.workload_match_640(seed=9, alpha="random")
        want_map, _, _ = cvm.match_map(parent, sub)
        want_t = cvm.match(parent, sub)
        rh, rw = want_map.shape
        frame = synth.flat_rect_scene(640, 400, np.random.default_rng(3), n_rects=25, channels=4)
        frame[..., 3] = 255
        subs = [frame[y:y + hh, x:x + ww].copy() for (x, y, ww, hh) in [(300, 150, 40, 30), (420, 220, 24, 24)]]
        ms = [cvm.NewMatcher(s) for s in subs]
        fl = cvm.NewFleet(*ms)
        want_all = [tuple(r) for r in fl.find_all(frame)]
        want_direct = [tuple(r) for r in fl.find_all(frame, direct=True)]
        want_find = [m.find(frame) for m in ms]
        dp, ds, df = dev(parent), dev(sub), dev(frame)
        dm = dev(np.zeros(rw * rh, np.float32))
        out = cvm.cvm_result()
        q = cvm.FrameStream(3)
        got = []
        tags = 0
        say("[frames] 4K frames in flight beside caller-stream calls")
        for r in range(4):
            for f in (big, shifted):
                q.find(m4, f, tag=tags)
                tags += 1
            rc = L.cvm_match_u8_device(dp, 640 * 4, 640, 480, ds, 48 * 4, 48, 48, 0, 4, dm, C.byref(out), stream)
            assert rc == 0, L.cvm_last_error_detail()
            gm = np.empty(rw * rh, np.float32)
            assert hip.hipStreamSynchronize(stream) == 0
            assert L.cvm_copy_d2h(C.c_void_p(gm.ctypes.data), dm, gm.nbytes) == 0
            assert np.array_equal(gm.view(np.uint32), want_map.ravel().view(np.uint32)), r
            assert out.tuple() == want_t, r
            assert [tuple(x) for x in fl.find_all_device(df.value, 640 * 4, 640, 400, stream=stream)] == want_all
            assert [tuple(x) for x in fl.find_all_device(df.value, 640 * 4, 640, 400, direct=True,
                                                         stream=stream)] == want_direct
            assert [m.find_device(df.value, 640 * 4, 640, 400, stream=stream) for m in ms] == want_find
            while q.pending() > 3:
                got.append(q.result())
        while q.pending():
            got.append(q.result())
        assert got == [(k, want_frames[k % 2]) for k in range(tags)]
    finally:
        for p in bufs:
            L.cvm_device_free(p)
        hip.hipStreamDestroy(stream)


def test_staged_uploads_after_a_stream_is_freed(cvm):
    """A frame stream's uploads go through the shared pinned staging ring on
    the stream's own copy stream; freeing the stream retires those staging
    slots first, so later host-buffer calls (whose staged copies wait on the
    slots) work: several streams created, used with 14.7 MB pageable frames
    and freed, each followed by staged host calls."""
    import gc

    frames, tpls = synth.workload_fleet(seed=14, frames=3, members=2)
    frames = [np.ascontiguousarray(f) for f in frames]
    m = cvm.NewMatcher(tpls[0])
    want = [m.find(f) for f in frames]
    say("[frames] create / use / free frame streams between staged host calls")
    for rep in range(4):
        q = cvm.FrameStream(2 + rep % 2)
        for i, f in enumerate(frames):
            q.find(m, f, tag=i)
        got = []
        while q.pending():
            got.append(q.result())
        assert got == list(enumerate(want)), rep
        del q
        gc.collect()
        assert [m.find(f) for f in frames] == want, rep
        fl = cvm.NewFleet(m)
        assert tuple(fl.find_all(frames[rep % 3])[0]) == want[rep % 3], rep
