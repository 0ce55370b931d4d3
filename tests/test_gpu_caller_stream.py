"""GPU: the device-resident entry points on a caller-created HIP stream (the
same HIP runtime the library uses, reached through ctypes): the caller's
stream orders the call, repeated calls are captured and replayed on that
stream, and every result equals the host-buffer call bit for bit."""
import ctypes as C
import sys

import numpy as np
import pytest

from cvmatch_amd import synth

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_match_and_fleet_on_caller_stream(cvm):
    L = cvm.lib()
    hip = C.CDLL("libamdhip64.so.7")  # already loaded by the library: the same runtime
    stream = C.c_void_p()
    assert hip.hipStreamCreateWithFlags(C.byref(stream), C.c_uint(1)) == 0  # non-blocking, default priority
    bufs = []

    def dev(a):
        p = C.c_void_p()
        assert L.cvm_device_alloc(max(a.nbytes, 4), C.byref(p)) == 0
        assert L.cvm_copy_h2d(p, C.c_void_p(a.ctypes.data), a.nbytes) == 0
        bufs.append(p)
        return p

    try:
        parent, sub, _ = synth.workload_match_640(seed=9, alpha="random")
        want_map, _, _ = cvm.match_map(parent, sub)
        want = cvm.match(parent, sub)
        rh, rw = want_map.shape
        dp, ds = dev(parent), dev(sub)
        dm = dev(np.full(rw * rh, -7.0, np.float32))
        out = cvm.cvm_result()
        say("[stream] cvm_match_u8_device on a caller stream, 5 calls")
        for k in range(5):  # eager, capture, replays
            rc = L.cvm_match_u8_device(dp, 640 * 4, 640, 480, ds, 48 * 4, 48, 48, 0, 4, dm, C.byref(out), stream)
            assert rc == 0, L.cvm_last_error_detail()
            assert hip.hipStreamSynchronize(stream) == 0
            got = np.empty(rw * rh, np.float32)
            assert L.cvm_copy_d2h(C.c_void_p(got.ctypes.data), dm, got.nbytes) == 0
            assert np.array_equal(got.view(np.uint32), want_map.ravel().view(np.uint32)), k
            assert out.tuple() == want, k
        # FindAll on a device view (an ROI of the frame) on the same stream
        frame = synth.flat_rect_scene(640, 400, np.random.default_rng(3), n_rects=25, channels=4)
        frame[..., 3] = 255
        subs = [frame[y:y + h, x:x + w].copy() for (x, y, w, h) in [(300, 150, 40, 30), (420, 220, 24, 24)]]
        fl = cvm.NewFleet(*[cvm.NewMatcher(s) for s in subs])
        x0, y0, x1, y1 = 250, 100, 560, 330
        want_roi = [tuple(r) for r in fl.find_all(frame[y0:y1, x0:x1])]
        df = dev(np.ascontiguousarray(frame))
        ptr = df.value + (y0 * 640 + x0) * 4
        got_roi = [tuple(r) for r in fl.find_all_device(ptr, 640 * 4, x1 - x0, y1 - y0, stream=stream)]
        assert got_roi == want_roi
        say("[stream] direct FindAll and Find on the device view")
        want_direct = [tuple(r) for r in fl.find_all(frame[y0:y1, x0:x1], direct=True)]
        got_direct = [tuple(r) for r in fl.find_all_device(ptr, 640 * 4, x1 - x0, y1 - y0, direct=True, stream=stream)]
        assert got_direct == want_direct
        for s in subs:
            m = cvm.NewMatcher(s)
            assert m.find_device(ptr, 640 * 4, x1 - x0, y1 - y0, stream=stream) == m.find(frame[y0:y1, x0:x1])
        # a gray Matcher on a gray device view (row pitch not a multiple of the width)
        gray = np.zeros((400, 656), np.uint8)
        gray[:, :640] = cvm.rgba_to_gray(frame)
        dg = dev(gray)
        gm = cvm.NewGrayMatcher(subs[0])
        gptr = dg.value + y0 * 656 + x0
        assert gm.find_device(gptr, 656, x1 - x0, y1 - y0, stream=stream) == gm.find(gray[y0:y1, x0:x1])
    finally:
        for p in bufs:
            L.cvm_device_free(p)
        hip.hipStreamDestroy(stream)
