"""GPU: captured call graphs replay the same bits as the eager calls.

A repeated call (device-resident Match, FindAll of either form, a frame
stream's Find) runs eagerly the first time its arguments are seen, is
captured into a hipGraph on the second sighting and replays afterwards; up
to four argument sets (frame buffers) keep their graphs.  Every replay must
equal the eager result, read the frame buffer's current contents, and stay
correct when other calls (which may grow the context's buffers) run in
between."""
import ctypes as C
import sys

import numpy as np
import pytest

from cvmatch_amd import synth

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture()
def dev(cvm):
    L = cvm.lib()
    bufs = []

    def put(a, p=None):
        if p is None:
            p = C.c_void_p()
            assert L.cvm_device_alloc(max(a.nbytes, 4), C.byref(p)) == 0
            bufs.append(p)
        assert L.cvm_copy_h2d(p, C.c_void_p(a.ctypes.data), a.nbytes) == 0
        return p

    yield put
    for p in bufs:
        L.cvm_device_free(p)


@pytest.mark.parametrize("direct", [False, True])
def test_fleet_graph_replays(cvm, dev, direct):
    frames, tpls = synth.workload_fleet(seed=9, frames=5, members=24, w=960, h=540)
    frames = [np.ascontiguousarray(f) for f in frames]
    fl = cvm.NewFleet(*[cvm.NewMatcher(t) for t in tpls])
    want = [[tuple(r) for r in fl.find_all(f, direct=direct)] for f in frames]
    bufs = [dev(frames[k]) for k in range(3)]
    say(f"[graphs] FindAll direct={direct}: 3 device frames cycled 4 times")
    for rep in range(4):
        for k, b in enumerate(bufs):
            got = [tuple(r) for r in fl.find_all_device(b.value, 960 * 4, 960, 540, direct=direct)]
            assert got == want[k], (rep, k)
        # another call in between that grows this thread's context buffers
        parent, sub, _ = synth.workload_match_640(seed=rep, alpha="random")
        cvm.match_map(parent, sub)
    # new contents in a buffer whose FindAll is captured: the replay reads them
    dev(frames[4], bufs[1])
    for _ in range(2):
        assert [tuple(r) for r in fl.find_all_device(bufs[1].value, 960 * 4, 960, 540, direct=direct)] == want[4]
    # host frames (uploaded into the context's frame buffer): replays too
    for rep in range(3):
        for k in (0, 3):
            assert [tuple(r) for r in fl.find_all(frames[k], direct=direct)] == want[k], (rep, k)


def test_device_match_graph_alternating_buffers(cvm, dev):
    L = cvm.lib()
    cases = [synth.workload_match_640(seed=s, alpha="random")[:2] for s in (21, 22)]
    want = []
    for p, t in cases:
        m, _, _ = cvm.match_map(p, t)
        want.append((m, cvm.match(p, t)))
    rw, rh = 640 - 48 + 1, 480 - 48 + 1
    ins = [(dev(p), dev(t), dev(np.zeros(rw * rh, np.float32))) for p, t in cases]
    out = cvm.cvm_result()
    say("[graphs] device Match on two buffer sets, alternating 5 times")
    for rep in range(5):
        for k, (dp, ds, dm) in enumerate(ins):
            rc = L.cvm_match_u8_device(dp, 640 * 4, 640, 480, ds, 48 * 4, 48, 48, 0, 4, dm, C.byref(out), None)
            assert rc == 0, L.cvm_last_error_detail()
            got = np.empty(rw * rh, np.float32)
            assert L.cvm_copy_d2h(C.c_void_p(got.ctypes.data), dm, got.nbytes) == 0
            assert np.array_equal(got.view(np.uint32), want[k][0].ravel().view(np.uint32)), (rep, k)
            assert out.tuple() == want[k][1], (rep, k)


@pytest.mark.parametrize("mode", ["1", "2", "0"])
def test_device_match_graph_or_eager(cvm, dev, monkeypatch, mode):
    """A small synchronous device call replays its graph (1), launches
    eagerly (2) or times both and keeps the faster (0, the default): every
    call gives the eager first call's bits, through the capture, the trials
    and the calls after the choice."""
    monkeypatch.setenv("CVMATCH_GRAPH_MODE", mode)
    L = cvm.lib()
    p, t = synth.workload_match_640(seed=23 + int(mode), alpha="random")[:2]
    want_m, _, _ = cvm.match_map(p, t)
    want_r = cvm.match(p, t)
    rw, rh = 640 - 48 + 1, 480 - 48 + 1
    dp, ds, dm = dev(p), dev(t), dev(np.zeros(rw * rh, np.float32))
    out = cvm.cvm_result()
    say(f"[graphs] device Match, graph mode {mode}: 14 calls")
    for rep in range(14):
        rc = L.cvm_match_u8_device(dp, 640 * 4, 640, 480, ds, 48 * 4, 48, 48, 0, 4, dm, C.byref(out), None)
        assert rc == 0, L.cvm_last_error_detail()
        got = np.empty(rw * rh, np.float32)
        assert L.cvm_copy_d2h(C.c_void_p(got.ctypes.data), dm, got.nbytes) == 0
        assert np.array_equal(got.view(np.uint32), want_m.ravel().view(np.uint32)), rep
        assert out.tuple() == tuple(want_r), rep


def test_frame_stream_graph_replays(cvm):
    frames, tpls = synth.workload_fleet(seed=10, frames=6, members=3, w=800, h=600)
    frames = [np.ascontiguousarray(f) for f in frames]
    ms = [cvm.NewMatcher(t) for t in tpls]
    want = [[m.find(f) for m in ms] for f in frames]
    want_d = [[m.find(f, direct=True) for m in ms] for f in frames]
    q = cvm.FrameStream(3)
    got, exp = [], []
    say("[graphs] frame stream: 6 frames x 3 matchers x 3 rounds, plain and direct")
    tag = 0
    for rep in range(3):
        for i, f in enumerate(frames):
            for j, m in enumerate(ms):
                d = (i + j + rep) % 2 == 1
                q.find(m, f, tag=tag, direct=d)
                exp.append((tag, (want_d if d else want)[i][j]))
                tag += 1
                while q.pending() > 3:
                    got.append(q.result())
    while q.pending():
        got.append(q.result())
    assert got == exp
