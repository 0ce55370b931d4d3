"""GPU: the allocation-free entry point cvm_match_u8_direct_device_noalloc
gives exactly what cvm_match_u8_direct_device gives (map bits and six-tuple)
on the same device buffers, over consecutive calls whose second input is
smaller than the first (the shared static scratch keeps the previous call's
contents), with and without a map, gray and RGBA, on the null stream and on
a caller's stream; a shape whose scratch exceeds the array is refused
before any work."""
import ctypes as C
import sys

import numpy as np
import pytest

from cvmatch_amd import synth

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


class Dev:
    def __init__(self, L):
        self.L, self.bufs = L, []

    def put(self, a):
        p = C.c_void_p()
        assert self.L.cvm_device_alloc(max(a.nbytes, 4), C.byref(p)) == 0
        assert self.L.cvm_copy_h2d(p, C.c_void_p(a.ctypes.data), a.nbytes) == 0
        self.bufs.append(p)
        return p

    def get(self, p, n):
        out = np.empty(n, np.float32)
        assert self.L.cvm_copy_d2h(C.c_void_p(out.ctypes.data), p, out.nbytes) == 0
        return out

    def free(self):
        for p in self.bufs:
            self.L.cvm_device_free(p)


def cases():
    big, bsub, _ = synth.workload_map_4k(seed=3, alpha="random")
    small, ssub, _ = synth.workload_match_640(seed=1, alpha="random")
    gray, gsub = synth.workload_gray_ui_1080p(seed=2)[:2]
    tiny = np.ascontiguousarray(gray[100:300, 200:520])
    tsub = np.ascontiguousarray(tiny[40:64, 50:82])
    # (parent, template, cn, step): RGBA with the alpha rule, then gray
    return [(big, bsub, 0, 4), (small, ssub, 0, 4), (gray, gsub[:, :64].copy(), 1, 1), (tiny, tsub, 1, 1)]


@pytest.mark.parametrize("with_map", [True, False])
@pytest.mark.parametrize("on_stream", [False, True])
def test_noalloc_equals_direct_device(cvm, with_map, on_stream):
    L = cvm.lib()
    hip = C.CDLL("libamdhip64.so.7")  # the runtime the library has loaded
    stream = C.c_void_p()
    if on_stream:
        assert hip.hipStreamCreateWithFlags(C.byref(stream), C.c_uint(1)) == 0
    D = Dev(L)
    try:
        for (p, t, cn, step) in cases():  # each case follows a larger one
            ih, iw = p.shape[:2]
            th, tw = t.shape[:2]
            rw, rh = iw - tw + 1, ih - th + 1
            say(f"[noalloc] {iw}x{ih} step {step} vs {tw}x{th} cn {cn} map={with_map} stream={on_stream}")
            dp, ds = D.put(p), D.put(t)
            dm_ref = D.put(np.full(rw * rh, -3.0, np.float32)) if with_map else None
            dm_new = D.put(np.full(rw * rh, -5.0, np.float32)) if with_map else None
            want, got = cvm.cvm_result(), cvm.cvm_result()
            rc = L.cvm_match_u8_direct_device(dp, iw * step, iw, ih, ds, tw * step, tw, th, cn, step, dm_ref,
                                              C.byref(want), stream if on_stream else None)
            assert rc == 0, L.cvm_last_error_detail()
            for k in range(2):
                rc = L.cvm_match_u8_direct_device_noalloc(dp, iw * step, iw, ih, ds, tw * step, tw, th, cn, step,
                                                          dm_new, C.byref(got), stream if on_stream else None)
                assert rc == 0, L.cvm_last_error_detail()
                assert got.tuple() == want.tuple(), (k, got.tuple(), want.tuple())
                if with_map:
                    a, b = D.get(dm_ref, rw * rh), D.get(dm_new, rw * rh)
                    assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), k
    finally:
        D.free()
        if on_stream:
            hip.hipStreamDestroy(stream)


def test_noalloc_refuses_oversized_scratch(cvm):
    """A 7680x4320 gray frame with no map needs a 132 MB C' map: refused
    with CVM_E_UNSUPPORTED before any device work (the pointers are never
    read)."""
    L = cvm.lib()
    out = cvm.cvm_result()
    fake = C.c_void_p(256)  # never dereferenced: the size check comes first
    rc = L.cvm_match_u8_direct_device_noalloc(fake, 7680, 7680, 4320, fake, 16, 16, 16, 1, 1, None, C.byref(out),
                                              None)
    assert rc == cvm.CVM_E_UNSUPPORTED
    assert "allocation-free call needs" in L.cvm_last_error_detail().decode()
