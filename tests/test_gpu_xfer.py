"""GPU: host-buffer transfers through the pinned staging chunks.

Pageable host buffers of 1 MiB or more move in 4 MiB pinned chunks copied by
host threads (four chunks in flight); smaller ones and cvm_host_alloc memory
go straight to the DMA engine.  The bytes must arrive unchanged for every
size around the chunk edges, and no thread count or pinned input may change
an output bit of Match, MatchMap, Find or FindAll (the reference's Threads
guarantee, threads_test.go:62-81).
"""
import ctypes as C
import sys

import numpy as np
import pytest

from cvmatch_amd import synth

pytestmark = pytest.mark.gpu

MIB = 1 << 20


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture
def threads(cvm):
    saved = cvm.Threads
    yield cvm
    cvm.Threads = saved


def dalloc(L, n):
    p = C.c_void_p()
    assert L.cvm_device_alloc(n, C.byref(p)) == 0
    return p


@pytest.mark.parametrize("nthreads", [1, 16])
def test_round_trip_chunk_edges(threads, nthreads):
    cvm = threads
    cvm.Threads = nthreads
    L = cvm._L()
    rng = np.random.default_rng(nthreads)
    sizes = [1, 4095, MIB - 1, MIB, MIB + 1, 4 * MIB - 1, 4 * MIB, 4 * MIB + 7, 16 * MIB + 3, 21 * MIB + 11]
    big = max(sizes)
    d = dalloc(L, big + 64)
    try:
        for n in sizes:
            say(f"[xfer] round trip {n} bytes, {nthreads} threads")
            src = rng.integers(0, 256, size=n, dtype=np.uint8)
            assert L.cvm_copy_h2d(d, C.c_void_p(src.ctypes.data), n) == 0
            back = np.full(n + 16, 0xEE, dtype=np.uint8)
            assert L.cvm_copy_d2h(C.c_void_p(back.ctypes.data), d, n) == 0
            assert np.array_equal(back[:n], src), n
            assert np.all(back[n:] == 0xEE), n
    finally:
        L.cvm_device_free(d)


def test_pinned_arrays(cvm):
    L = cvm.lib()
    n = 9 * MIB + 5
    a = cvm.host_array(n)
    b = cvm.host_array((n,), np.uint8)
    a[:] = np.random.default_rng(2).integers(0, 256, size=n, dtype=np.uint8)
    d = dalloc(L, n)
    try:
        assert L.cvm_copy_h2d(d, C.c_void_p(a.ctypes.data), n) == 0
        assert L.cvm_copy_d2h(C.c_void_p(b.ctypes.data), d, n) == 0
        assert np.array_equal(a, b)
        # a pageable source into the same device buffer, read back pinned
        c = np.random.default_rng(3).integers(0, 256, size=n, dtype=np.uint8)
        assert L.cvm_copy_h2d(d, C.c_void_p(c.ctypes.data), n) == 0
        assert L.cvm_copy_d2h(C.c_void_p(b.ctypes.data), d, n) == 0
        assert np.array_equal(c, b)
    finally:
        L.cvm_device_free(d)
    with pytest.raises(cvm.CvmatchError):
        cvm._check(L.cvm_host_free(C.c_void_p(c.ctypes.data)))


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


def test_threads_and_pinned_do_not_change_bits(threads):
    """4K RGBA MatchMap / Match, a Matcher and a Fleet on host buffers: the
    same bits at 1, 3 and 16 copy threads and from pinned inputs."""
    cvm = threads
    parent, sub, _ = synth.workload_map_4k(seed=5, alpha="random")
    ph, pw = parent.shape[:2]
    pin = cvm.host_array(parent.shape)
    pin[:] = parent
    m = cvm.NewMatcher(sub)
    fl = cvm.NewFleet(m, cvm.NewMatcher(parent[100:164, 900:980].copy()))
    ref = None
    for nthreads, src in [(1, parent), (3, parent), (16, parent), (16, pin), (0, parent[:, :])]:
        cvm.Threads = nthreads
        say(f"[xfer] 4K MatchMap, {nthreads} threads, pinned={src is pin}")
        res, w, h = cvm.match_map(src, sub)
        got = (bits(res), cvm.match(src, sub), m.find(src), [tuple(r) for r in fl.find_all(src)])
        if ref is None:
            ref = got
            continue
        assert np.array_equal(got[0], ref[0]), nthreads
        assert got[1:] == ref[1:], nthreads
    assert (w, h) == (pw - sub.shape[1] + 1, ph - sub.shape[0] + 1)
