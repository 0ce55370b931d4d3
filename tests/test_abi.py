"""The C-ABI library loads and exports every entry point include/*.h declares,
and its host-only planner agrees with the oracle -- CPU only, no compute calls."""
import ctypes
import glob
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\**\s*(cvm_[A-Za-z0-9_]+)\s*\(", text, re.M):
            names.append(m.group(1))
    return sorted(set(names))


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ["cvm_match_u8", "cvm_match", "cvm_match_map", "cvm_match_gray", "cvm_matcher_new",
                 "cvm_matcher_find", "cvm_fleet_new", "cvm_fleet_find_all", "cvm_last_error_detail"]:
        assert must in names


def test_library_exports_every_declared_symbol(cvm):
    lib = ctypes.CDLL(cvm.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    # the Python binding covers every declared entry point
    assert set(declared_functions()) <= set(cvm.SIGNATURES)


def test_planner_matches_oracle(cvm, orc):
    rng = np.random.default_rng(3)
    shapes = [(128, 128, 3713, 2033), (96, 32, 1825, 1049), (48, 48, 593, 433), (16, 16, 2545, 1425),
              (1, 1, 50, 40), (400, 1, 1, 2000), (1, 400, 2000, 1)]
    for _ in range(200):
        tw, th = (int(v) for v in rng.integers(1, 300, 2))
        rw, rh = (int(v) for v in rng.integers(1, 3000, 2))
        shapes.append((tw, th, rw, rh))
    for s in shapes:
        assert cvm.plan(*s) == orc.plan(*s), s


def test_stats_cap_matches(cvm, orc):
    for cn in range(1, 5):
        assert cvm.stats_cap(cn) == orc.stats_cap(cn)


def test_version(cvm):
    assert b"gfx950" in cvm.lib().cvm_version()


def test_threads_knob(cvm):
    """threads() cvmatch.go:52-63: n <= 0 is automatic, others clamp to
    [1, 16].  It sizes host copy threads only (host-side, no GPU call)."""
    L = cvm.lib()
    try:
        auto = L.cvm_set_threads(0)
        assert 1 <= auto <= 16
        assert L.cvm_set_threads(-7) == auto
        assert L.cvm_set_threads(5) == 5 and L.cvm_threads() == 5
        assert L.cvm_set_threads(100) == 16
        assert L.cvm_set_threads(1) == 1
    finally:
        L.cvm_set_threads(0)
        cvm._threads_pushed = None


def test_copy_pool_splits_exactly(cvm):
    """The host copy threads that feed the pinned staging chunks: every byte
    of [off, off+n) arrives, nothing outside it is written, for sizes around
    the piece boundaries and every thread count."""
    L = cvm.lib()
    rng = np.random.default_rng(8)
    src = rng.integers(0, 256, size=(6 << 20) + 4096, dtype=np.uint8)
    dst = np.empty_like(src)
    piece = 192 << 10
    sizes = [0, 1, 63, 64, 65, piece - 1, piece, piece + 1, 2 * piece + 7, (1 << 20) + 3, (4 << 20) + 5, 6 << 20]
    sizes += [int(v) for v in rng.integers(0, 6 << 20, 20)]
    for i, n in enumerate(sizes):
        th = 1 + i % 16
        off = int(rng.integers(0, 64))
        dst.fill(0xEE)
        rc = L.cvm_selftest_copy(ctypes.c_void_p(dst.ctypes.data + off), ctypes.c_void_p(src.ctypes.data + off), n, th)
        assert rc == 0
        assert np.array_equal(dst[off:off + n], src[off:off + n]), (n, th, off)
        assert not np.any(dst[:off] != 0xEE) and not np.any(dst[off + n:] != 0xEE), (n, th, off)


def test_peak_argument_errors(cvm):
    """Argument checks of the peak entry points fail before any GPU work."""
    L = cvm.lib()
    img = np.zeros((40, 50, 4), np.uint8)
    tpl = np.zeros((8, 8, 4), np.uint8)
    xy = np.zeros(16, np.int32)
    vals = np.zeros(8, np.float32)
    cnt = ctypes.c_int32()
    out = cvm.cvm_result()
    P = ctypes.c_void_p

    def call(rx=2, ry=2, cap=8, xyp=P(xy.ctypes.data), vp=P(vals.ctypes.data), cn=0, step=4, cntp=ctypes.byref(cnt)):
        return L.cvm_match_peaks(P(img.ctypes.data), 200, 50, 40, P(tpl.ctypes.data), 32, 8, 8, cn, step, 0.5, rx, ry,
                                 xyp, vp, cap, cntp, ctypes.byref(out))

    assert call(rx=-1) == cvm.CVM_E_BAD_ARGS
    assert call(ry=-3) == cvm.CVM_E_BAD_ARGS
    assert call(cap=-1) == cvm.CVM_E_BAD_ARGS
    assert call(xyp=None) == cvm.CVM_E_BAD_ARGS
    assert call(cntp=None) == cvm.CVM_E_BAD_ARGS
    assert call(cn=0, step=1) == cvm.CVM_E_BAD_ARGS
    assert L.cvm_find_peaks_device(None, 10, 10, 0.5, 1, 1, None, None, 0, ctypes.byref(cnt), None) == \
        cvm.CVM_E_NIL_IMAGE
    assert L.cvm_find_peaks_device(P(8), 0, 10, 0.5, 1, 1, None, None, 0, ctypes.byref(cnt), None) == \
        cvm.CVM_E_EMPTY_IMAGE


def test_tuner_nested_measurement_does_not_deadlock(cvm, tmp_path, monkeypatch):
    """A tuner measurement that consults the tuner for another key (the mixed
    form's crossover runs direct calls, which pick their own waves) must not
    wait on the tuner's lock: run in a thread with a deadline."""
    import threading

    monkeypatch.setenv("CVMATCH_TUNE_FILE", str(tmp_path / "tune.txt"))
    L = cvm.lib()
    out = []
    t = threading.Thread(target=lambda: out.append(L.cvm_selftest_tuner()), daemon=True)
    t.start()
    t.join(timeout=20)
    assert not t.is_alive(), "the tuner deadlocked on a nested measurement"
    assert out == [2]
