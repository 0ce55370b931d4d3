"""cvmatch_amd -- TM_CCOEFF_NORMED template matching on MI355X (gfx950).

A Python mirror of the reference Go API (cvmatch.go, matcher.go, fleet.go)
over the C ABI of ``libcvmatch_hip.so`` (declared in include/cvmatch_hip.h):

    match(parent, sub)      -> (minV, minX, minY, maxV, maxX, maxY)   Match cvmatch.go:75
    match_map(parent, sub)  -> (res[h, w] float32, w, h)              MatchMap cvmatch.go:100
    match_gray(parent, sub) -> six-tuple                              MatchGray cvmatch.go:128
    Matcher(sub) / GrayMatcher(sub) .find(parent)                    matcher.go:45-130
    Fleet(*matchers).find_all(parent) -> [Result]                    fleet.go:85-332

Images are numpy uint8 arrays: (H, W, 4) interleaved RGBA or (H, W) gray.
Row-strided views (``parent[y0:y1, x0:x1]``) are passed zero-copy with their
row stride, like the reference's sub-images.  Errors raise ``CvmatchError``
carrying the reference's panic text.  Every call runs on the GPU; there is
no CPU path: if the HIP library is not built, the import of any entry point
fails loudly.
"""
from __future__ import annotations

import ctypes as C
import os
import weakref
from collections import namedtuple

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcvmatch_hip.so")
_lib = None

Result = namedtuple("Result", "MinV MinX MinY MaxV MaxX MaxY")

CVM_OK, CVM_E_BAD_ARGS, CVM_E_STATS_CAP, CVM_E_EMPTY_FLEET, CVM_E_MIXED_FLEET = 0, 1, 2, 3, 4
CVM_E_HIP, CVM_E_UNSUPPORTED, CVM_E_NIL_IMAGE, CVM_E_EMPTY_IMAGE = 5, 6, 7, 8

# Worker-count knob of the reference (cvmatch.go:36-63), with its
# CVMATCH_THREADS environment variable read once at import.  Here it sizes the
# host threads that copy pageable buffers through pinned staging (0 or less:
# automatic; clamped to [1, 16]); it never changes an output bit.
try:
    Threads = max(int(os.environ.get("CVMATCH_THREADS", "0")), 0)
except ValueError:
    Threads = 0
_threads_pushed = None


class CvmatchError(RuntimeError):
    def __init__(self, code: int, detail: str):
        super().__init__(detail)
        self.code = code


class cvm_result(C.Structure):
    _fields_ = [
        ("min_v", C.c_float),
        ("min_x", C.c_int32),
        ("min_y", C.c_int32),
        ("max_v", C.c_float),
        ("max_x", C.c_int32),
        ("max_y", C.c_int32),
    ]

    def tuple(self):
        return (self.min_v, self.min_x, self.min_y, self.max_v, self.max_x, self.max_y)


# exported symbols and their ctypes signatures
_P, _I32, _I64, _SZ = C.c_void_p, C.c_int32, C.c_int64, C.c_size_t
_RES = C.POINTER(cvm_result)
SIGNATURES = {
    "cvm_last_error_detail": ([], C.c_char_p),
    "cvm_version": ([], C.c_char_p),
    "cvm_match_u8": ([_P, _I64, _I32, _I32, _P, _I64, _I32, _I32, _I32, _I32, _P, _RES], C.c_int),
    "cvm_match": ([_P, _I64, _I32, _I32, _P, _I64, _I32, _I32, _RES], C.c_int),
    "cvm_match_map": ([_P, _I64, _I32, _I32, _P, _I64, _I32, _I32, _P, _RES], C.c_int),
    "cvm_match_gray": ([_P, _I64, _I32, _I32, _I32, _P, _I64, _I32, _I32, _I32, _RES], C.c_int),
    "cvm_alpha_const": ([_P, _I64, _I32, _I32, C.POINTER(_I32)], C.c_int),
    "cvm_rgba_to_gray": ([_P, _I64, _I32, _I32, _P], C.c_int),
    "cvm_stats_cap": ([_I32], _I64),
    "cvm_plan": ([_I32, _I32, _I32, _I32, _P], _I64),
    "cvm_matcher_new": ([_P, _I64, _I32, _I32, _I32, C.POINTER(_P)], C.c_int),
    "cvm_matcher_find": ([_P, _P, _I64, _I32, _I32, _RES], C.c_int),
    "cvm_matcher_free": ([_P], None),
    "cvm_fleet_new": ([_P, _I32, C.POINTER(_P)], C.c_int),
    "cvm_fleet_find_all": ([_P, _P, _I64, _I32, _I32, _P], C.c_int),
    "cvm_fleet_free": ([_P], None),
    "cvm_match_u8_device": ([_P, _I64, _I32, _I32, _P, _I64, _I32, _I32, _I32, _I32, _P, _RES, _P], C.c_int),
    "cvm_match_u8_direct": ([_P, _I64, _I32, _I32, _P, _I64, _I32, _I32, _I32, _I32, _P, _RES], C.c_int),
    "cvm_match_u8_direct_device": ([_P, _I64, _I32, _I32, _P, _I64, _I32, _I32, _I32, _I32, _P, _RES, _P],
                                   C.c_int),
    "cvm_match_u8_direct_device_noalloc": ([_P, _I64, _I32, _I32, _P, _I64, _I32, _I32, _I32, _I32, _P, _RES, _P],
                                           C.c_int),
    "cvm_fleet_find_all_mixed": ([_P, _P, _I64, _I32, _I32, _P], C.c_int),
    "cvm_fleet_find_all_mixed_device": ([_P, _P, _I64, _I32, _I32, _P, _P], C.c_int),
    "cvm_fleet_mixed_plan": ([_P, _P, _P, _I32, C.POINTER(_I32)], C.c_int),
    "cvm_tune_report": ([_P, _I32], C.c_int),
    "cvm_matcher_find_direct": ([_P, _P, _I64, _I32, _I32, _RES], C.c_int),
    "cvm_fleet_find_all_direct": ([_P, _P, _I64, _I32, _I32, _P], C.c_int),
    "cvm_device_alloc": ([_SZ, C.POINTER(_P)], C.c_int),
    "cvm_device_free": ([_P], C.c_int),
    "cvm_copy_h2d": ([_P, _P, _SZ], C.c_int),
    "cvm_copy_d2h": ([_P, _P, _SZ], C.c_int),
    "cvm_synchronize": ([], C.c_int),
    "cvm_device_copy": ([_P, _P, _SZ], C.c_int),
    "cvm_profile_enable": ([_I32], C.c_int),
    "cvm_profile_reset": ([], C.c_int),
    "cvm_profile_read": ([_P, _P, _P, _I32], C.c_int),
    "cvm_selftest_fp": ([_P, _P, _P, _I32, _P, _P, _P], C.c_int),
    "cvm_twiddles": ([_I32, _P, _P, C.POINTER(_I32)], C.c_int),
    "cvm_match_peaks": ([_P, _I64, _I32, _I32, _P, _I64, _I32, _I32, _I32, _I32, C.c_float, _I32, _I32, _P, _P,
                         _I32, C.POINTER(_I32), _RES], C.c_int),
    "cvm_find_peaks_device": ([_P, _I32, _I32, C.c_float, _I32, _I32, _P, _P, _I32, C.POINTER(_I32), _P], C.c_int),
    "cvm_matcher_find_device": ([_P, _P, _I64, _I32, _I32, _RES, _P], C.c_int),
    "cvm_fleet_find_all_device": ([_P, _P, _I64, _I32, _I32, _P, _P], C.c_int),
    "cvm_fleet_find_all_direct_device": ([_P, _P, _I64, _I32, _I32, _P, _P], C.c_int),
    "cvm_rgba_to_gray_device": ([_P, _I64, _I32, _I32, _P, _P], C.c_int),
    "cvm_frames_new": ([_I32, C.POINTER(_P)], C.c_int),
    "cvm_frames_find": ([_P, _P, _P, _I64, _I32, _I32, _I32, _I64], C.c_int),
    "cvm_frames_result": ([_P, C.POINTER(_I64), _RES], C.c_int),
    "cvm_frames_pending": ([_P], _I32),
    "cvm_frames_free": ([_P], None),
    "cvm_set_threads": ([_I32], _I32),
    "cvm_threads": ([], _I32),
    "cvm_host_alloc": ([_SZ, C.POINTER(_P)], C.c_int),
    "cvm_host_free": ([_P], C.c_int),
    "cvm_selftest_copy": ([_P, _P, _SZ, _I32], C.c_int),
    "cvm_selftest_tuner": ([], C.c_int),
}


def lib():
    """Load libcvmatch_hip.so (built by cvmatch_amd/build.py).  Raises if absent."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run `python cvmatch_amd/build.py` (hipcc, gfx950)")
        L = C.CDLL(LIB_PATH)
        for name, (args, res) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def _L():
    """lib(), with the module's Threads pushed to the library when it changed
    (the reference reads its Threads variable at every call)."""
    global _threads_pushed
    L = lib()
    if Threads != _threads_pushed:
        L.cvm_set_threads(int(Threads))
        _threads_pushed = Threads
    return L


def threads() -> int:
    """threads() cvmatch.go:52-63: the resolved copy-thread count."""
    return int(_L().cvm_threads())


def host_array(shape, dtype=np.uint8) -> np.ndarray:
    """A numpy array in pinned host memory (cvm_host_alloc), freed with the
    array.  Frames filled here move to the GPU at the DMA rate, unstaged."""
    dt = np.dtype(dtype)
    count = int(np.prod(shape))
    nbytes = max(count * dt.itemsize, 1)
    p = C.c_void_p()
    _check(lib().cvm_host_alloc(nbytes, C.byref(p)))
    buf = (C.c_uint8 * nbytes).from_address(p.value)
    fin = weakref.finalize(buf, lib().cvm_host_free, C.c_void_p(p.value))
    fin.atexit = False  # the process's exit releases it
    return np.frombuffer(buf, dtype=dt, count=count).reshape(shape)


def tune_report() -> dict:
    """The autotuner's decisions so far (cvm_tune_report)."""
    import json

    L = lib()
    n = L.cvm_tune_report(None, 0)
    buf = C.create_string_buffer(n + 1)
    L.cvm_tune_report(C.cast(buf, C.c_void_p), n + 1)
    return json.loads(buf.value.decode() or "{}")


def _check(rc: int):
    if rc != CVM_OK:
        raise CvmatchError(rc, lib().cvm_last_error_detail().decode())


def _ptr(a: np.ndarray):
    return C.c_void_p(a.ctypes.data)


def view(img: np.ndarray):
    """(pointer-owner array, stride, w, h, step) of an RGBA (H,W,4) or gray (H,W)
    uint8 image.  Row-strided views are kept zero-copy; anything else is made
    contiguous."""
    if img is None:
        raise CvmatchError(CVM_E_NIL_IMAGE, "cvmatch: nil image")
    img = np.asarray(img)
    if img.dtype != np.uint8 or img.ndim not in (2, 3) or (img.ndim == 3 and img.shape[2] != 4):
        raise CvmatchError(CVM_E_BAD_ARGS, "cvmatch: images are (H, W, 4) RGBA or (H, W) gray uint8")
    h, w = img.shape[:2]
    if w <= 0 or h <= 0:
        raise CvmatchError(CVM_E_EMPTY_IMAGE, "cvmatch: empty image")
    step = 4 if img.ndim == 3 else 1
    inner_ok = img.strides[1] == step and (img.ndim == 2 or img.strides[2] == 1)
    if not inner_ok or img.strides[0] < w * step:
        img = np.ascontiguousarray(img)
    return img, int(img.strides[0]), w, h, step


# ---------------------------------------------------------------- one-shot


def match_u8(img, istride, iw, ih, tpl, tstride, tw, th, cn, step, want_map=False, direct=False):
    """matchU8 impl.go:1126 on flat uint8 buffers; returns (map or None, six-tuple).

    direct=True: the direct form (exact int8-MFMA correlation, FP64 scores
    rounded to float32); cn=0 then selects Match's RGBA alpha rule."""
    img = np.ascontiguousarray(img, dtype=np.uint8)
    tpl = np.ascontiguousarray(tpl, dtype=np.uint8)
    rw, rh = iw - tw + 1, ih - th + 1
    res = np.empty((rh, rw), dtype=np.float32) if want_map and rw > 0 and rh > 0 else None
    out = cvm_result()
    fn = _L().cvm_match_u8_direct if direct else _L().cvm_match_u8
    _check(fn(_ptr(img), istride, iw, ih, _ptr(tpl), tstride, tw, th, cn, step,
              _ptr(res) if res is not None else None, C.byref(out)))
    return res, out.tuple()


def match_direct(parent, sub, want_map=False):
    """The direct form of Match / MatchGray on (H, W, 4) RGBA (alpha rule) or
    (H, W) gray arrays: (map or None, six-tuple)."""
    p, ps, pw, ph, pst = view(parent)
    s, ss, sw, sh, sst = view(sub)
    if pst != sst:
        raise CvmatchError(CVM_E_BAD_ARGS, "cvmatch: parent and template formats differ")
    if sw > pw or sh > ph:
        raise CvmatchError(CVM_E_BAD_ARGS, f"cvmatch: bad match arguments ({sw}x{sh} in {pw}x{ph})")
    rw, rh = pw - sw + 1, ph - sh + 1
    res = np.empty((rh, rw), dtype=np.float32) if want_map else None
    out = cvm_result()
    _check(_L().cvm_match_u8_direct(_ptr(p), ps, pw, ph, _ptr(s), ss, sw, sh, 0 if pst == 4 else 1, pst,
                                     _ptr(res) if res is not None else None, C.byref(out)))
    return res, out.tuple()


def match(parent, sub):
    """Match cvmatch.go:75: RGBA TM_CCOEFF_NORMED + minMaxLoc six-tuple."""
    p, ps, pw, ph, pst = view(parent)
    s, ss, sw, sh, sst = view(sub)
    if pst != 4 or sst != 4:
        raise CvmatchError(CVM_E_BAD_ARGS, "cvmatch: Match takes RGBA (H, W, 4) images")
    out = cvm_result()
    _check(_L().cvm_match(_ptr(p), ps, pw, ph, _ptr(s), ss, sw, sh, C.byref(out)))
    return out.tuple()


def match_map(parent, sub):
    """MatchMap cvmatch.go:100: (res[h, w] float32, w, h)."""
    p, ps, pw, ph, pst = view(parent)
    s, ss, sw, sh, sst = view(sub)
    if pst != 4 or sst != 4:
        raise CvmatchError(CVM_E_BAD_ARGS, "cvmatch: MatchMap takes RGBA (H, W, 4) images")
    if sw > pw or sh > ph:
        raise CvmatchError(CVM_E_BAD_ARGS,
                           f"cvmatch: bad match arguments ({sw}x{sh} in {pw}x{ph}, cn=4 step=4)")
    w, h = pw - sw + 1, ph - sh + 1
    res = np.empty((h, w), dtype=np.float32)
    out = cvm_result()
    _check(_L().cvm_match_map(_ptr(p), ps, pw, ph, _ptr(s), ss, sw, sh, _ptr(res), C.byref(out)))
    return res, w, h


def match_peaks(parent, sub, thresh=0.999, rx=None, ry=None, cap=4096):
    """Every occurrence of sub in parent (README.md:82-90): threshold + local-
    maximum suppression over the match map, the recipe TestFindAllOccurrences
    (cvmatch_test.go:320-381) verifies, with the map kept on the GPU.  RGBA
    inputs take Match's alpha rule, gray inputs MatchGray's single channel.
    rx / ry default to half the template (the test's 16 x 12 for 32 x 24).
    Returns (peaks, total, six-tuple): peaks = [(x, y, score)] in raster
    order, at most cap of them; total counts them all."""
    p, ps, pw, ph, pst = view(parent)
    s, ss, sw, sh, sst = view(sub)
    if pst != sst:
        raise CvmatchError(CVM_E_BAD_ARGS, "cvmatch: parent and template must both be RGBA or both gray")
    rx = sw // 2 if rx is None else rx
    ry = sh // 2 if ry is None else ry
    xy = np.zeros((max(cap, 1), 2), dtype=np.int32)
    vals = np.zeros(max(cap, 1), dtype=np.float32)
    count = C.c_int32()
    out = cvm_result()
    cn = 0 if pst == 4 else 1
    _check(_L().cvm_match_peaks(_ptr(p), ps, pw, ph, _ptr(s), ss, sw, sh, cn, pst, float(thresh), rx, ry,
                                _ptr(xy), _ptr(vals), cap, C.byref(count), C.byref(out)))
    k = min(count.value, cap)
    return [(int(xy[i, 0]), int(xy[i, 1]), float(vals[i])) for i in range(k)], count.value, out.tuple()


def find_peaks_device(d_map, w, h, thresh, rx, ry, d_xy, d_vals, cap, stream=None) -> int:
    """Peaks of a device-resident w x h map into device buffers d_xy (2*cap
    int32) / d_vals (cap float32); returns the total count."""
    count = C.c_int32()
    _check(_L().cvm_find_peaks_device(d_map, w, h, float(thresh), rx, ry, d_xy, d_vals, cap, C.byref(count),
                                      stream))
    return count.value


def match_gray(parent, sub):
    """MatchGray cvmatch.go:128: gray inputs as is, RGBA inputs via BT.601."""
    p, ps, pw, ph, pst = view(parent)
    s, ss, sw, sh, sst = view(sub)
    out = cvm_result()
    _check(_L().cvm_match_gray(_ptr(p), ps, pw, ph, pst, _ptr(s), ss, sw, sh, sst, C.byref(out)))
    return out.tuple()


def alpha_const(pix) -> bool:
    """alphaConst cvmatch.go:146."""
    p, ps, w, h, st = view(pix)
    if st != 4:
        raise CvmatchError(CVM_E_BAD_ARGS, "cvmatch: alpha scan needs RGBA")
    out = C.c_int32()
    _check(_L().cvm_alpha_const(_ptr(p), ps, w, h, C.byref(out)))
    return bool(out.value)


def rgba_to_gray(pix) -> np.ndarray:
    """rgbToGray cvmatch.go:222."""
    p, ps, w, h, st = view(pix)
    if st != 4:
        raise CvmatchError(CVM_E_BAD_ARGS, "cvmatch: gray conversion needs RGBA")
    out = np.empty((h, w), dtype=np.uint8)
    _check(_L().cvm_rgba_to_gray(_ptr(p), ps, w, h, _ptr(out)))
    return out


def stats_cap(cn: int) -> int:
    return int(lib().cvm_stats_cap(cn))


def plan(tw, th, rw, rh):
    out = np.zeros(4, dtype=np.int32)
    cost = lib().cvm_plan(tw, th, rw, rh, _ptr(out))
    return tuple(int(v) for v in out), int(cost)


# ----------------------------------------------------------------- Matcher


class Matcher:
    """NewMatcher matcher.go:45 (RGBA) / NewGrayMatcher matcher.go:57 (gray=True).

    ``find(parent)`` is bit-identical to ``match`` / ``match_gray``."""

    def __init__(self, sub, gray: bool = False):
        s, ss, sw, sh, sst = view(sub)
        if gray and sst == 4:
            s = rgba_to_gray(s)
            ss, sst = sw, 1
        if not gray and sst != 4:
            raise CvmatchError(CVM_E_BAD_ARGS, "cvmatch: NewMatcher takes an RGBA template")
        self.w, self.h, self.step = sw, sh, sst
        h = C.c_void_p()
        _check(_L().cvm_matcher_new(_ptr(s), ss, sw, sh, sst, C.byref(h)))
        self._h = h

    def find(self, parent, direct=False):
        p, ps, pw, ph, pst = view(parent)
        if self.step == 1 and pst == 4:
            p = rgba_to_gray(p)
            ps, pst = pw, 1
        if pst != self.step:
            raise CvmatchError(CVM_E_BAD_ARGS, "cvmatch: parent format does not match the matcher")
        out = cvm_result()
        fn = _L().cvm_matcher_find_direct if direct else _L().cvm_matcher_find
        _check(fn(self._h, _ptr(p), ps, pw, ph, C.byref(out)))
        return out.tuple()

    Find = find

    def find_device(self, d_img, stride, w, h, stream=None):
        """Find on a frame (or a region of one) already in device memory:
        d_img is the device address of the view's first pixel, stride its
        row pitch in bytes; the view has the matcher's format."""
        out = cvm_result()
        _check(_L().cvm_matcher_find_device(self._h, C.c_void_p(int(d_img)), stride, w, h, C.byref(out), stream))
        return out.tuple()

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and _lib is not None:
            _lib.cvm_matcher_free(h)
            self._h = None


def NewMatcher(sub) -> Matcher:
    return Matcher(sub, gray=False)


def NewGrayMatcher(sub) -> Matcher:
    return Matcher(sub, gray=True)


# ------------------------------------------------------------------- Fleet


class Fleet:
    """NewFleet fleet.go:85: many prepared matchers against one frame."""

    def __init__(self, *ms: Matcher):
        if len(ms) == 1 and isinstance(ms[0], (list, tuple)):
            ms = tuple(ms[0])
        self._ms = list(ms)
        arr = (C.c_void_p * max(len(ms), 1))(*[m._h.value for m in ms])
        h = C.c_void_p()
        _check(_L().cvm_fleet_new(C.cast(arr, C.c_void_p), len(ms), C.byref(h)))
        self._h = h
        self.step = ms[0].step

    def find_all(self, parent, direct=False, mixed=False):
        """FindAll; direct=True: the direct form, mixed=True: each member on
        the form the autotuner picked for this frame geometry."""
        p, ps, pw, ph, pst = view(parent)
        if self.step == 1 and pst == 4:
            p = rgba_to_gray(p)
            ps, pst = pw, 1
        if pst != self.step:
            raise CvmatchError(CVM_E_BAD_ARGS, "cvmatch: parent format does not match the fleet")
        n = len(self._ms)
        out = (cvm_result * n)()
        L = _L()
        fn = L.cvm_fleet_find_all_mixed if mixed else L.cvm_fleet_find_all_direct if direct else L.cvm_fleet_find_all
        _check(fn(self._h, _ptr(p), ps, pw, ph, C.cast(out, C.c_void_p)))
        return [Result(*out[i].tuple()) for i in range(n)]

    FindAll = find_all

    def find_all_device(self, d_img, stride, w, h, direct=False, stream=None, mixed=False):
        """FindAll on a frame (or a region of one) already in device memory:
        d_img is the device address of the view's first pixel, stride its
        row pitch in bytes; the view has the fleet's format."""
        n = len(self._ms)
        out = (cvm_result * n)()
        L = _L()
        fn = (L.cvm_fleet_find_all_mixed_device if mixed else
              L.cvm_fleet_find_all_direct_device if direct else L.cvm_fleet_find_all_device)
        _check(fn(self._h, C.c_void_p(int(d_img)), stride, w, h, C.cast(out, C.c_void_p), stream))
        return [Result(*out[i].tuple()) for i in range(n)]

    def mixed_plan(self):
        """The last mixed FindAll's routing: (routed flags, gate deviations,
        crossover side) -- see cvm_fleet_mixed_plan."""
        n = len(self._ms)
        routed = (C.c_int32 * n)()
        gate = (C.c_float * n)()
        xo = C.c_int32()
        _check(_L().cvm_fleet_mixed_plan(self._h, C.cast(routed, C.c_void_p), C.cast(gate, C.c_void_p), n,
                                         C.byref(xo)))
        return [bool(v) for v in routed], [float(v) for v in gate], int(xo.value)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and _lib is not None:
            _lib.cvm_fleet_free(h)
            self._h = None


class FrameStream:
    """Find over a stream of frames with the upload of each frame overlapping
    the matching of the previous one (cvm_frames_*): ``find`` submits and
    returns once the frame has been copied out of its buffer; ``result``
    returns (tag, six-tuple) of the oldest submitted frame.  Each result
    equals ``matcher.find(frame)``."""

    def __init__(self, depth: int = 2):
        h = C.c_void_p()
        _check(_L().cvm_frames_new(depth, C.byref(h)))
        self._h = h
        self._keep = {}

    def find(self, matcher: Matcher, frame, tag: int = 0, direct: bool = False):
        p, ps, pw, ph, pst = view(frame)
        if matcher.step == 1 and pst == 4:
            p = rgba_to_gray(p)
            ps, pst = pw, 1
        if pst != matcher.step:
            raise CvmatchError(CVM_E_BAD_ARGS, "cvmatch: parent format does not match the matcher")
        _check(_L().cvm_frames_find(self._h, matcher._h, _ptr(p), ps, pw, ph, int(direct), int(tag)))
        self._keep[id(matcher)] = matcher  # the Python object outlives its frames

    def result(self):
        tag = C.c_int64()
        out = cvm_result()
        _check(_L().cvm_frames_result(self._h, C.byref(tag), C.byref(out)))
        return int(tag.value), out.tuple()

    def pending(self) -> int:
        return int(_L().cvm_frames_pending(self._h))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and _lib is not None:
            _lib.cvm_frames_free(h)
            self._h = None


def NewFleet(*ms) -> Fleet:
    return Fleet(*ms)


Match, MatchMap, MatchGray = match, match_map, match_gray
