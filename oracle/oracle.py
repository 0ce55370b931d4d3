"""ctypes binding of the CPU oracle (oracle/cvm_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
and by ``bench.py``'s cpu_baseline leg, never by the product package.  The
functions restate the reference's scalar op sequence (see the C file's header
for the file:line map) plus an FP64 exact-integer oracle.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libcvm_oracle.so")
_lib = None


class OrcResult(C.Structure):
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


ERRORS = {1: "bad match arguments", 2: "stats cap", 3: "out of memory", 4: "empty fleet", 5: "mixed fleet"}


class OracleError(RuntimeError):
    def __init__(self, code):
        super().__init__(ERRORS.get(code, f"error {code}"))
        self.code = code


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc)."""
    cmd = ["make", "-C", _HERE]
    if force:
        cmd.append("-B")
    subprocess.run(cmd, check=True, stdout=subprocess.DEVNULL)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        P = C.c_void_p
        i32, i64 = C.c_int32, C.c_int64
        L.orc_sincospi.argtypes = [C.c_int, C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_float)]
        L.orc_tri_len.argtypes = [C.c_int]
        L.orc_tri_len.restype = C.c_int
        L.orc_tri_twiddles.argtypes = [C.c_int, P]
        L.orc_tri_twiddles.restype = C.c_int
        L.orc_bitrev.argtypes = [C.c_int, P]
        L.orc_fft_r4.argtypes = [P, C.c_int, C.c_int]
        L.orc_cols_r4.argtypes = [P, C.c_int, C.c_int, C.c_int, C.c_int]
        L.orc_plan.argtypes = [C.c_int] * 4 + [P]
        L.orc_plan.restype = i64
        L.orc_shrink_dh.argtypes = [C.c_int] * 4
        L.orc_shrink_dh.restype = C.c_int
        L.orc_stats_cap.argtypes = [C.c_int]
        L.orc_stats_cap.restype = i64
        L.orc_templ_stats.argtypes = [P, i64, C.c_int, C.c_int, C.c_int, C.c_int, P, P]
        L.orc_match_u8.argtypes = [P, i64, C.c_int, C.c_int, P, i64, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, P,
                                   C.POINTER(OrcResult)]
        L.orc_match_u8.restype = C.c_int
        L.orc_match_rgba.argtypes = [P, i64, C.c_int, C.c_int, P, i64, C.c_int, C.c_int, C.c_int, P,
                                     C.POINTER(OrcResult)]
        L.orc_match_rgba.restype = C.c_int
        L.orc_matcher_find.argtypes = [P, i64, C.c_int, C.c_int, P, C.c_int, C.c_int, C.c_int, C.c_int, P,
                                       C.POINTER(OrcResult)]
        L.orc_matcher_find.restype = C.c_int
        L.orc_alpha_const.argtypes = [P, i64, C.c_int, C.c_int]
        L.orc_alpha_const.restype = C.c_int
        L.orc_rgba_to_gray.argtypes = [P, i64, C.c_int, C.c_int, P]
        L.orc_bucketize.argtypes = [P, P, P, C.c_int, P, P]
        L.orc_bucketize.restype = C.c_int
        L.orc_fleet_find_all.argtypes = [P, i64, C.c_int, C.c_int, C.c_int, P, P, P, C.c_int, P]
        L.orc_fleet_find_all.restype = C.c_int
        L.orc_fp64_map.argtypes = [P, i64, C.c_int, C.c_int, P, i64, C.c_int, C.c_int, C.c_int, C.c_int, P]
        L.orc_fp64_map.restype = C.c_int
        L.orc_fp64_points.argtypes = [P, i64, C.c_int, C.c_int, P, i64, C.c_int, C.c_int, C.c_int, C.c_int, P, P,
                                      C.c_int, P]
        L.orc_fp64_points.restype = C.c_int
        L.orc_golden_hash.argtypes = [P, C.c_size_t, C.POINTER(OrcResult)]
        L.orc_golden_hash.restype = C.c_uint64
        L.orc_find_peaks.argtypes = [P, C.c_int, C.c_int, C.c_float, C.c_int, C.c_int, P, P, C.c_int]
        L.orc_find_peaks.restype = C.c_int
        _lib = L
    return _lib


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def _u8(a) -> np.ndarray:
    a = np.asarray(a)
    assert a.dtype == np.uint8
    return a


# ------------------------------------------------------------------ twiddles


def sincospi(j: int, half: int):
    c, s = C.c_float(), C.c_float()
    lib().orc_sincospi(j, half, C.byref(c), C.byref(s))
    return np.float32(c.value), np.float32(s.value)


def tri_twiddles(n: int) -> np.ndarray:
    L = lib()
    m = L.orc_tri_len(n)
    out = np.zeros(max(m, 1), dtype=np.complex64)
    L.orc_tri_twiddles(n, _ptr(out))
    return out[:m]


def bitrev(n: int) -> np.ndarray:
    out = np.zeros(n, dtype=np.int32)
    lib().orc_bitrev(n, _ptr(out))
    return out


def fft_r4(x: np.ndarray, inverse: bool = False) -> np.ndarray:
    a = np.ascontiguousarray(x, dtype=np.complex64).copy()
    lib().orc_fft_r4(_ptr(a), a.size, int(inverse))
    return a


def cols_r4(d: np.ndarray, inverse: bool = False, use_rmap: bool = False) -> np.ndarray:
    a = np.ascontiguousarray(d, dtype=np.complex64).copy()
    n, width = a.shape
    lib().orc_cols_r4(_ptr(a), n, width, int(inverse), int(use_rmap))
    return a


# ---------------------------------------------------------------- planner


def plan(tw: int, th: int, rw: int, rh: int):
    out = np.zeros(4, dtype=np.int32)
    cost = lib().orc_plan(tw, th, rw, rh, _ptr(out))
    return tuple(int(v) for v in out), int(cost)


def shrink_dh(tw: int, th: int, rw: int, rh: int) -> int:
    return int(lib().orc_shrink_dh(tw, th, rw, rh))


def stats_cap(cn: int) -> int:
    return int(lib().orc_stats_cap(cn))


def templ_stats(tpl, tstride, tw, th, cn, step):
    tpl = _u8(tpl)
    tsum = np.zeros(4, dtype=np.int64)
    var = np.zeros(1, dtype=np.int64)
    lib().orc_templ_stats(_ptr(tpl), tstride, tw, th, cn, step, _ptr(tsum), _ptr(var))
    return tsum, int(var[0])


# ------------------------------------------------------------------ matching


def match_u8(img, istride, iw, ih, tpl, tstride, tw, th, cn, step, threads=1, want_map=True):
    """matchU8 restated.  Returns (map or None, six-tuple)."""
    img, tpl = _u8(img), _u8(tpl)
    rw, rh = iw - tw + 1, ih - th + 1
    res = np.empty(rw * rh, dtype=np.float32) if want_map and rw > 0 and rh > 0 else None
    r = OrcResult()
    err = lib().orc_match_u8(_ptr(img), istride, iw, ih, _ptr(tpl), tstride, tw, th, cn, step, threads,
                             _ptr(res) if res is not None else None, C.byref(r))
    if err:
        raise OracleError(err)
    return (res.reshape(rh, rw) if res is not None else None), r.tuple()


def match_rgba(parent: np.ndarray, sub: np.ndarray, threads=1, want_map=True):
    """Match / MatchMap restated on (H, W, 4) uint8 arrays."""
    parent, sub = np.ascontiguousarray(_u8(parent)), np.ascontiguousarray(_u8(sub))
    ih, iw = parent.shape[:2]
    th, tw = sub.shape[:2]
    rw, rh = iw - tw + 1, ih - th + 1
    res = np.empty(rw * rh, dtype=np.float32) if want_map else None
    r = OrcResult()
    err = lib().orc_match_rgba(_ptr(parent), iw * 4, iw, ih, _ptr(sub), tw * 4, tw, th, threads,
                               _ptr(res) if res is not None else None, C.byref(r))
    if err:
        raise OracleError(err)
    return (res.reshape(rh, rw) if res is not None else None), r.tuple()


def match_gray(parent: np.ndarray, sub: np.ndarray, threads=1, want_map=True):
    """MatchGray restated on (H, W) uint8 arrays (already gray)."""
    parent, sub = np.ascontiguousarray(_u8(parent)), np.ascontiguousarray(_u8(sub))
    ih, iw = parent.shape
    th, tw = sub.shape
    return match_u8(parent, iw, iw, ih, sub, tw, tw, th, 1, 1, threads, want_map)


def matcher_find(parent: np.ndarray, sub: np.ndarray, threads=1, want_map=False):
    """Matcher.Find restated: step 4 for (H, W, 4) templates, 1 for (H, W)."""
    parent, sub = np.ascontiguousarray(_u8(parent)), np.ascontiguousarray(_u8(sub))
    step = 4 if sub.ndim == 3 else 1
    ih, iw = parent.shape[:2]
    th, tw = sub.shape[:2]
    rw, rh = iw - tw + 1, ih - th + 1
    res = np.empty(rw * rh, dtype=np.float32) if want_map else None
    r = OrcResult()
    err = lib().orc_matcher_find(_ptr(parent), iw * step, iw, ih, _ptr(sub), tw, th, step, threads,
                                 _ptr(res) if res is not None else None, C.byref(r))
    if err:
        raise OracleError(err)
    return (res.reshape(rh, rw) if res is not None else None), r.tuple()


def alpha_const(pix: np.ndarray) -> bool:
    pix = np.ascontiguousarray(_u8(pix))
    h, w = pix.shape[:2]
    return bool(lib().orc_alpha_const(_ptr(pix), w * 4, w, h))


def rgba_to_gray(pix: np.ndarray) -> np.ndarray:
    pix = np.ascontiguousarray(_u8(pix))
    h, w = pix.shape[:2]
    out = np.empty((h, w), dtype=np.uint8)
    lib().orc_rgba_to_gray(_ptr(pix), w * 4, w, h, _ptr(out))
    return out


def bucketize(sizes, flat=None):
    n = len(sizes)
    ws = np.array([s[0] for s in sizes], dtype=np.int32)
    hs = np.array([s[1] for s in sizes], dtype=np.int32)
    fl = np.array(flat if flat is not None else [0] * n, dtype=np.int32)
    dims = np.zeros(3 * max(n, 1), dtype=np.int32)
    mem = np.zeros(max(n, 1), dtype=np.int32)
    nb = lib().orc_bucketize(_ptr(ws), _ptr(hs), _ptr(fl), n, _ptr(dims), _ptr(mem))
    out, at = [], 0
    for b in range(nb):
        tw, th, cnt = (int(v) for v in dims[3 * b:3 * b + 3])
        out.append((tw, th, [int(m) for m in mem[at:at + cnt]]))
        at += cnt
    return out


def fleet_find_all(parent: np.ndarray, subs):
    """Fleet.FindAll restated; subs are (h, w, 4) or (h, w) uint8 arrays."""
    parent = np.ascontiguousarray(_u8(parent))
    step = 4 if parent.ndim == 3 else 1
    ih, iw = parent.shape[:2]
    subs = [np.ascontiguousarray(_u8(s)) for s in subs]
    n = len(subs)
    ptrs = (C.c_void_p * max(n, 1))(*[s.ctypes.data for s in subs])
    tws = np.array([s.shape[1] for s in subs], dtype=np.int32)
    ths = np.array([s.shape[0] for s in subs], dtype=np.int32)
    out = (OrcResult * max(n, 1))()
    err = lib().orc_fleet_find_all(_ptr(parent), iw * step, iw, ih, step, C.cast(ptrs, C.c_void_p), _ptr(tws),
                                   _ptr(ths), n, C.cast(out, C.c_void_p))
    if err:
        raise OracleError(err)
    return [out[i].tuple() for i in range(n)]


def fp64_map(img, istride, iw, ih, tpl, tstride, tw, th, cn, step) -> np.ndarray:
    img, tpl = _u8(img), _u8(tpl)
    rw, rh = iw - tw + 1, ih - th + 1
    out = np.empty(rw * rh, dtype=np.float64)
    err = lib().orc_fp64_map(_ptr(img), istride, iw, ih, _ptr(tpl), tstride, tw, th, cn, step, _ptr(out))
    if err:
        raise OracleError(err)
    return out.reshape(rh, rw)


def fp64_points(img, istride, iw, ih, tpl, tstride, tw, th, cn, step, xs, ys) -> np.ndarray:
    img, tpl = _u8(img), _u8(tpl)
    xs = np.ascontiguousarray(xs, dtype=np.int32)
    ys = np.ascontiguousarray(ys, dtype=np.int32)
    out = np.empty(xs.size, dtype=np.float64)
    err = lib().orc_fp64_points(_ptr(img), istride, iw, ih, _ptr(tpl), tstride, tw, th, cn, step, _ptr(xs), _ptr(ys),
                                xs.size, _ptr(out))
    if err:
        raise OracleError(err)
    return out


def _window_sums(a: np.ndarray, th: int, tw: int) -> np.ndarray:
    """Exact int64 sums of every th x tw window of a (h, w) int64 array."""
    c = np.zeros((a.shape[0] + 1, a.shape[1] + 1), dtype=np.int64)
    np.cumsum(np.cumsum(a, axis=0), axis=1, out=c[1:, 1:])
    return c[th:, tw:] - c[:-th, tw:] - c[th:, :-tw] + c[:-th, :-tw]


def fp64_map_fft(parent: np.ndarray, sub: np.ndarray, cn: int) -> np.ndarray:
    """The FP64 oracle's full map (same values, bit for bit, as fp64_map) at
    sizes the brute force cannot reach: the window x template sums are the
    exact integers, obtained by rounding a float64 FFT correlation whose
    rounding residue is asserted to stay below 1/8 (the products are < 2^32,
    so float64 leaves a margin of many orders); window sums come from exact
    int64 integral images; the score is the C oracle's fp64_score, op for op
    (the TM_CCOEFF_NORMED guard ladder of cvmatch_test.go:57-76).
    parent / sub: (h, w) gray (cn 1) or (h, w, 4) RGBA (first cn channels)."""
    p = parent if parent.ndim == 3 else parent[..., None]
    t = sub if sub.ndim == 3 else sub[..., None]
    ih, iw = p.shape[:2]
    th, tw = t.shape[:2]
    rh, rw = ih - th + 1, iw - tw + 1
    area = tw * th
    spec = None
    s = []
    s2 = np.zeros((rh, rw), dtype=np.int64)
    tsum = []
    for k in range(cn):
        pk = p[..., k].astype(np.int64)
        tk = t[..., k].astype(np.int64)
        fp = np.fft.rfft2(pk.astype(np.float64))
        ft = np.fft.rfft2(tk.astype(np.float64), s=(ih, iw))
        spec = fp * np.conj(ft) if spec is None else spec + fp * np.conj(ft)
        s.append(_window_sums(pk, th, tw))
        s2 += _window_sums(pk * pk, th, tw)
        tsum.append(int(tk.sum()))
    cf = np.fft.irfft2(spec, s=(ih, iw))[:rh, :rw]
    corr = np.rint(cf)
    resid = float(np.max(np.abs(cf - corr))) if corr.size else 0.0
    assert resid < 0.125, f"float64 correlation residue {resid} too close to 1/2 for exact rounding"
    corr = corr.astype(np.int64)
    tt = [t[..., k].astype(np.int64) for k in range(cn)]
    varSum = sum(area * int((x * x).sum()) - int(x.sum()) ** 2 for x in tt)
    if varSum == 0:
        return np.ones((rh, rw), dtype=np.float64)
    cross = np.zeros((rh, rw), dtype=np.int64)
    sq = np.zeros((rh, rw), dtype=np.int64)
    for k in range(cn):
        cross += s[k] * tsum[k]
        sq += s[k] * s[k]
    numI = area * corr - cross
    idiff = area * s2 - sq
    diff2 = idiff.astype(np.float64) / float(area)
    lim = np.minimum(10.0 * 1.1920928955078125e-07 * s2.astype(np.float64), 0.5)
    denI = np.sqrt(idiff.astype(np.float64)) * np.sqrt(float(varSum))
    nf = numI.astype(np.float64)
    an = np.abs(nf)
    with np.errstate(divide="ignore", invalid="ignore"):
        q = nf / denI
    out = np.where(an < denI, q, np.where(an < denI * 1.125, np.where(numI > 0, 1.0, -1.0), 0.0))
    out[~(diff2 > lim)] = 0.0
    return out


def golden_hash(res: np.ndarray, tup) -> int:
    res = np.ascontiguousarray(res, dtype=np.float32).ravel()
    r = OrcResult(*tup)
    return int(lib().orc_golden_hash(_ptr(res), res.size, C.byref(r)))


def find_peaks(resp: np.ndarray, thresh: float, rx: int, ry: int, cap: int | None = None):
    """Threshold + local-maximum suppression (cvmatch_test.go:346-365) on a
    (h, w) float32 map: (count, xy[n, 2], vals[n]) in raster order, n =
    min(count, cap)."""
    resp = np.ascontiguousarray(resp, dtype=np.float32)
    h, w = resp.shape
    if cap is None:
        cap = w * h
    xy = np.zeros((max(cap, 1), 2), dtype=np.int32)
    vals = np.zeros(max(cap, 1), dtype=np.float32)
    n = lib().orc_find_peaks(_ptr(resp), w, h, float(thresh), rx, ry, _ptr(xy), _ptr(vals), cap)
    k = min(n, cap)
    return n, xy[:k], vals[:k]
