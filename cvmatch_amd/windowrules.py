"""Window rules: many templates, grouped by region of interest, against a
stream of captured window frames -- the consumer of the Fleet that the
reference's windowrules example builds (examples/windowrules/engine.go).

Same rules, same verdicts: templates deduplicated per (template id, gray),
one Fleet per (ROI, gray) group, a memoised sub-fleet for view sizes that
not every member fits, ModeFirst (templates in order, the first passing one
wins) and ModeAll (every template reported, the best passing one wins,
first on ties), ROIs canonicalised and clipped to the frame, templates that
cannot fit degrading to "not matched".

What differs is where the frame lives: Run uploads the captured frame once
(pinned staging), converts it to gray once on the GPU when any gray group
exists, and runs every group's Fleet on a device view of its ROI (pointer +
row stride), so a frame costs one PCIe transfer whatever the number of
groups.  Each group's results equal FindAll on the host view bit for bit.
"""
from __future__ import annotations

import ctypes as C
import threading
from dataclasses import dataclass, field
from typing import List, Optional, Tuple

import numpy as np

from . import CvmatchError, Fleet, Matcher, _check, _L, lib, stats_cap, view

ModeFirst = 0  # templates tried in order, first one passing wins
ModeAll = 1  # every template evaluated, best passing one wins

Rect = Tuple[int, int, int, int]  # (x0, y0, x1, y1), half-open like image.Rectangle


@dataclass
class Template:
    """Template engine.go:44-47: an id and its RGBA pixels ((h, w, 4) uint8)."""
    ID: int
    Img: Optional[np.ndarray]


@dataclass
class Rule:
    """Rule engine.go:52-67.  ROI in window coordinates; (0, 0, 0, 0) is the
    whole window."""
    ID: int
    Name: str = ""
    ROI: Rect = (0, 0, 0, 0)
    Gray: bool = False
    Threshold: float = 0.0
    Mode: int = ModeFirst
    Templates: List[Template] = field(default_factory=list)


@dataclass(frozen=True)
class Hit:
    """Hit engine.go:77-81: a template's best position, window coordinates."""
    TemplateID: int
    Score: float
    Rect: Rect


@dataclass
class RuleResult:
    """RuleResult engine.go:87-93."""
    RuleID: int
    Name: str
    Matched: bool = False
    Best: Optional[Hit] = None
    All: List[Hit] = field(default_factory=list)


class EngineError(ValueError):
    pass


def resolve_roi(r: Rect) -> Rect:
    """resolveROI engine.go:231-241: the zero rectangle is the whole window;
    anything else is canonicalised and must have area."""
    if tuple(r) == (0, 0, 0, 0):
        return (0, 0, 0, 0)
    x0, y0, x1, y1 = r
    c = (min(x0, x1), min(y0, y1), max(x0, x1), max(y0, y1))
    if c[0] >= c[2] or c[1] >= c[3]:
        raise EngineError(f"degenerate ROI {tuple(r)} has no area")
    return c


def _intersect(a: Rect, b: Rect) -> Rect:
    r = (max(a[0], b[0]), max(a[1], b[1]), min(a[2], b[2]), min(a[3], b[3]))
    return r if r[0] < r[2] and r[1] < r[3] else (0, 0, 0, 0)


class _Member:
    def __init__(self, m: Matcher, w: int, h: int):
        self.m, self.w, self.h = m, w, h
        self.refs: List[Tuple[int, int]] = []


class _Group:
    """One (ROI, gray) class sharing a Fleet (group engine.go:110-122)."""

    def __init__(self, roi: Rect, gray: bool):
        self.roi, self.gray = roi, gray
        self.members: List[_Member] = []
        self.fleet: Optional[Fleet] = None
        self.all_idx: List[int] = []
        self.sub_wh = (0, 0)
        self.sub: Optional[Fleet] = None
        self.sub_idx: List[int] = []

    def fit(self, w: int, h: int):
        """fit engine.go:351-379: the full fleet when every member fits the
        w x h view, else a sub-fleet of those that do, memoised on the last
        size (None when none fits)."""
        fits = [i for i, m in enumerate(self.members) if m.w <= w and m.h <= h]
        if len(fits) == len(self.members):
            return self.fleet, self.all_idx
        if self.sub_wh != (w, h):
            self.sub_wh, self.sub, self.sub_idx = (w, h), None, []
            if fits:
                self.sub = Fleet(*[self.members[i].m for i in fits])
                self.sub_idx = fits
        return self.sub, self.sub_idx


class Engine:
    """NewEngine engine.go:136-224 / Run engine.go:276-345."""

    def __init__(self, rules: List[Rule]):
        # private copies: a caller mutating its rule list cannot desync us
        rs = [Rule(r.ID, r.Name, tuple(r.ROI), r.Gray, r.Threshold, r.Mode, list(r.Templates)) for r in rules]
        # validation in the reference's order, before any GPU work
        sizes = {}
        for r in rs:
            if r.Mode not in (ModeFirst, ModeAll):
                raise EngineError(f"windowrules: rule {r.ID}: unknown mode {r.Mode}")
            try:
                r.ROI = resolve_roi(r.ROI)
            except EngineError as e:
                raise EngineError(f"windowrules: rule {r.ID}: {e}") from None
            for t in r.Templates:
                if t.Img is None:
                    raise EngineError(f"windowrules: rule {r.ID}: template {t.ID} has no image")
                h, w = t.Img.shape[:2]
                if w <= 0 or h <= 0:
                    raise EngineError(f"windowrules: rule {r.ID}: template {t.ID} is empty")
                key = (t.ID, r.Gray)
                if key not in sizes:
                    if w * h > stats_cap(1 if r.Gray else 4):
                        raise EngineError(f"windowrules: rule {r.ID}: template {t.ID}: cvmatch: template area "
                                          f"{w}x{h} exceeds the exact-statistics bound")
                    sizes[key] = (w, h)
                elif sizes[key] != (w, h):
                    fw, fh = sizes[key]
                    raise EngineError(f"windowrules: rule {r.ID}: template {t.ID} redefined at {w}x{h} "
                                      f"(first seen {fw}x{fh})")
        self._rules = rs
        matchers = {}
        groups = {}
        self._groups: List[_Group] = []
        for ri, r in enumerate(rs):
            for ti, t in enumerate(r.Templates):
                key = (t.ID, r.Gray)
                if key not in matchers:
                    h, w = t.Img.shape[:2]
                    try:
                        m = Matcher(np.ascontiguousarray(t.Img), gray=r.Gray)
                    except CvmatchError as e:
                        raise EngineError(f"windowrules: rule {r.ID}: template {t.ID}: {e}") from None
                    matchers[key] = _Member(m, w, h)
                rec = matchers[key]
                gk = (r.ROI, r.Gray)
                g = groups.get(gk)
                if g is None:
                    g = groups[gk] = _Group(r.ROI, r.Gray)
                    self._groups.append(g)
                mem = next((x for x in g.members if x.m is rec.m), None)
                if mem is None:
                    mem = _Member(rec.m, rec.w, rec.h)
                    g.members.append(mem)
                mem.refs.append((ri, ti))
        for g in self._groups:
            g.fleet = Fleet(*[m.m for m in g.members])
            g.all_idx = list(range(len(g.members)))
        self._dev = C.c_void_p()
        self._dev_cap = 0
        # Run is safe for concurrent use (engine.go:123): runs of one engine
        # share its device frame buffer and sub-fleet memos, so they take
        # turns (the GPU work of every call is ordered by the library anyway)
        self._lock = threading.Lock()

    def _device(self, nbytes: int) -> int:
        L = lib()
        if nbytes > self._dev_cap:
            if self._dev.value:
                _check(L.cvm_device_free(self._dev))
                self._dev = C.c_void_p()
                self._dev_cap = 0
            _check(L.cvm_device_alloc(nbytes, C.byref(self._dev)))
            self._dev_cap = nbytes
        return self._dev.value

    def run(self, frame) -> List[RuleResult]:
        """Run engine.go:276-345 on an (h, w, 4) RGBA frame (window origin at
        (0, 0)): verdicts in rule order."""
        with self._lock:
            return self._run(frame)

    def _run(self, frame) -> List[RuleResult]:
        p, stride, fw, fh, step = view(frame)
        if step != 4:
            raise EngineError("windowrules: frames are RGBA (h, w, 4) arrays")
        L = _L()
        span = (fh - 1) * stride + fw * 4
        gray_off = (span + 255) // 256 * 256
        any_gray = any(g.gray for g in self._groups)
        base = self._device(gray_off + (fw * fh if any_gray else 0))
        _check(L.cvm_copy_h2d(C.c_void_p(base), C.c_void_p(p.ctypes.data), span))
        if any_gray:
            _check(L.cvm_rgba_to_gray_device(C.c_void_p(base), stride, fw, fh, C.c_void_p(base + gray_off), None))
        self._frame = (base, gray_off, stride, fw, fh)
        out = [[None] * len(r.Templates) for r in self._rules]
        for g in self._groups:
            v = self._group_view(g)
            if v is None:
                continue  # ROI outside the frame: its templates stay unevaluated
            ptr, vstride, (x0, y0, x1, y1) = v
            fleet, idx = g.fit(x1 - x0, y1 - y0)
            if fleet is None:
                continue  # nothing fits this ROI at this frame size
            for i, res in enumerate(fleet.find_all_device(ptr, vstride, x1 - x0, y1 - y0)):
                mem = g.members[idx[i]]
                pos = (x0 + res.MaxX, y0 + res.MaxY)
                for ri, ti in mem.refs:
                    out[ri][ti] = (res.MaxV, pos, mem.w, mem.h)
        results = []
        for ri, r in enumerate(self._rules):
            rr = RuleResult(r.ID, r.Name)
            for ti, t in enumerate(r.Templates):
                o = out[ri][ti]
                if o is None:
                    continue
                score, (px, py), w, h = o
                hit = Hit(t.ID, score, (px, py, px + w, py + h))
                passed = score >= np.float32(r.Threshold)
                if r.Mode == ModeFirst:
                    if passed:
                        rr.Matched, rr.Best = True, hit
                        break
                    continue
                rr.All.append(hit)
                # strictly greater keeps the first occurrence on ties
                if passed and (rr.Best is None or hit.Score > rr.Best.Score):
                    rr.Matched, rr.Best = True, hit
            results.append(rr)
        return results

    Run = run

    def _group_view(self, g: _Group):
        """(device address, row stride, rect) of a group's region in the last
        uploaded frame (the gray copy for gray groups), or None when the ROI
        misses the frame."""
        base, gray_off, stride, fw, fh = self._frame
        rect = (0, 0, fw, fh)
        if g.roi != (0, 0, 0, 0):
            rect = _intersect(g.roi, rect)
            if rect == (0, 0, 0, 0):
                return None
        x0, y0 = rect[:2]
        if g.gray:
            return base + gray_off + y0 * fw + x0, fw, rect
        return base + y0 * stride + x0 * 4, stride, rect

    def __del__(self):
        d = getattr(self, "_dev", None)
        if d is not None and d.value:
            try:
                lib().cvm_device_free(d)
            except Exception:
                pass


def NewEngine(rules: List[Rule]) -> Engine:
    return Engine(rules)
