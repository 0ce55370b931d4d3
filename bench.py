"""Benchmark: 4K RGBA MatchMap (TM_CCOEFF_NORMED + minMaxLoc) on one MI355X.

Workload: parent 3840x2160 RGBA (i.i.d. bytes through a 5x5 box blur, seeded),
template 128x128 cropped from it with sigma-2 noise on RGB; the full 3713x2033
float32 response map and the six-tuple are produced every step.  Inputs are
resident in HBM before the timed region; each step is one complete call
(template statistics, template spectrum, correlation, normalisation, extrema,
24-byte result back to the host).

Prints ONE JSON line.  value = Gwindows/s (windows = result elements).
roofline: the call's algorithmic HBM bytes (parent + template + map, 8.40 B
per window) / the average duration of its dominant kernel, from HIP events
recorded on the library's streams (in a profiling pass after the timed one);
kernel_stream_efficiency: that kernel's own compulsory bytes (the spectra it
must stream) / the same duration.  vs_fp64 fields: |score - FP64 score| at
seeded windows and at the reported extrema.
cpu_baseline: the CPU oracle (a C restatement of the reference's op
sequence) timed on this host on the same workload.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import math
import os
import threading
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def tiles(tw, th, rw, rh, plan):
    (dftW, dftH, blockW, blockH), _ = plan
    ntx = (rw + blockW - 1) // blockW
    nty = (rh + blockH - 1) // blockH
    lastY0 = ((rh - 1) // blockH) * blockH
    dh2 = 0
    if lastY0 > 0:
        n2 = 1
        while n2 < rh - lastY0 + th - 1:
            n2 <<= 1
        n2 = max(2, n2)
        if n2 < dftH:
            dh2 = n2
    out = []
    for t in range(ntx * nty):
        tx, ty = t % ntx, t // ntx
        x0, y0 = tx * blockW, ty * blockH
        bw = rw - x0 if tx == ntx - 1 else blockW
        bh = rh - y0 if ty == nty - 1 else blockH
        n = dh2 if (dh2 and y0 == lastY0) else dftH
        out.append((x0, y0, bw, bh, n))
    return out, dftW


def stage_bytes(iw, ih, tw, th, cn, plan):
    """Algorithmic (compulsory) HBM bytes of one call of each stage:
    row_fwd: each tile's pixels (with halo) once + its spectrum rows written;
    cols: spectrum rows in and out + the template spectrum once per channel
    (and its shrunk-band copy); row_inv: spectrum rows in + the map written;
    normalize: the parent once + the correlation read + the map written."""
    rw, rh = iw - tw + 1, ih - th + 1
    ts, dftW = tiles(tw, th, rw, rh, plan)
    hw = dftW // 2 + 1
    b = {"row_fwd": 0, "cols": 0, "row_inv": 0}
    heights = set()
    for (x0, y0, bw, bh, n) in ts:
        loadW, loadH = bw + tw - 1, bh + th - 1
        loaded = min(n, (loadH + 1) & ~1)
        need = min(n, (bh + 1) & ~1)
        b["row_fwd"] += loadW * loadH * 4 + cn * loaded * hw * 8
        b["cols"] += cn * (loaded + need) * hw * 8
        b["row_inv"] += cn * need * hw * 8 + bw * bh * 4
        heights.add(n)
    b["cols"] += sum(cn * n * hw * 8 for n in heights)
    b["normalize"] = iw * ih * 4 + 2 * rw * rh * 4
    return b


def fp64_scores(img, tpl, cn, pts):
    """The FP64 reference the deviations are measured against: TM_CCOEFF_NORMED
    from exact integer window statistics (Python integers), the numerator,
    denominator and normOne's guard ladder (impl.go:816-849) evaluated in
    binary64.  img (H, W) or (H, W, C) uint8, tpl likewise; cn channels;
    pts = [(x, y)] window origins."""
    th, tw = tpl.shape[:2]
    area = tw * th
    T = tpl.reshape(th, tw, -1)[..., :cn].astype(np.int64)
    tsum = T.sum(axis=(0, 1))
    var_sum = int((area * (T * T).sum(axis=(0, 1)) - tsum * tsum).sum())
    out = np.empty(len(pts), np.float64)
    for i, (x, y) in enumerate(pts):
        if var_sum == 0:
            out[i] = 1.0
            continue
        P = img[y:y + th, x:x + tw].reshape(th, tw, -1)[..., :cn].astype(np.int64)
        s = P.sum(axis=(0, 1))
        num = area * int((P * T).sum()) - int((s * tsum).sum())
        idiff = area * int((P * P).sum()) - int((s * s).sum())
        s2 = int((P * P).sum())
        lim = min(10.0 * 2.0 ** -23 * float(s2), 0.5)
        if not float(idiff) / float(area) > lim:
            out[i] = 0.0
            continue
        den = math.sqrt(float(idiff)) * math.sqrt(float(var_sum))
        an = abs(float(num))
        out[i] = float(num) / den if an < den else ((1.0 if num > 0 else -1.0) if an < den * 1.125 else 0.0)
    return out


def delta_stats(got, ref):
    d = np.abs(np.asarray(got, np.float64) - np.asarray(ref, np.float64))
    return {"max": float(d.max()) if d.size else 0.0, "mean": float(d.mean()) if d.size else 0.0, "n": int(d.size)}


def map_delta_vs_fp64(m, img, tpl, cn, samples=4096, seed=0, extra=()):
    """max / mean |map - FP64| over seeded windows plus the given ones."""
    rh, rw = m.shape
    rng = np.random.default_rng(seed)
    pts = list(zip(rng.integers(0, rw, samples).tolist(), rng.integers(0, rh, samples).tolist())) + list(extra)
    ref = fp64_scores(img, tpl, cn, pts)
    return delta_stats([m[y, x] for (x, y) in pts], ref)


def cpu_baseline(parent, sub, budget_s=3.0):
    from oracle import oracle as O

    threads = max(1, min(16, os.cpu_count() or 1))
    O.match_rgba(parent, sub, threads=threads, want_map=True)  # warm-up
    times = []
    t_end = time.perf_counter() + budget_s
    while time.perf_counter() < t_end or len(times) < 2:
        t0 = time.perf_counter()
        O.match_rgba(parent, sub, threads=threads, want_map=True)
        times.append(time.perf_counter() - t0)
        if len(times) >= 20:
            break
    ih, iw = parent.shape[:2]
    th, tw = sub.shape[:2]
    windows = (iw - tw + 1) * (ih - th + 1)
    best = min(times)
    # the same call on one thread, best of 3
    one = []
    for _ in range(3):
        t0 = time.perf_counter()
        O.match_rgba(parent, sub, threads=1, want_map=True)
        one.append(time.perf_counter() - t0)
    host = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            host = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), host)
    except OSError:
        pass
    return {
        "value": windows / best / 1e9,
        "unit": "Gwindows/s",
        "cores": threads,
        "kind": "port",
        "sample": f"full 3840x2160 RGBA 128x128 MatchMap, best of {len(times)} runs after 1 warm-up "
                  f"({best * 1e3:.1f} ms/run) with {threads} threads",
        "single_thread": {"value": windows / min(one) / 1e9, "ms": round(min(one) * 1e3, 1), "runs": 3},
        "host_cpu": host,
    }


def fleet_leg(frames_timed=63, solo_frames=8, fp64_frames=(0, 21, 42, 63)):
    """The batched Fleet over the seeded frame stream: 256 templates U[16,128]^2
    (half with sigma-2 noise) on 64 frames of 2560x1440 RGBA, six results per
    member out.  templates*frames/s of the direct (mixed-precision: exact
    int8-MFMA correlation, FP64 scores) Fleet and of the plain Fleet (the
    reference's bucketed float32 FFT form, bit-exact to it), with the frames
    resident in HBM (the headline form) and as host frames; the plain GPU form
    is 256 solo Finds per frame on the same device-resident frames (host-frame
    solo Finds, which upload the frame once per member, are reported apart).
    Deviations: both forms' reported extrema against FP64 scores on four
    frames, and direct against plain over every timed frame."""
    import cvmatch_amd as cvm
    from cvmatch_amd import synth

    L = cvm.lib()
    frames, tpls = synth.workload_fleet()
    frames = [np.ascontiguousarray(f) for f in frames]
    ms = [cvm.NewMatcher(t) for t in tpls]
    fl = cvm.NewFleet(*ms)
    n = len(ms)
    fh, fw = frames[0].shape[:2]
    dfr = []
    for f in frames:
        p = C.c_void_p()
        assert L.cvm_device_alloc(f.nbytes, C.byref(p)) == 0
        assert L.cvm_copy_h2d(p, f.ctypes.data_as(C.c_void_p), f.nbytes) == 0
        dfr.append(p)
    nt = min(frames_timed, len(frames) - 1)
    out = {"workload": "64-frame 2560x1440 RGBA flat-rectangle stream (alpha 255), 256 templates U[16,128]^2",
           "frames_timed": nt}
    res = {}
    for key, direct, mixed in (("direct", True, False), ("plain_fleet", False, False), ("mixed", False, True)):
        for where in ("device", "host"):
            if mixed and where == "host":
                continue
            if where == "device":
                def call(f):
                    return fl.find_all_device(dfr[f].value, fw * 4, fw, fh, direct=direct, mixed=mixed)
            else:
                def call(f):
                    return fl.find_all(frames[f], direct=direct)
            call(0)  # geometry / fragment caches
            got = {0: call(0)}
            t0 = time.perf_counter()
            for f in range(1, 1 + nt):
                got[f] = call(f)
            dt = (time.perf_counter() - t0) / nt
            out[f"{key}_{where}"] = {"tpl_frames_per_s": round(n / dt, 1), "ms_per_frame": round(dt * 1e3, 3)}
            if where == "device":
                res[key] = got
    rep = cvm.tune_report()
    out["direct_form"] = {k.split("@")[0]: {"choice": v["choice"], "ms_per_candidate": v["ms"], "source": v["source"]}
                          for k, v in rep.items()
                          if (k.startswith("fleet.fused/") or k.startswith("fleet.fused.merge/")) and f"{fw}x{fh}" in k}
    # the direct form's unfused launches (each member's C' map written to HBM
    # and read back by the scoring launch), forced on a Fleet of its own
    os.environ["CVMATCH_FLEET_FUSED"] = "0"
    try:
        flu = cvm.NewFleet(*ms)
        nu = min(nt, 16)
        flu.find_all_device(dfr[0].value, fw * 4, fw, fh, direct=True)
        flu.find_all_device(dfr[0].value, fw * 4, fw, fh, direct=True)
        same = True
        t0 = time.perf_counter()
        for f in range(1, 1 + nu):
            same &= flu.find_all_device(dfr[f].value, fw * 4, fw, fh, direct=True) == res["direct"][f]
        dt = (time.perf_counter() - t0) / nu
        out["direct_unfused_device"] = {"tpl_frames_per_s": round(n / dt, 1), "ms_per_frame": round(dt * 1e3, 3),
                                        "frames_timed": nu, "equals_fused": bool(same)}
        del flu
    finally:
        del os.environ["CVMATCH_FLEET_FUSED"]
    for m in ms:  # each member's first Find tunes and caches its shape's choices: untimed
        m.find_device(dfr[0].value, fw * 4, fw, fh)
    t0 = time.perf_counter()
    for f in range(1, 1 + solo_frames):
        for m in ms:
            m.find_device(dfr[f].value, fw * 4, fw, fh)
    dt = (time.perf_counter() - t0) / solo_frames
    out["plain_solo_device"] = {"tpl_frames_per_s": round(n / dt, 1), "ms_per_frame": round(dt * 1e3, 3),
                                "frames_timed": solo_frames}
    t0 = time.perf_counter()
    for f in range(1, 3):
        for m in ms:
            m.find(frames[f])
    dt = (time.perf_counter() - t0) / 2
    out["plain_solo_host_frames"] = {"tpl_frames_per_s": round(n / dt, 1), "ms_per_frame": round(dt * 1e3, 3),
                                     "frames_timed": 2, "note": "each Find uploads the 14.7 MB frame again"}
    routed, gate, xo = fl.mixed_plan()
    cand = [g for g in gate if g >= 0]
    out["mixed_plan"] = {"crossover_side": xo, "routed_to_fft": int(sum(routed)), "candidates": len(cand),
                         "gate_bound": 1e-5,
                         "candidate_fft_dev_min": min(cand) if cand else None,
                         "candidate_fft_dev_median": float(np.median(cand)) if cand else None,
                         "mixed_equals_direct": res["mixed"] == res["direct"]}
    solo = out["plain_solo_device"]["tpl_frames_per_s"]
    out["direct_over_plain_solo"] = round(out["direct_device"]["tpl_frames_per_s"] / solo, 3)
    out["plain_fleet_over_plain_solo"] = round(out["plain_fleet_device"]["tpl_frames_per_s"] / solo, 3)
    out["mixed_over_plain_solo"] = round(out["mixed_device"]["tpl_frames_per_s"] / solo, 3)
    # deviations
    live = [i for i, t in enumerate(tpls) if t[..., :3].std() > 0 or t[..., 3].std() > 0]
    dvp = [abs(res["direct"][f][i][k] - res["plain_fleet"][f][i][k])
           for f in res["direct"] for i in live for k in (0, 3)]
    same_loc = np.mean([res["direct"][f][i][4:6] == res["plain_fleet"][f][i][4:6]
                        for f in res["direct"] for i in live])
    out["direct_vs_plain"] = {"max_abs_delta_extrema": float(max(dvp)), "mean_abs_delta_extrema": float(np.mean(dvp)),
                              "max_location_agreement": round(float(same_loc), 4), "frames": len(res["direct"])}
    for key in ("direct", "plain_fleet"):
        got, ref = [], []
        for f in fp64_frames:
            if f not in res[key]:
                continue
            for i in live:
                r = res[key][f][i]
                cn = 3  # alpha 255 in every frame and template: the alpha rule picks 3
                v = fp64_scores(frames[f], tpls[i], cn, [(r[1], r[2]), (r[4], r[5])])
                got += [r[0], r[3]]
                ref += list(v)
        out[f"{key}_vs_fp64_at_extrema"] = dict(delta_stats(got, ref), frames=list(fp64_frames))
    for p in dfr:
        L.cvm_device_free(p)
    del fl, ms
    return out


def host_leg(parent, sub, reps=10):
    """The drop-in path: host buffers in, six-tuple (and map) out, PCIe
    included.  MatchMap writes a freshly allocated result per call, as the
    reference's does (cvmatch.go:100-117); the reused-result and pinned-input
    rows isolate the page faults and the staging copy.  Never the headline."""
    import cvmatch_amd as cvm

    rw, rh = parent.shape[1] - sub.shape[1] + 1, parent.shape[0] - sub.shape[0] + 1
    pin = cvm.host_array(parent.shape)
    pin[:] = parent
    res = np.empty((rh, rw), dtype=np.float32)
    L = cvm._L()
    p, ps, pw, ph, _ = cvm.view(parent)
    s, ss, sw, sh, _ = cvm.view(sub)
    out = cvm.cvm_result()

    def map_reused():
        cvm._check(L.cvm_match_map(cvm._ptr(p), ps, pw, ph, cvm._ptr(s), ss, sw, sh, cvm._ptr(res), C.byref(out)))

    legs = {}
    for key, fn in (("match_map", lambda: cvm.match_map(parent, sub)), ("match_map_reused_result", map_reused),
                    ("match", lambda: cvm.match(parent, sub)), ("match_pinned_input", lambda: cvm.match(pin, sub))):
        fn()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        dt = (time.perf_counter() - t0) / reps
        legs[key] = {"ms": round(dt * 1e3, 3), "Gwindows_per_s": round(rw * rh / dt / 1e9, 3)}
    legs["copy_threads"] = cvm.threads()
    return legs


def peaks_leg(L, d_map, rw, rh, tw, th, reps=20):
    """Every occurrence over the device-resident 4K map: threshold 0.5 and the
    template's half-size window (the recipe of cvmatch_test.go:346-365), peak
    list and count back on the host."""
    cap = 4096
    d_xy, d_v = C.c_void_p(), C.c_void_p()
    assert L.cvm_device_alloc(8 * cap, C.byref(d_xy)) == 0 and L.cvm_device_alloc(4 * cap, C.byref(d_v)) == 0
    cnt = C.c_int32()

    def call():
        if L.cvm_find_peaks_device(d_map, rw, rh, 0.5, tw // 2, th // 2, d_xy, d_v, cap, C.byref(cnt), None) != 0:
            raise RuntimeError(L.cvm_last_error_detail().decode())

    call()
    t0 = time.perf_counter()
    for _ in range(reps):
        call()
    dt = (time.perf_counter() - t0) / reps
    L.cvm_device_free(d_xy)
    L.cvm_device_free(d_v)
    return {"ms": round(dt * 1e3, 4), "thresh": 0.5, "rx": tw // 2, "ry": th // 2, "count": cnt.value,
            "map_bytes": rw * rh * 4}


def windowrules_leg(reps=20):
    """The window-rules engine over the reference example's demo window
    (1280x800 RGBA, 4 rules in 3 (ROI, gray) groups, 7 templates): one Run =
    one frame upload, one gray conversion, three device-view FindAlls and the
    verdicts.  Beside it, the same groups run as the reference's Run does,
    each FindAll on its own host view (one upload per group, gray views
    converted separately)."""
    import cvmatch_amd as cvm
    from cvmatch_amd import synth
    from cvmatch_amd import windowrules as W

    base = synth.window_frame()
    R = synth.WINDOW_RECTS

    def tpl(i, r):
        return W.Template(i, synth.crop(base, r))

    panel = tpl(105, R["panel"])
    rules = [
        W.Rule(1, "confirm-button", Mode=W.ModeFirst, Threshold=0.95,
               Templates=[W.Template(100, synth.decoy_widget(120, 36, 99)), tpl(101, R["ok"])]),
        W.Rule(2, "tray-status", ROI=(1150, 0, 1280, 48), Gray=True, Mode=W.ModeAll, Threshold=0.95,
               Templates=[tpl(102, R["wifi"]), tpl(103, R["mute"]), W.Template(104, synth.decoy_widget(24, 24, 77))]),
        W.Rule(3, "main-panel", ROI=(180, 180, 520, 420), Threshold=0.95, Templates=[panel]),
        W.Rule(4, "oversized-roi", ROI=(0, 0, 40, 40), Threshold=0.95, Templates=[panel]),
    ]
    eng = W.NewEngine(rules)
    frames = [base, synth.window_frame(move_button=True)]

    def host_groups(frame):
        for g in eng._groups:
            x0, y0, x1, y1 = (0, 0, frame.shape[1], frame.shape[0]) if g.roi == (0, 0, 0, 0) else g.roi
            fleet, _ = g.fit(x1 - x0, y1 - y0)
            if fleet is not None:
                fleet.find_all(frame[y0:y1, x0:x1])

    out = {"workload": "windowrules demo window 1280x800 RGBA, 4 rules, 3 groups, 7 templates"}
    for key, fn in (("engine_run", eng.run), ("per_group_host_findall", host_groups)):
        fn(frames[0])
        fn(frames[1])
        t0 = time.perf_counter()
        for i in range(reps):
            fn(frames[i % 2])
        out[key + "_ms"] = round((time.perf_counter() - t0) / reps * 1e3, 3)
    return out


def configs_leg(reps=15, small_only=False):
    """The other configurations, device resident (inputs in HBM, the map
    written on the device, graphs replayed like the headline's step):
    640x480 RGBA Match 48x48 (cn = 4), the 1920x1080 gray UI MatchGray 96x32,
    the 2560x1440 gray direct-vs-FFT sweep (ms per call with the map for
    square templates 8..256; the direct form is the FP64-exact one) and the
    4K RGBA 128x128 call in both forms.  small_only: the two small calls'
    times alone (profiling)."""
    import cvmatch_amd as cvm
    from cvmatch_amd import synth

    L = cvm.lib()
    out = cvm.cvm_result()
    bufs = []

    def dev(a):
        p = C.c_void_p()
        assert L.cvm_device_alloc(max(a.nbytes, 4), C.byref(p)) == 0
        assert L.cvm_copy_h2d(p, a.ctypes.data_as(C.c_void_p), a.nbytes) == 0
        bufs.append(p)
        return p

    def timed(fn):
        for _ in range(3):  # eager, capture, replay
            fn()
        L.cvm_synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        L.cvm_synchronize()
        return (time.perf_counter() - t0) / reps

    def call(direct, d_img, istride, iw, ih, d_tpl, tstride, tw, th, cn, step, d_map):
        f = L.cvm_match_u8_direct_device if direct else L.cvm_match_u8_device
        rc = f(d_img, istride, iw, ih, d_tpl, tstride, tw, th, cn, step, d_map, C.byref(out), None)
        if rc != 0:
            raise RuntimeError(L.cvm_last_error_detail().decode())

    def host_map(dm, rw, rh):
        m = np.empty((rh, rw), np.float32)
        assert L.cvm_copy_d2h(m.ctypes.data_as(C.c_void_p), dm, m.nbytes) == 0
        return m

    def extrema_pts():
        return [(out.min_x, out.min_y), (out.max_x, out.max_y)]

    res = {}
    parent, sub, _ = synth.workload_match_640(seed=1, alpha="random")
    rw, rh = 640 - 48 + 1, 480 - 48 + 1
    dp, ds = dev(parent), dev(sub)
    dm = dev(np.zeros(rw * rh, np.float32))
    dt = timed(lambda: call(False, dp, 640 * 4, 640, 480, ds, 48 * 4, 48, 48, 0, 4, dm))
    res["match_640x480_rgba_48"] = {"ms": round(dt * 1e3, 4), "Gwindows_per_s": round(rw * rh / dt / 1e9, 3)}
    if not small_only:
        res["match_640x480_rgba_48"]["vs_fp64"] = map_delta_vs_fp64(host_map(dm, rw, rh), parent, sub, 4, seed=11,
                                                                    extra=extrema_pts())
    gray, gsub = synth.workload_gray_ui_1080p(seed=2)[:2]
    rw, rh = 1920 - 96 + 1, 1080 - 32 + 1
    dg, dgs = dev(gray), dev(gsub)
    dm = dev(np.zeros(rw * rh, np.float32))
    dt = timed(lambda: call(False, dg, 1920, 1920, 1080, dgs, 96, 96, 32, 1, 1, dm))
    res["match_gray_1920x1080_96x32"] = {"ms": round(dt * 1e3, 4), "Gwindows_per_s": round(rw * rh / dt / 1e9, 3)}
    if small_only:
        for p in bufs:
            L.cvm_device_free(p)
        return res
    res["match_gray_1920x1080_96x32"]["vs_fp64"] = map_delta_vs_fp64(host_map(dm, rw, rh), gray, gsub, 1, seed=12,
                                                                     extra=extrema_pts())
    img, _ = synth.workload_direct_1440p(seed=4, sizes=())
    di = dev(img)
    dm = dev(np.zeros(2560 * 1440, np.float32))
    sweep = []
    for sz in (8, 16, 24, 32, 48, 64, 96, 128, 192, 256):
        t = np.ascontiguousarray(img[100:100 + sz, 200:200 + sz])
        dtp = dev(t)
        fft = timed(lambda: call(False, di, 2560, 2560, 1440, dtp, sz, sz, sz, 1, 1, dm))
        dirt = timed(lambda: call(True, di, 2560, 2560, 1440, dtp, sz, sz, sz, 1, 1, dm))
        row = {"template": sz, "fft_ms": round(fft * 1e3, 4), "direct_ms": round(dirt * 1e3, 4)}
        if sz in (16, 24, 32, 48):  # the mixed-precision configuration: the direct map against FP64
            rw, rh = 2560 - sz + 1, 1440 - sz + 1
            row["direct_vs_fp64"] = map_delta_vs_fp64(host_map(dm, rw, rh), img, t, 1, samples=2048, seed=sz,
                                                      extra=extrema_pts())
        sweep.append(row)
    res["direct_vs_fft_gray_2560x1440"] = sweep
    cross = [r["template"] for r in sweep if r["direct_ms"] > r["fft_ms"]]
    res["crossover_template"] = cross[0] if cross else None
    # the headline's 4K RGBA shape (128x128, cn = 4) in both forms
    parent, sub, _ = synth.workload_map_4k(seed=3, alpha="random")
    rw, rh = 3840 - 128 + 1, 2160 - 128 + 1
    dp, ds = dev(parent), dev(sub)
    dm = dev(np.zeros(rw * rh, np.float32))
    fft = timed(lambda: call(False, dp, 3840 * 4, 3840, 2160, ds, 128 * 4, 128, 128, 0, 4, dm))
    dirt = timed(lambda: call(True, dp, 3840 * 4, 3840, 2160, ds, 128 * 4, 128, 128, 0, 4, dm))
    res["direct_vs_fft_rgba_3840x2160_128"] = {"fft_ms": round(fft * 1e3, 4), "direct_ms": round(dirt * 1e3, 4)}
    for p in bufs:
        L.cvm_device_free(p)
    return res


def concurrency_leg(calls=480, threads=8):
    """640x480 RGBA Match 48x48 (device resident, the six-tuple back on the
    host) from one host thread, then from `threads` threads at once: calls/s.
    Every call is synchronous; concurrent calls take separate contexts of the
    library's pool (own streams, scratch and captured graph), so they overlap
    on the device instead of queueing behind one lock."""
    import cvmatch_amd as cvm
    from cvmatch_amd import synth

    L = cvm.lib()
    parent, sub, _ = synth.workload_match_640(seed=1, alpha="random")
    rw, rh = 640 - 48 + 1, 480 - 48 + 1
    bufs = []

    def dev(nbytes, src=None):
        p = C.c_void_p()
        assert L.cvm_device_alloc(nbytes, C.byref(p)) == 0
        if src is not None:
            assert L.cvm_copy_h2d(p, src.ctypes.data_as(C.c_void_p), nbytes) == 0
        bufs.append(p)
        return p

    dp, ds = dev(parent.nbytes, parent), dev(sub.nbytes, sub)
    maps = [dev(rw * rh * 4) for _ in range(threads)]
    outs = [cvm.cvm_result() for _ in range(threads)]

    def call(i):
        rc = L.cvm_match_u8_device(dp, 640 * 4, 640, 480, ds, 48 * 4, 48, 48, 0, 4, maps[i], C.byref(outs[i]), None)
        if rc != 0:
            raise RuntimeError(L.cvm_last_error_detail().decode())

    for _ in range(3):
        call(0)
    t0 = time.perf_counter()
    for _ in range(calls):
        call(0)
    single = calls / (time.perf_counter() - t0)
    want = outs[0].tuple()
    gate = threading.Barrier(threads + 1)
    errors = []

    def worker(i):
        try:
            for _ in range(3):  # this thread's context: eager, capture, replay
                call(i)
            gate.wait()
            for _ in range(calls // threads):
                call(i)
        except Exception as e:  # reported in the leg's output
            errors.append(repr(e))
            gate.abort()

    ts = [threading.Thread(target=worker, args=(i,)) for i in range(threads)]
    for t in ts:
        t.start()
    try:
        gate.wait()
    except threading.BrokenBarrierError:  # a worker failed: its error is reported below
        pass
    t0 = time.perf_counter()
    for t in ts:
        t.join()
    multi = (calls // threads) * threads / (time.perf_counter() - t0)
    same = all(o.tuple() == want for o in outs)
    for p in bufs:
        L.cvm_device_free(p)
    return {"workload": "640x480 RGBA Match 48x48, device resident, synchronous calls",
            "single_thread_calls_per_s": round(single, 1), "threads": threads,
            "multi_thread_calls_per_s": round(multi, 1), "speedup": round(multi / single, 3),
            "results_identical": same, "errors": errors}


def frames_leg(parent, sub, n=20):
    """A stream of 4K RGBA host frames through a prepared Matcher (128x128):
    synchronous Find per frame against the frame stream (cvm_frames_*), where
    the upload of frame i+1 overlaps the matching of frame i."""
    import cvmatch_amd as cvm

    m = cvm.NewMatcher(sub)
    shifted = np.ascontiguousarray(np.roll(parent, 3, axis=1))
    fr = [parent, shifted]
    out = {"workload": "4K RGBA host frames (two alternating), 128x128 Matcher", "frames": n}
    m.find(fr[0])
    m.find(fr[1])
    t0 = time.perf_counter()
    for i in range(n):
        m.find(fr[i % 2])
    out["sync_find_ms_per_frame"] = round((time.perf_counter() - t0) / n * 1e3, 3)
    for depth in (2, 3):
        q = cvm.FrameStream(depth)
        q.find(m, fr[0])
        q.result()
        t0 = time.perf_counter()
        for i in range(n):
            q.find(m, fr[i % 2], tag=i)
            if q.pending() >= depth:
                q.result()
        while q.pending():
            q.result()
        out[f"stream_depth{depth}_ms_per_frame"] = round((time.perf_counter() - t0) / n * 1e3, 3)
    return out


def copy_calibration(mib):
    import cvmatch_amd as cvm

    L = cvm.lib()
    n = mib << 20
    a, b = C.c_void_p(), C.c_void_p()
    assert L.cvm_device_alloc(n, C.byref(a)) == 0 and L.cvm_device_alloc(n, C.byref(b)) == 0
    L.cvm_profile_enable(1)
    for _ in range(5):
        assert L.cvm_device_copy(b, a, n) == 0
    names = C.create_string_buffer(32 * 8)
    tot = (C.c_double * 8)()
    cnt = (C.c_int64 * 8)()
    k = L.cvm_profile_read(names, tot, cnt, 8)
    for i in range(k):
        nm = names.raw[32 * i:32 * (i + 1)].split(b"\0")[0].decode()
        if nm == "copy":
            ms = tot[i] / cnt[i]
            print(json.dumps({"copy_bytes": n, "avg_ms": ms, "GBps_read_plus_write": 2 * n / ms / 1e6}), flush=True)
    L.cvm_device_free(a)
    L.cvm_device_free(b)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-fleet", action="store_true", help="skip the Fleet leg")
    ap.add_argument("--no-host", action="store_true", help="skip the host-buffer (PCIe-inclusive) leg")
    ap.add_argument("--fleet-frames", type=int, default=63, help="frames timed per Fleet form")
    ap.add_argument("--fleet-only", action="store_true", help="only run the Fleet leg (profiling)")
    ap.add_argument("--small-only", action="store_true",
                    help="only time the 640x480 RGBA Match and the 1080p gray MatchGray (profiling)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run the plain copy kernel over MIB MiB (calibrates FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.fleet_only:
        print(json.dumps({"fleet": fleet_leg(args.fleet_frames)}), flush=True)
        return
    if args.small_only:
        print(json.dumps({"configs": configs_leg(reps=args.steps, small_only=True)}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the path shards per frame or per "
              f"fleet member)", file=sys.stderr)
        sys.exit(2)

    import cvmatch_amd as cvm
    from cvmatch_amd import synth

    L = cvm.lib()
    parent, sub, (cx, cy) = synth.workload_map_4k(seed=3, alpha="random")
    ih, iw = parent.shape[:2]
    th, tw = sub.shape[:2]
    rw, rh = iw - tw + 1, ih - th + 1
    windows = rw * rh

    def dalloc(n):
        p = C.c_void_p()
        if L.cvm_device_alloc(n, C.byref(p)) != 0:
            raise RuntimeError(L.cvm_last_error_detail().decode())
        return p

    d_img, d_tpl, d_map = dalloc(parent.nbytes), dalloc(sub.nbytes), dalloc(windows * 4)
    L.cvm_copy_h2d(d_img, parent.ctypes.data_as(C.c_void_p), parent.nbytes)
    L.cvm_copy_h2d(d_tpl, sub.ctypes.data_as(C.c_void_p), sub.nbytes)
    out = cvm.cvm_result()

    def step():
        rc = L.cvm_match_u8_device(d_img, iw * 4, iw, ih, d_tpl, tw * 4, tw, th, 0, 4, d_map, C.byref(out), None)
        if rc != 0:
            raise RuntimeError(L.cvm_last_error_detail().decode())

    for _ in range(args.warmup):
        step()
    L.cvm_synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    L.cvm_synchronize()
    elapsed = time.perf_counter() - t0

    # per-kernel times from HIP events on the library's streams, in a
    # separate pass so the event records stay out of the timed region
    L.cvm_profile_reset()
    L.cvm_profile_enable(1)
    for _ in range(max(5, min(args.steps, 20))):
        step()
    L.cvm_synchronize()
    L.cvm_profile_enable(0)

    names = C.create_string_buffer(32 * 64)
    tot = (C.c_double * 64)()
    cnt = (C.c_int64 * 64)()
    n = L.cvm_profile_read(names, tot, cnt, 64)
    stages = {}
    for i in range(n):
        nm = names.raw[32 * i:32 * (i + 1)].split(b"\0")[0].decode()
        stages[nm] = {"avg_ms": tot[i] / max(cnt[i], 1), "count": int(cnt[i])}

    # parity spot check of the timed result against the planted crop
    assert (out.max_x, out.max_y) == (cx, cy), (out.tuple(), (cx, cy))

    ms = elapsed / args.steps * 1e3
    value = windows * args.steps / elapsed / 1e9
    cn = 4 if not cvm.alpha_const(parent) else 3
    plan = cvm.plan(tw, th, rw, rh)
    sb = stage_bytes(iw, ih, tw, th, cn, plan)
    kern = {k: v for k, v in stages.items() if k in sb}
    dom = max(kern, key=lambda k: kern[k]["avg_ms"])
    dom_s = kern[dom]["avg_ms"] * 1e-3
    # algorithmic bytes of one call (8.40 B per window here): the parent and
    # the template read once, the map written once
    bytes_alg = iw * ih * 4 + tw * th * 4 + windows * 4
    achieved = bytes_alg / dom_s / 1e9
    traffic = None
    traffic_src = "profiles/pipeline_4k_rgba_matchmap.md (rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE; calibrated)"
    tf = os.path.join(ROOT, "profiles", "traffic_4k_rgba.json")
    if os.path.exists(tf):
        try:
            traffic = json.load(open(tf)).get(dom)
        except Exception:
            traffic = None
    # the headline map against FP64 at seeded windows and both extrema
    hmap = np.empty((rh, rw), np.float32)
    L.cvm_copy_d2h(hmap.ctypes.data_as(C.c_void_p), d_map, hmap.nbytes)
    vs64 = map_delta_vs_fp64(hmap, parent, sub, cn, samples=2048, seed=3,
                             extra=[(out.min_x, out.min_y), (out.max_x, out.max_y)])
    line = {
        "metric": "Gwindows/s + %HBM roofline, 4K RGBA Match; Fleet tpl*frames/s at |d|<=1e-5",
        "value": round(value, 4),
        "unit": "Gwindows/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8->f32",
        "data": "synthetic (seeded PCG64)",
        "config": {"workload": "MatchMap 3840x2160 RGBA (box-blurred i.i.d.) vs 128x128 crop + sigma2 noise, cn=%d" % cn,
                   "windows_per_step": windows, "plan": list(plan[0]), "inputs": "HBM-resident"},
        "hbm_frac_pipeline": round(bytes_alg / (ms * 1e-3) / (HBM_PEAK_GBS * 1e9), 5),
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
                     "traffic_source": traffic_src, "alg_bytes_per_launch": bytes_alg,
                     "alg_bytes_def": "parent + template + map of the call (8.40 B/window x 7,548,529 windows)",
                     "avg_launch_ms": round(kern[dom]["avg_ms"], 5)},
        "kernel_stream_efficiency": {"kernel": dom, "stage_bytes": sb[dom],
                                     "GBps": round(sb[dom] / dom_s / 1e9, 1),
                                     "frac": round(sb[dom] / dom_s / 1e9 / HBM_PEAK_GBS, 4),
                                     "def": "the kernel's own compulsory spectrum/pixel/map bytes per launch"},
        "vs_fp64": vs64,
        "stages_ms": {k: round(v["avg_ms"], 5) for k, v in stages.items()},
    }
    def say(msg):  # progress on stderr; the JSON line is the only stdout
        print(f"[bench] {msg}", file=sys.stderr, flush=True)

    say(f"headline {ms:.4f} ms/step")
    line["peaks"] = peaks_leg(L, d_map, rw, rh, tw, th)
    say("peaks leg done")
    if not args.no_host:
        line["configs"] = configs_leg()
        say("configs leg done")
    if not args.no_host:
        line["host_buffers"] = host_leg(parent, sub)
        say("host-buffer leg done")
        line["frame_stream"] = frames_leg(parent, sub)
        say("frame-stream leg done")
        line["windowrules"] = windowrules_leg()
        say("windowrules leg done")
    if not args.no_host:
        line["concurrency"] = concurrency_leg()
        say("concurrency leg done")
    if not args.no_fleet:
        line["fleet"] = fleet_leg(args.fleet_frames)
        say("fleet leg done")
    rep = cvm.tune_report()
    kinds = {}
    for k, v in rep.items():
        kind = k.split("/")[0]
        kinds.setdefault(kind, {}).setdefault(str(v["choice"]), 0)
        kinds[kind][str(v["choice"])] += 1
    line["tuner"] = {"decisions": len(rep), "choice_counts_by_kind": kinds,
                     # the fused-or-not decision is in fleet.direct_form (the fleet leg's
                     # forced unfused run overrides the key afterwards)
                     "fleet_2560x1440": {k.split("@")[0]: {"choice": v["choice"], "ms": v["ms"], "source": v["source"]}
                                         for k, v in rep.items()
                                         if k.startswith("fleet.") and "2560x1440" in k
                                         and not k.startswith("fleet.fused/")
                                         and not k.startswith("fleet.fused.merge/")},
                     "headline_4k": {k.split("@")[0]: v["choice"] for k, v in rep.items() if "3840x2160" in k},
                     "sources": sorted({v["source"] for v in rep.values()})}
    if not args.no_cpu:
        cb = cpu_baseline(parent, sub)
        line["cpu_baseline"] = cb
        line["gpu_over_cpu"] = round(value / cb["value"], 2)
    print(json.dumps(line), flush=True)
    for p in (d_img, d_tpl, d_map):
        L.cvm_device_free(p)


if __name__ == "__main__":
    main()
