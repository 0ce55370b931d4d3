"""Writes tests/golden/golden.json.

  python tests/golden/make_golden.py

What each value is computed with:
  plan_pins, golden_110  transcribed from the reference's tests
                         (plan_test.go:12-29; impl_golden_test.go:70-86,
                         golden_hashes_test.go:21): data, nothing computed
  cases[*].hash/tuple    the CPU oracle's map at 1 and at 4 threads (both must
                         agree), hashed as goldenHash does
                         (impl_golden_test.go:27-44).  These are recorded
                         values, as the reference records its own goldens: a
                         later change that moves any bit of any map fails them
  cases[*].fp64          scores at sampled windows from refMatch's float64
                         steps (cvmatch_test.go:13-81) restated below in
                         numpy, independent of the oracle

The inputs are seeded PCG64 draws, so the reference's hashes #101-#109 (made
from Go's math/rand) do not apply to them; #110 needs no generator and is the
reference's own value.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)

import fixtures as F  # noqa: E402

PLAN_PINS = [  # plan_test.go:12-29: tw, th, rw, rh -> w, h, dw, dh
    (128, 128, 1793, 953, 1024, 512, 897, 385),
    (96, 32, 1505, 969, 512, 1024, 417, 969),
    (32, 32, 1889, 1049, 512, 256, 481, 225),
    (256, 256, 3585, 1905, 2048, 1024, 1793, 769),
    (24, 24, 1577, 977, 256, 512, 233, 489),
    (80, 80, 433, 401, 512, 512, 433, 401),
    (64, 64, 449, 449, 512, 512, 449, 449),
    (100, 100, 769, 501, 512, 256, 413, 157),
    (128, 128, 625, 473, 512, 512, 385, 385),
    (64, 64, 1, 1, 64, 64, 1, 1),
    (1, 1, 50, 40, 4, 2, 4, 2),
    (33, 257, 8, 44, 64, 512, 8, 44),
    (17, 13, 81, 49, 64, 64, 48, 49),
    (1, 400, 2000, 1, 32, 512, 32, 1),
    (400, 1, 1, 2000, 512, 2, 1, 2),
]

GOLDEN_110 = {  # impl_golden_test.go:75-86, golden_hashes_test.go:21
    "iw": 200, "ih": 150, "tw": 32, "th": 32, "fill": 128,
    "img_at": 75 * 200 + 100, "img_val": 129, "tpl_at": 5 * 32 + 7, "tpl_val": 129,
    "cn": 1, "step": 1, "hash": "0x99534488138e0635",
}

SHAPES = [  # goldenCases impl_golden_test.go:61-69: id, iw, ih, tw, th, cn, step, ipad, tpad
    (101, 97, 61, 17, 13, 1, 1, 0, 0),
    (102, 97, 61, 17, 13, 4, 4, 0, 0),
    (103, 120, 90, 24, 18, 3, 4, 0, 0),
    (104, 300, 200, 31, 29, 1, 1, 7, 3),
    (105, 64, 64, 64, 64, 1, 1, 0, 0),
    (106, 50, 40, 1, 1, 4, 4, 0, 0),
    (107, 640, 400, 96, 32, 3, 4, 0, 0),
    (108, 128, 96, 5, 4, 3, 3, 0, 0),
    (109, 40, 300, 33, 257, 1, 1, 0, 0),
]

POINTS = 64


def ref_points(img, istride, iw, ih, tpl, tstride, tw, th, cn, step, xs, ys):
    """refMatch (cvmatch_test.go:13-81) in float64 at windows (xs[i], ys[i])."""
    T = np.lib.stride_tricks.as_strided(tpl, (th, tw, step), (tstride, step, 1))[..., :cn].astype(np.float64)
    area = float(tw * th)
    inv_area = 1.0 / area
    mean = T.sum(axis=(0, 1)) * inv_area
    templ_norm = float(np.sum((T * T).sum(axis=(0, 1)) * inv_area - mean * mean))
    if templ_norm < 2.220446049250313e-16:
        return np.ones(len(xs))
    templ_norm = np.sqrt(templ_norm * area)
    out = np.empty(len(xs))
    for i, (x, y) in enumerate(zip(xs, ys)):
        W = np.lib.stride_tricks.as_strided(img[y * istride + x * step:], (th, tw, step),
                                            (istride, step, 1))[..., :cn].astype(np.float64)
        corr = float(np.sum(W * T))
        wsum = W.sum(axis=(0, 1))
        wsum2 = float(np.sum(W * W))
        num = corr - float(np.sum(wsum * mean))
        wmean2 = float(np.sum(wsum * wsum)) * inv_area
        diff2 = max(wsum2 - wmean2, 0.0)
        den = np.sqrt(diff2) * templ_norm if diff2 > min(0.5, 10 * 1.1920929e-07 * wsum2) else 0.0
        if abs(num) < den:
            num /= den
        elif abs(num) < den * 1.125:
            num = float(np.copysign(1.0, num))
        else:
            num = 0.0
        out[i] = num
    return out


def main():
    from oracle import oracle as O

    O.lib()
    cases = []
    for cid, iw, ih, tw, th, cn, step, ipad, tpad in SHAPES:
        case = {"id": cid, "iw": iw, "ih": ih, "tw": tw, "th": th, "cn": cn, "step": step, "ipad": ipad, "tpad": tpad}
        img, istride, tpl, tstride = F.case_inputs(case)
        m1, t1 = O.match_u8(img, istride, iw, ih, tpl, tstride, tw, th, cn, step, threads=1)
        m4, t4 = O.match_u8(img, istride, iw, ih, tpl, tstride, tw, th, cn, step, threads=4)
        assert t1 == t4 and np.array_equal(m1.view(np.uint32), m4.view(np.uint32)), cid
        h = F.fnv_hash(m1, t1)
        assert h == O.golden_hash(m1, t1), cid
        rw, rh = iw - tw + 1, ih - th + 1
        rng = np.random.default_rng(7000 + cid)
        xs = [int(v) for v in rng.integers(0, rw, POINTS)] + [t1[1], t1[4]]
        ys = [int(v) for v in rng.integers(0, rh, POINTS)] + [t1[2], t1[5]]
        fp = ref_points(img, istride, iw, ih, tpl, tstride, tw, th, cn, step, xs, ys)
        case.update({"hash": f"{h:#018x}", "tuple": F.tuple_to_json(t1),
                     "fp64": {"xs": xs, "ys": ys, "scores": [float(v).hex() for v in fp]}})
        cases.append(case)
        print(f"case {cid}: {rw}x{rh} map, hash {h:#018x}")
    rows = lambda items: ",\n  ".join(json.dumps(v) for v in items)  # noqa: E731
    with open(F.PATH, "w") as f:
        f.write('{\n "plan_pins": [\n  ' + rows(PLAN_PINS) + "\n ],\n")
        f.write(' "golden_110": ' + json.dumps(GOLDEN_110) + ",\n")
        f.write(' "cases": [\n  ' + rows(cases) + "\n ]\n}\n")
    assert F.load()["cases"] == cases
    print("wrote", F.PATH)


if __name__ == "__main__":
    main()
