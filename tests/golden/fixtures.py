"""Committed golden fixtures and the two helpers that read them.

golden.json holds data only:
  plan_pins    the reference's 15 planner pins (plan_test.go:12-29)
  golden_110   the reference's RNG-free canary (impl_golden_test.go:70-86) and
               its recorded hash (golden_hashes_test.go:21)
  cases        the reference's golden shapes (impl_golden_test.go:60-69) on
               inputs from numpy's PCG64 (seed = case id); per case the
               FNV-1a-64 map hash and six-tuple, plus FP64 scores at sampled
               windows from the float64 brute force of cvmatch_test.go:13-81
make_golden.py rewrites the file; see its docstring for what each value is
computed with.
"""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PATH = os.path.join(HERE, "golden.json")

FNV_OFFSET = 14695981039346656037
FNV_PRIME = 1099511628211
MASK64 = (1 << 64) - 1


def load() -> dict:
    with open(PATH) as f:
        return json.load(f)


def case_inputs(case: dict):
    """The seeded inputs of one golden shape: (img, istride, tpl, tstride)."""
    iw, ih, tw, th = case["iw"], case["ih"], case["tw"], case["th"]
    step = case["step"]
    istride, tstride = (iw + case["ipad"]) * step, (tw + case["tpad"]) * step
    rng = np.random.default_rng(case["id"])
    img = rng.integers(0, 256, istride * ih, dtype=np.uint8)
    tpl = rng.integers(0, 256, tstride * th, dtype=np.uint8)
    return img, istride, tpl, tstride


def canary_110(spec: dict):
    iw, ih, tw, th = spec["iw"], spec["ih"], spec["tw"], spec["th"]
    img = np.full(iw * ih, spec["fill"], np.uint8)
    tpl = np.full(tw * th, spec["fill"], np.uint8)
    img[spec["img_at"]] = spec["img_val"]
    tpl[spec["tpl_at"]] = spec["tpl_val"]
    return img, iw, ih, tpl, tw, th


def fnv_hash(res: np.ndarray, tup) -> int:
    """goldenHash (impl_golden_test.go:27-44): FNV-1a 64 over the little-endian
    bytes of every map float's bits, then of min value, min x, min y, max
    value, max x, max y as 32-bit words."""
    words = np.ascontiguousarray(res, dtype=np.float32).ravel().view(np.uint32)
    tail = np.array([tup[0], tup[3]], np.float32).view(np.uint32)
    extra = np.array([tail[0], tup[1], tup[2], tail[1], tup[4], tup[5]], np.int64).astype(np.uint32)
    data = np.concatenate([words, extra]).astype("<u4").tobytes()
    h = FNV_OFFSET
    for b in data:
        h = ((h ^ b) * FNV_PRIME) & MASK64
    return h


def tuple_of(entry) -> tuple:
    """Six-tuple from its JSON form (floats as hex strings)."""
    return (float.fromhex(entry[0]), entry[1], entry[2], float.fromhex(entry[3]), entry[4], entry[5])


def tuple_to_json(t) -> list:
    return [float(np.float32(t[0])).hex(), int(t[1]), int(t[2]), float(np.float32(t[3])).hex(), int(t[4]), int(t[5])]
