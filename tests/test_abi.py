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
