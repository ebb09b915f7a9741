"""The C-ABI library (rt_amd/librt_hip.so) without a GPU: it loads, exports
every entry point include/rt_hip.h declares, agrees with the header's
layout constants, validates schemes on the host and generates/compiles the
per-scheme kernels (hiprtc runs without a device)."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from rt_amd import lib, schemes
from rt_amd import scheme as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "rt_hip.h")


def declared_functions(path):
    text = open(path).read()
    return sorted(set(re.findall(r"^int (rt_hip_\w+)\(", text, re.M)))


def test_library_exports_every_declared_entry_point():
    names = declared_functions(HEADER)
    assert "rt_hip_trace" in names and len(names) >= 8
    out = subprocess.run(["nm", "-D", "--defined-only", lib.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (\w+)", out))
    missing = [n for n in names if n not in exported]
    assert not missing, f"declared but not exported: {missing}"
    L = lib.lib()
    for n in names:
        assert hasattr(L, n)


def _enum_values(text):
    vals = {}
    for name, val in re.findall(r"#define (RT_[A-Z0-9_]+) (\d+)", text):
        vals[name] = int(val)
    for name, val in re.findall(r"\b(RT_[A-Z0-9_]+)\s*=\s*([A-Z0-9_+* ]+?)\s*(?:,|}|/\*|\n)", text):
        vals[name] = int(eval(val, {}, dict(vals)))
    return vals


def test_python_layout_matches_header():
    v = _enum_values(open(HEADER).read())
    pairs = {"RT_SURF_STRIDE": S.SURF_STRIDE, "RT_BEAM_STRIDE": S.BEAM_STRIDE, "RT_STATS_LEN": S.STATS_LEN,
             "RT_S_C": S.S_C, "RT_S_A": S.S_A, "RT_S_CR": S.S_CR, "RT_S_DIST": S.S_DIST, "RT_S_ANG": S.S_ANG,
             "RT_S_SIN": S.S_SIN, "RT_S_COS": S.S_COS, "RT_S_GRCONST": S.S_GRCONST, "RT_S_N2VAL": S.S_N2VAL,
             "RT_S_CCD_XDIM": S.S_CCD_XDIM, "RT_S_SPVAL": S.S_SPVAL, "RT_B_ANGPAR": S.B_ANGPAR,
             "RT_B_COMPOSITE": S.B_COMPOSITE, "RT_ST_SURF": S.ST_SURF, "RT_MAX_SURF": S.RT_MAX_SURF,
             "RT_TRACE_FLUX": S.RT_TRACE_FLUX, "RT_DETECTOR": S.RT_DETECTOR, "RT_TRANSPARENT": S.RT_TRANSPARENT}
    for k, py in pairs.items():
        assert v[k] == py, k


def test_packed_angles_and_trig():
    sch = schemes.echelle(1, 4)
    t = sch.table.reshape(sch.n_surf, S.SURF_STRIDE)
    for s, surf in enumerate(sch.surfaces):
        for j in range(3):
            rad = surf.angles[j] * S.DEG2RAD
            rad = (rad / S.DEG2RAD) * S.DEG2RAD      # surface.cpp:201,227 + beam.cpp:609-611
            assert t[s, S.S_ANG + j] == rad
            assert t[s, S.S_SIN + j] == np.sin(rad) and t[s, S.S_COS + j] == np.cos(rad)


def test_validate_scheme():
    sch = schemes.echelle(1, 4)
    assert lib.validate(sch) == 0
    bad = schemes.echelle(1, 4)
    bad.table = bad.table.copy()
    bad.table[S.S_NPOLY] = 11
    assert lib.validate(bad) == 4 and "aspheric" in lib.last_error()
    bad.table[S.S_NPOLY] = 0
    bad.table[S.S_CLASS] = 7
    assert lib.validate(bad) == 4
    lens = 5 * S.SURF_STRIDE
    bad = schemes.echelle(1, 4)
    bad.table = bad.table.copy()
    bad.table[lens + S.S_N2VAL] = len(bad.pool) + 3          # spline reference past the pool
    assert lib.validate(bad) == 4 and "refractive index" in lib.last_error()
    L = lib.lib()
    assert L.rt_hip_validate_scheme(None, 0, None, 0) == 4


@pytest.mark.parametrize("make,mode", [(lambda: schemes.echelle(1, 4), S.RT_BEAM_GAUSS),
                                       (lambda: schemes.czerny_turner(1, 5), -1)])
def test_generated_kernel_compiles(make, mode):
    rc, size, text = lib.jit_check(make(), mode)
    assert rc == 0, text[-2000:]
    assert size > 10000
    assert text.count("rt::surface_step(") == make().n_surf
