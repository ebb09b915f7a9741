"""Algorithmic FP64 operation counts of the fused trace.

Convention: every +, -, *, / and sqrt written in a stage formula counts 1
(an FMA is its multiply and its add, 2); negation, fabs, comparisons,
integer work and wave-uniform subexpressions (per-surface constants such as
(k+1)c^2, half sizes, pixel area) count 0; sin/cos/log/atan2 count 1.
The counts are read off the stage formulas of rt_amd/csrc/rt_device.h
(which restate the reference kernels) and multiplied by the kernel's own
counters: rays entering each surface's intersection stage, Newton
iterations per surface, rays lost at intersection and at mask/action.
"""
from __future__ import annotations

import numpy as np

from . import scheme as S

NORMALIZE = 9           # |c|^2 (3 mul, 2 add), one reciprocal square root, 3 mul
PLANE_INT = 5           # s0 = -Z/cZ; X += cX s0; Y += cY s0
CONIC_QUAD = 39         # A, B, C, D, sqrt, stable q, root, update
SPLINE = 21             # splint with a binary search (no FP in the search)
CCD_MAP = 8             # two pixel coordinates + the edge test
FLUX = 25               # QE spline + weight
BEAM_POS_CIRCLE = 12
BEAM_DIR = {S.RT_BEAM_PARALLEL: 0, S.RT_BEAM_CONIC: 5, S.RT_BEAM_GAUSS: 36}
BEAM_LAMBDA = 2


def transform_cost(rec) -> int:
    c = 3 if (rec[S.S_DIST] != 0 or rec[S.S_DIST + 1] != 0 or rec[S.S_DIST + 2] != 0) else 0
    return c + 12 * int(sum(1 for j in range(3) if rec[S.S_ANG + j] != 0.0))


def newton_iter_cost(rec) -> int:
    """one Newton iteration: the point (6), sag with sqrt and divide, the
    aspheric polynomial and its derivative by one Horner pass (4N - 3), dF/ds,
    the update and the convergence test"""
    n = int(rec[S.S_NPOLY])
    return (27 + 4 * n) if int(rec[S.S_CLASS]) == S.RT_CONIC else (39 + 4 * n)


def is_newton(rec) -> bool:
    cls, n, c = int(rec[S.S_CLASS]), int(rec[S.S_NPOLY]), rec[S.S_C]
    return cls == S.RT_TORIC or (cls == S.RT_CONIC and n > 0)


def intersect_cost(rec) -> int:
    cls, n, c = int(rec[S.S_CLASS]), int(rec[S.S_NPOLY]), rec[S.S_C]
    if cls == S.RT_PLANE or (cls == S.RT_CONIC and n == 0 and c == 0.0):
        return PLANE_INT
    if cls == S.RT_CONIC and n == 0:
        return CONIC_QUAD
    return 1            # s0; the iterations are counted separately


def derivs_cost(rec) -> int:
    n = int(rec[S.S_NPOLY])
    return (9 + 2 * n) if int(rec[S.S_CLASS]) == S.RT_CONIC else (20 + 4 * n)


def fn_cost(kind, table: bool = False) -> int:
    """a spline-valued function: SPLINE per evaluation, 0 when the kernel
    loads it from the per-wavelength table"""
    return SPLINE if kind == S.RT_FN_SPLINE and not table else 0


def action_cost(rec, n1_reused: bool = False, table: bool = False) -> int:
    """reflection / refraction / diffraction with the surface normal; n1(lambda)
    is not counted when the previous action already evaluated the same table,
    neither index when the per-wavelength table supplies it"""
    t, cls, n = int(rec[S.S_TYPE]), int(rec[S.S_CLASS]), int(rec[S.S_NPOLY])
    if t == S.RT_MIRROR:
        if cls == S.RT_PLANE:
            return 0
        return (26 + 2 * n) if cls == S.RT_CONIC else (37 + 4 * n)
    idx = (0 if n1_reused else fn_cost(rec[S.S_N1KIND], table)) + fn_cost(rec[S.S_N2KIND], table) + 1
    if t == S.RT_LENS:
        if cls == S.RT_PLANE:
            return idx + 11
        return idx + ((36 + 2 * n) if cls == S.RT_CONIC else (47 + 4 * n))
    if t == S.RT_GRATING:
        refl = int(rec[S.S_GRTYPE]) == S.RT_REFLECTIVE
        order = 2 if int(rec[S.S_ORDMODE]) == S.RT_ORDER_BLAZE else 0
        base = idx + order + 3
        if cls == S.RT_PLANE:
            return base + (10 if refl else 19)
        if cls == S.RT_CONIC:
            return base + (55 + 2 * n if refl else 63 + 2 * n)
        return base + (68 + 4 * n if refl else 76 + 4 * n)
    return 0


def n1_reuse(table) -> list:
    """per surface: does its action take n1 from the previous action's n2 (the
    same table at the same wavelength)?  The rule of the kernel generator."""
    out, prev = [], None
    for rec in table:
        t = int(rec[S.S_TYPE])
        opt = t in (S.RT_LENS, S.RT_GRATING)
        n1s = rec[S.S_N1KIND] == S.RT_FN_SPLINE
        out.append(bool(opt and n1s and prev is not None and rec[S.S_N1VAL] == prev))
        if t != S.RT_AUXTYPE:
            prev = rec[S.S_N2VAL] if (opt and rec[S.S_N2KIND] == S.RT_FN_SPLINE) else None
    return out


def mask_cost(rec) -> int:
    return 5 if int(rec[S.S_SHAPE]) == S.RT_CIRCLE else 0


def trace_flops(scheme: "S.Scheme", stats: np.ndarray, generated: bool = True,
                generated_kernel: bool = True, table: bool = False) -> float:
    """Algorithmic FP64 operations of the trace(s) that produced `stats`
    (`generated`: the beam was generated in the kernel; `generated_kernel`:
    the per-scheme kernel ran, which reuses repeated refractive indices;
    `table`: that kernel read the spline-valued functions from the
    per-wavelength table, so their evaluations count 0 per ray)."""
    t = np.asarray(scheme.table).reshape(scheme.n_surf, S.SURF_STRIDE)
    st = np.asarray(stats).astype(np.float64)
    total = 0.0
    if generated:
        b = scheme.beam
        total += st[S.ST_RAYS] * (BEAM_POS_CIRCLE + BEAM_DIR[b.type] + BEAM_LAMBDA)
    reuse = n1_reuse(t) if generated_kernel else [False] * scheme.n_surf
    for s in range(scheme.n_surf):
        rec = t[s]
        if int(rec[S.S_TYPE]) == S.RT_AUXTYPE:
            continue
        n_in = st[S.ST_SURF + 4 * s]
        iters = st[S.ST_SURF + 4 * s + 1]
        after_int = n_in - st[S.ST_SURF + 4 * s + 2]
        after_act = after_int - st[S.ST_SURF + 4 * s + 3]
        total += n_in * (transform_cost(rec) + NORMALIZE + intersect_cost(rec))
        if is_newton(rec):
            total += iters * newton_iter_cost(rec)
        total += after_int * mask_cost(rec)
        total += after_act * (action_cost(rec, reuse[s], table) + NORMALIZE)
        if int(rec[S.S_TYPE]) == S.RT_DETECTOR and rec[S.S_CCD_XDIM] > 0:
            flux = FLUX - (fn_cost(rec[S.S_QEKIND], False) if table else 0)
            total += after_act * CCD_MAP + st[S.ST_CCD_HITS] * flux
    return total


def ray_surface(scheme: "S.Scheme", stats: np.ndarray) -> int:
    st = np.asarray(stats)
    return int(sum(int(st[S.ST_SURF + 4 * s]) for s in range(scheme.n_surf)))
