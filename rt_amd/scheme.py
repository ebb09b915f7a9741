"""Scheme description and packing for the MI355X rt engine.

A scheme is a sequence of surfaces in the form the reference's XML loader
builds them (common-lib/scheme.cpp:406-664 -> Surface::SetParams /
SetGeometry / SetDistance / SetAngles, surface.cpp:85-202) plus a beam
(scheme.cpp:681-856 -> Beam::SetParams / Create, beam.cpp:67-485).  This
module turns that description into the packed table of include/rt_hip.h:
RT_SURF_STRIDE doubles per surface and a spline pool.

Units follow the reference: lengths in mm, wavelengths and grating constants
in um, surface angles in degrees (converted as surface.cpp:201 and
beam.cpp:605-611 do), cone angles in degrees, Gaussian FWHM in arcsec.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple, Union

import numpy as np

# ---- constants of include/rt_hip.h ---------------------------------------
RT_PLANE, RT_CONIC, RT_TORIC, RT_AUXCLASS = 0, 1, 2, 3
RT_MIRROR, RT_LENS, RT_GRATING, RT_STOP, RT_SHADE, RT_DETECTOR, RT_AUXTYPE = range(7)
RT_CIRCLE, RT_RECTANGLE = 0, 1
RT_REFLECTIVE, RT_TRANSPARENT = 0, 1
RT_ORDER_FIXED, RT_ORDER_BLAZE = 0, 1
RT_FN_CONST, RT_FN_SPLINE = 0, 1
RT_BEAM_PARALLEL, RT_BEAM_CONIC, RT_BEAM_GAUSS = 0, 1, 2
RT_RANGE_RANDOM, RT_RANGE_UNIFORM, RT_RANGE_MONO = 0, 1, 2

RT_MAX_SURF = 32
RT_MAX_POLY = 10
RT_MAX_SPLINE_KNOTS = 1024

(S_CLASS, S_TYPE, S_SHAPE, S_C, S_K, S_NPOLY) = range(6)
S_A = 6
S_CR = 16
S_SIZE0, S_SIZE1, S_CEN0, S_CEN1 = 17, 18, 19, 20
S_DIST, S_ANG, S_SIN, S_COS = 21, 24, 27, 30
S_GRTYPE, S_ORDER, S_GRCONST = 33, 34, 35
S_ORDMODE, S_BLAZEK, S_ORDMIN, S_ORDMAX = 36, 37, 38, 39
S_N1KIND, S_N1VAL, S_N2KIND, S_N2VAL = 40, 41, 42, 43
S_CCD_XDIM, S_CCD_YDIM, S_CCD_XPIX, S_CCD_YPIX, S_CCD_XSIZE, S_CCD_YSIZE = 44, 45, 46, 47, 48, 49
S_QEKIND, S_QEVAL, S_SPKIND, S_SPVAL = 50, 51, 52, 53
SURF_STRIDE = 56

(B_TYPE, B_SHAPE, B_RANGE_DISTR, B_NPER, B_P1, B_P2, B_P3, B_ANGPAR) = range(8)
B_CEN, B_RANGE, B_NLAMBDA, B_COMPOSITE, B_ANGUNIT = 8, 11, 13, 14, 15
BEAM_STRIDE = 16

RT_TRACE_GENERATE, RT_TRACE_STORE, RT_TRACE_HIST, RT_TRACE_FLUX = 1, 2, 4, 8

ST_RAYS, ST_GOOD, ST_CCD_HITS, ST_CCD_OFFCHIP, ST_CCD_EDGE, ST_SPLINE_OOB, ST_ADJ_SKIP = range(7)
ST_SURF = 16
STATS_LEN = ST_SURF + 4 * RT_MAX_SURF

RT_PI = 3.14159265359          # common-lib/num_defs.h:20
DEG2RAD = RT_PI / 180.0        # num_defs.h:31


def surface_angle_rad(deg: float) -> float:
    """Angle handed to beam_transform for a surface angle given in degrees.

    Surface::SetAngles stores deg*deg2rad (surface.cpp:201), GetAngles
    divides by deg2rad (:227) and Beam::Transform multiplies again
    (beam.cpp:609-611); the round trip is kept bit for bit.
    """
    rad = deg * DEG2RAD
    return (rad / DEG2RAD) * DEG2RAD


# ---- splines ----------------------------------------------------------------

def natural_spline_y2(x: np.ndarray, y: np.ndarray) -> np.ndarray:
    """Natural cubic-spline second derivatives by the serial Numerical Recipes
    recurrence (the semantics of spline_init, rt_cuda_engine_spline.cu:42-145,
    computed serially)."""
    n = len(x)
    y2 = np.zeros(n)
    u = np.zeros(n)
    for i in range(1, n - 1):
        sig = (x[i] - x[i - 1]) / (x[i + 1] - x[i - 1])
        p = sig * y2[i - 1] + 2.0
        y2[i] = (sig - 1.0) / p
        ui = (y[i + 1] - y[i]) / (x[i + 1] - x[i]) - (y[i] - y[i - 1]) / (x[i] - x[i - 1])
        u[i] = (6.0 * ui / (x[i + 1] - x[i - 1]) - sig * u[i - 1]) / p
    for k in range(n - 2, 0, -1):
        y2[k] = y2[k] * y2[k + 1] + u[k]
    return y2


@dataclass
class Spline:
    """A tabulated function (TabulatedFunction, tabulated_function.h:18-54)."""
    x: np.ndarray
    y: np.ndarray
    name: str = ""

    def __post_init__(self):
        self.x = np.asarray(self.x, dtype=np.float64)
        self.y = np.asarray(self.y, dtype=np.float64)
        if len(self.x) < 4:  # tabulated_function.cpp:100
            raise ValueError("a tabulated function needs at least 4 points")
        if len(self.x) > RT_MAX_SPLINE_KNOTS:
            raise ValueError("too many knots")
        if np.any(np.diff(self.x) <= 0):
            raise ValueError("knots must be strictly increasing")
        self.y2 = natural_spline_y2(self.x, self.y)

    def __call__(self, xv: float) -> float:
        """TabulatedFunction::operator[] (tabulated_function.cpp:173-204)."""
        x, y, y2 = self.x, self.y, self.y2
        if xv < x[0] or xv > x[-1]:
            raise ValueError("argument out of range")
        il, ir = 0, len(x)
        while ir - il > 1:
            i = (il + ir) >> 1
            if x[i] == xv:
                return float(y[i])
            if x[i] > xv:
                ir = i
            else:
                il = i
        h = x[ir] - x[il]
        a = (x[ir] - xv) / h
        b = (xv - x[il]) / h
        return float(a * y[il] + b * y[ir] + ((a * a * a - a) * y2[il] + (b * b * b - b) * y2[ir]) * (h * h) / 6.0)


FnRef = Union[float, Spline]


def sellmeier(lam_um: np.ndarray, B: Sequence[float], C: Sequence[float]) -> np.ndarray:
    """3-term Sellmeier index n(lambda), C in um^2."""
    l2 = np.asarray(lam_um, dtype=np.float64) ** 2
    s = 1.0 + sum(b * l2 / (l2 - c) for b, c in zip(B, C))
    return np.sqrt(s)


# published Sellmeier coefficients (B1..B3, C1..C3 in um^2)
GLASSES = {
    # SCHOTT N-BK7 data sheet
    "N-BK7": ((1.03961212, 0.231792344, 1.01046945), (0.00600069867, 0.0200179144, 103.560653)),
    # fused silica, Malitson (1965)
    "F_SILICA": ((0.6961663, 0.4079426, 0.8974794), (0.0684043 ** 2, 0.1162414 ** 2, 9.896161 ** 2)),
    # calcium fluoride, Malitson (1963)
    "CAF2": ((0.5675888, 0.4710914, 3.8484723), (0.050263605 ** 2, 0.1003909 ** 2, 34.649040 ** 2)),
    # SCHOTT N-F2 data sheet
    "N-F2": ((1.39757037, 0.159201403, 1.2686543), (0.00995906143, 0.0546931752, 119.248346)),
}


def glass_spline(name: str, lo: float = 0.35, hi: float = 0.95, knots: int = 64) -> Spline:
    """n(lambda) of a catalogue glass tabulated at `knots` points."""
    B, C = GLASSES[name]
    x = np.linspace(lo, hi, knots)
    return Spline(x, sellmeier(x, B, C), name=name)


def qe_spline(lo: float = 0.30, hi: float = 1.00, knots: int = 200) -> Spline:
    """A smooth synthetic back-illuminated CCD quantum-efficiency curve."""
    x = np.linspace(lo, hi, knots)
    y = 0.92 * np.exp(-((x - 0.62) / 0.30) ** 4) + 0.03
    return Spline(x, y, name="QE")


# ---- surfaces and beams -----------------------------------------------------

@dataclass
class Surface:
    cls: int
    type: int
    shape: int = RT_CIRCLE
    params: Sequence[float] = (0.0, 0.0, 0)      # [c, k, N, a1..aN, (cR)]
    size: Tuple[float, float] = (0.0, 0.0)
    center: Tuple[float, float] = (0.0, 0.0)
    distance: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    angles: Tuple[float, float, float] = (0.0, 0.0, 0.0)    # degrees: about Y, X, Z
    grating_type: int = RT_REFLECTIVE
    order: int = 1
    gr_const: float = 1.0
    order_mode: int = RT_ORDER_FIXED
    blaze_k: float = 0.0
    order_range: Tuple[int, int] = (-1000000, 1000000)
    n1: FnRef = 1.0
    n2: FnRef = 1.0
    ccd: Optional[dict] = None      # xdim, ydim, xpix, ypix (mm), xsize, ysize (mm)
    qe: FnRef = 1.0
    spectrum: FnRef = 1.0
    comment: str = ""

    def validate(self):
        p = list(self.params)
        if self.cls == RT_CONIC:               # surface.cpp:113-129
            if len(p) < 3 or len(p) < 3 + int(p[2]):
                raise ValueError("invalid conic params")
        if self.cls == RT_TORIC:               # surface.cpp:96-112
            if len(p) < 4 or len(p) < 4 + int(p[2]):
                raise ValueError("invalid toric params")
        if len(p) >= 3 and int(p[2]) > RT_MAX_POLY:
            raise ValueError("too many aspheric coefficients")


@dataclass
class Beam:
    type: int = RT_BEAM_PARALLEL
    shape: int = RT_CIRCLE
    range_distr: int = RT_RANGE_UNIFORM
    params: Sequence[float] = (1000, 0.0, 10.0)   # N, inner, outer, [ratio], [cone deg | FWHM arcsec]
    center: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    range: Tuple[float, float] = (0.5, 0.6)
    nlambda: int = 1
    composite: bool = False

    @property
    def n_rays(self) -> int:
        n = int(self.params[0])
        return n * self.nlambda if self.composite else n

    def pack(self) -> np.ndarray:
        b = np.zeros(BEAM_STRIDE)
        p = list(self.params)
        b[B_TYPE], b[B_SHAPE], b[B_RANGE_DISTR] = self.type, self.shape, self.range_distr
        b[B_NPER], b[B_P1], b[B_P2] = p[0], p[1], p[2]
        if self.shape == RT_RECTANGLE:
            b[B_P3] = p[3]
            if self.type != RT_BEAM_PARALLEL:
                b[B_ANGPAR] = p[4]
        elif self.type != RT_BEAM_PARALLEL:
            b[B_ANGPAR] = p[3]
        b[B_CEN:B_CEN + 3] = self.center
        b[B_RANGE:B_RANGE + 2] = self.range
        b[B_NLAMBDA] = self.nlambda
        b[B_COMPOSITE] = 1.0 if self.composite else 0.0
        return b


@dataclass
class Scheme:
    surfaces: List[Surface]
    beam: Beam
    name: str = ""
    table: np.ndarray = field(init=False, repr=False)
    pool: np.ndarray = field(init=False, repr=False)

    def __post_init__(self):
        self.table, self.pool = pack_surfaces(self.surfaces)

    @property
    def n_surf(self) -> int:
        return len(self.surfaces)

    def detector(self) -> Optional[Surface]:
        for s in self.surfaces:
            if s.type == RT_DETECTOR and s.ccd:
                return s
        return None


def pack_surfaces(surfaces: Sequence[Surface]) -> Tuple[np.ndarray, np.ndarray]:
    """Packed table and spline pool of include/rt_hip.h."""
    if len(surfaces) > RT_MAX_SURF:
        raise ValueError("too many surfaces")
    table = np.zeros((len(surfaces), SURF_STRIDE))
    pool: List[float] = []
    offsets = {}

    def ref(fn: FnRef) -> Tuple[float, float]:
        if isinstance(fn, Spline):
            key = id(fn)
            if key not in offsets:
                offsets[key] = len(pool)
                k = len(fn.x)
                pool.append(float(k))
                pool.extend(fn.x.tolist())
                pool.extend(fn.y.tolist())
                pool.extend(fn.y2.tolist())
            return float(RT_FN_SPLINE), float(offsets[key])
        return float(RT_FN_CONST), float(fn)

    for i, s in enumerate(surfaces):
        s.validate()
        r = table[i]
        r[S_CLASS], r[S_TYPE], r[S_SHAPE] = s.cls, s.type, s.shape
        p = list(s.params) + [0.0] * 3
        if s.cls in (RT_CONIC, RT_TORIC):
            r[S_C], r[S_K] = p[0], p[1]
            npoly = int(p[2])
            r[S_NPOLY] = npoly
            r[S_A:S_A + npoly] = p[3:3 + npoly]
            if s.cls == RT_TORIC:
                r[S_CR] = list(s.params)[3 + npoly]
        r[S_SIZE0], r[S_SIZE1] = s.size
        r[S_CEN0], r[S_CEN1] = s.center
        r[S_DIST:S_DIST + 3] = s.distance
        for j in range(3):
            a = surface_angle_rad(s.angles[j])
            r[S_ANG + j] = a
            r[S_SIN + j] = math.sin(a)
            r[S_COS + j] = math.cos(a)
        r[S_GRTYPE], r[S_ORDER], r[S_GRCONST] = s.grating_type, s.order, s.gr_const
        r[S_ORDMODE], r[S_BLAZEK] = s.order_mode, s.blaze_k
        r[S_ORDMIN], r[S_ORDMAX] = s.order_range
        r[S_N1KIND], r[S_N1VAL] = ref(s.n1)
        r[S_N2KIND], r[S_N2VAL] = ref(s.n2)
        if s.ccd:
            c = s.ccd
            r[S_CCD_XDIM], r[S_CCD_YDIM] = c["xdim"], c["ydim"]
            r[S_CCD_XPIX], r[S_CCD_YPIX] = c["xpix"], c["ypix"]
            r[S_CCD_XSIZE] = c.get("xsize", c["xdim"] * c["xpix"])
            r[S_CCD_YSIZE] = c.get("ysize", c["ydim"] * c["ypix"])
            r[S_QEKIND], r[S_QEVAL] = ref(s.qe)
            r[S_SPKIND], r[S_SPVAL] = ref(s.spectrum)
    if not pool:
        pool = [0.0]
    return table.reshape(-1).copy(), np.asarray(pool, dtype=np.float64)


# ---- geometry helper: aim the next surface along a chief ray ----------------

def _rot_y(a, v):
    c, s = math.cos(a), math.sin(a)
    return np.array([v[0] * c - v[2] * s, v[1], v[0] * s + v[2] * c])


def _rot_x(a, v):
    c, s = math.cos(a), math.sin(a)
    return np.array([v[0], v[1] * c - v[2] * s, v[1] * s + v[2] * c])


def aim(direction, d: float, tilt_y: float = 0.0, tilt_x: float = 0.0):
    """(distance, angles_deg) that put the next surface vertex `d` mm along the
    unit `direction` (expressed in the current surface frame) with its axis
    along that direction, then tilted by tilt_y / tilt_x degrees."""
    v = np.asarray(direction, dtype=np.float64)
    v = v / np.linalg.norm(v)
    ay = math.atan2(v[0], v[2])
    rho = math.hypot(v[0], v[2])
    ax = math.atan2(v[1], rho)
    dist = tuple(float(t) for t in d * v)
    return dist, (math.degrees(ay) + tilt_y, math.degrees(ax) + tilt_x, 0.0)


def frame_direction(v, angles_deg):
    """Direction `v` of the previous frame expressed in a frame rotated by angles_deg."""
    w = np.asarray(v, dtype=np.float64)
    w = _rot_y(math.radians(angles_deg[0]), w)
    w = _rot_x(math.radians(angles_deg[1]), w)
    return w
