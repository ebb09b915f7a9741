"""ctypes binding of the CPU oracle (oracle/librt_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and bench.py's
cpu_baseline leg.  The library is built by oracle/Makefile.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_dp = C.POINTER(C.c_double)
_u64p = C.POINTER(C.c_uint64)
_u32p = C.POINTER(C.c_uint32)
_i8p = C.POINTER(C.c_int8)


def lib():
    global _LIB
    if _LIB is None:
        _LIB = _load(os.environ.get("RT_ORACLE_LIB") or os.path.join(_HERE, "librt_oracle.so"))
    return _LIB


_FAST = None


def fast_lib():
    """The oracle built for speed (oracle/librt_oracle_fast.so: -O3, AVX2,
    strict IEEE): the timed CPU baseline of bench.py."""
    global _FAST
    if _FAST is None:
        _FAST = _load(os.path.join(_HERE, "librt_oracle_fast.so"))
    return _FAST


def _load(path):
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built (run `make -C oracle`)")
    L = C.CDLL(path)
    L.oracle_philox4x32_10.argtypes = [_u32p, _u32p, _u32p]
    L.oracle_u01.argtypes = [C.c_uint32, C.c_uint32]
    L.oracle_u01.restype = C.c_double
    L.oracle_spline_init.argtypes = [C.c_size_t, _dp, _dp, _dp]
    L.oracle_spline_eval.argtypes = [C.c_size_t, _dp, _dp, _dp, C.c_double, C.POINTER(C.c_int)]
    L.oracle_spline_eval.restype = C.c_double
    L.oracle_gen_ray.argtypes = [_dp, C.c_uint64, C.c_uint64] + [_dp] * 7
    L.oracle_trace.argtypes = [_dp, C.c_int, _dp, _dp, C.c_uint64, C.c_size_t, C.c_size_t,
                               _dp, _dp, _dp, _dp, _dp, _dp, _dp, _i8p, C.c_int,
                               _u32p, _dp, _u64p, _dp, _i8p, C.c_int]
    L.oracle_trace_iters.argtypes = L.oracle_trace.argtypes + [_u32p, C.c_int]
    L.oracle_beam_transform.argtypes = [C.c_size_t, _dp, _dp] + [_dp] * 6
    L.oracle_normalize.argtypes = [C.c_size_t, _dp, _dp, _dp]
    L.oracle_surface_intersection.argtypes = [_dp, C.c_size_t] + [_dp] * 6 + [_u64p, _i8p, _u64p]
    L.oracle_apply_mask.argtypes = [_dp, C.c_size_t, _dp, _dp, _u64p, _i8p]
    L.oracle_surface_reflection.argtypes = [_dp, C.c_size_t] + [_dp] * 6
    L.oracle_surface_refraction.argtypes = [_dp, C.c_size_t] + [_dp] * 7 + [_u64p, _i8p]
    L.oracle_surface_diffraction.argtypes = [_dp, C.c_size_t] + [_dp] * 8 + [_u64p, _i8p]
    L.oracle_ccd_coordinates.argtypes = [C.c_size_t, _dp, _dp, C.c_double, C.c_double,
                                         C.c_double, C.c_double, _u32p, _u32p]
    L.oracle_grating_energy_distr.argtypes = [C.c_size_t, _dp, C.c_long] + [C.c_double] * 6 + [_dp]
    _u8p = C.POINTER(C.c_uint8)
    L.oracle_trace_record.argtypes = ([_dp, C.c_int, _dp, _dp, C.c_uint64, C.c_size_t] + [_dp] * 7 +
                                      [_i8p, _u64p, _u8p, _dp, C.c_int])
    L.oracle_rearrange.argtypes = [_u8p, C.c_size_t, _u64p, _u64p]
    L.oracle_beam_order.argtypes = [C.c_size_t, _u64p, _u8p, _u8p] + [_dp] * 8
    L.oracle_beam_order.restype = None
    return L


def _p(a, t=_dp):
    if a is None:
        return None
    return a.ctypes.data_as(t)


def philox(ctr, key):
    c = np.asarray(ctr, dtype=np.uint32)
    k = np.asarray(key, dtype=np.uint32)
    o = np.zeros(4, dtype=np.uint32)
    lib().oracle_philox4x32_10(_p(c, _u32p), _p(k, _u32p), _p(o, _u32p))
    return o


def spline_init(x, y):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    y2 = np.zeros_like(x)
    rc = lib().oracle_spline_init(len(x), _p(x), _p(y), _p(y2))
    if rc != 0:
        raise RuntimeError(f"oracle_spline_init failed ({rc})")
    return y2


def spline_eval(x, y, y2, xv):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    y2 = np.ascontiguousarray(y2, dtype=np.float64)
    err = C.c_int(0)
    v = lib().oracle_spline_eval(len(x), _p(x), _p(y), _p(y2), float(xv), C.byref(err))
    return v, err.value


def gen_rays(beam_desc, seed, begin, n):
    b = np.ascontiguousarray(beam_desc, dtype=np.float64)
    out = np.zeros((7, n))
    vals = [C.c_double() for _ in range(7)]
    L = lib()
    for i in range(n):
        L.oracle_gen_ray(_p(b), seed, begin + i, *[C.byref(v) for v in vals])
        out[:, i] = [v.value for v in vals]
    return out


class TraceResult:
    def __init__(self, rays, flag, hist, flux, stats, dump, dump_flag, status):
        self.rays, self.flag, self.hist, self.flux = rays, flag, hist, flux
        self.stats, self.dump, self.dump_flag, self.status = stats, dump, dump_flag, status


def trace(scheme, seed=0, begin=0, n=None, rays=None, flag=None, flags=None, hist=False,
          flux=False, dump=False, nthreads=0, iters=None, iters_mode=0, fast=False):
    """Run the oracle's fused trace. `scheme` is an rt_amd.scheme.Scheme.
    With rays=None the beam is generated from (scheme.beam, seed).
    iters (n_surf x n uint32) with iters_mode 1 receives the Newton iteration
    count of every (surface, ray); with iters_mode 2 it forces them (finite
    differences at fixed iteration counts)."""
    from rt_amd import scheme as S
    table = np.ascontiguousarray(scheme.table, dtype=np.float64)
    pool = np.ascontiguousarray(scheme.pool, dtype=np.float64)
    beam = scheme.beam.pack()
    if n is None:
        n = scheme.beam.n_rays if rays is None else rays.shape[1]
    f = S.RT_TRACE_STORE
    if rays is None:
        f |= S.RT_TRACE_GENERATE
        R = np.zeros((7, n))
        flag = np.zeros(n, dtype=np.int8)
    else:
        R = np.ascontiguousarray(rays, dtype=np.float64).copy()
        flag = np.ones(n, dtype=np.int8) if flag is None else np.ascontiguousarray(flag, dtype=np.int8).copy()
    det = scheme.detector()
    H = F = None
    if det is not None and hist:
        f |= S.RT_TRACE_HIST
        H = np.zeros(int(det.ccd["ydim"]) * int(det.ccd["xdim"]), dtype=np.uint32)
    if det is not None and flux:
        f |= S.RT_TRACE_FLUX
        F = np.zeros(int(det.ccd["ydim"]) * int(det.ccd["xdim"]), dtype=np.float64)
    if flags is not None:
        f |= flags
    st = np.zeros(S.STATS_LEN, dtype=np.uint64)
    D = DF = None
    if dump:
        D = np.zeros((scheme.n_surf, 7, n))
        DF = np.zeros((scheme.n_surf, n), dtype=np.int8)
    rows = [R[i] for i in range(7)]
    if iters is not None:
        assert iters.dtype == np.uint32 and iters.shape == (scheme.n_surf, n) and iters.flags.c_contiguous
    rc = (fast_lib() if fast else lib()).oracle_trace_iters(_p(table), scheme.n_surf, _p(pool), _p(beam), seed,
                                                            begin, n,
                                  *[_p(r) for r in rows], _p(flag, _i8p), f,
                                  _p(H, _u32p), _p(F), _p(st, _u64p), _p(D), _p(DF, _i8p), nthreads,
                                  _p(iters, _u32p), int(iters_mode) if iters is not None else 0)
    return TraceResult(R, flag, H, F, st, D, DF, rc)


def stage(name, *args):
    return getattr(lib(), "oracle_" + name)(*args)


# ---- per-stage wrappers over (7, n) or (6, n) ray arrays ----------------------

def _rows(rays):
    R = np.ascontiguousarray(rays, dtype=np.float64).copy()
    return R, [R[k] for k in range(R.shape[0])]


def transform(angles, dist, rays):
    R, rows = _rows(rays)
    a = np.ascontiguousarray(angles, dtype=np.float64)
    d = np.ascontiguousarray(dist, dtype=np.float64)
    lib().oracle_beam_transform(R.shape[1], _p(a), _p(d), *[_p(r) for r in rows[:6]])
    return R


def normalize(rays):
    R, rows = _rows(rays)
    lib().oracle_normalize(R.shape[1], _p(rows[3]), _p(rows[4]), _p(rows[5]))
    return R


def intersect(rec, rays):
    R, rows = _rows(rays)
    n = R.shape[1]
    rec = np.ascontiguousarray(rec, dtype=np.float64)
    flag = np.ones(n, dtype=np.int8)
    nb = np.zeros(1, dtype=np.uint64)
    it = np.zeros(1, dtype=np.uint64)
    lib().oracle_surface_intersection(_p(rec), n, *[_p(r) for r in rows[:6]], _p(nb, _u64p),
                                      _p(flag, _i8p), _p(it, _u64p))
    return R, flag, int(nb[0]), int(it[0])


def mask(rec, X, Y):
    rec = np.ascontiguousarray(rec, dtype=np.float64)
    X = np.ascontiguousarray(X, dtype=np.float64)
    Y = np.ascontiguousarray(Y, dtype=np.float64)
    flag = np.ones(len(X), dtype=np.int8)
    nb = np.zeros(1, dtype=np.uint64)
    lib().oracle_apply_mask(_p(rec), len(X), _p(X), _p(Y), _p(nb, _u64p), _p(flag, _i8p))
    return flag


def reflect(rec, rays):
    R, rows = _rows(rays)
    rec = np.ascontiguousarray(rec, dtype=np.float64)
    lib().oracle_surface_reflection(_p(rec), R.shape[1], *[_p(r) for r in rows[:6]])
    return R


def refract(rec, rays, mu):
    R, rows = _rows(rays)
    n = R.shape[1]
    rec = np.ascontiguousarray(rec, dtype=np.float64)
    mu = np.ascontiguousarray(np.broadcast_to(mu, (n,)), dtype=np.float64)
    flag = np.ones(n, dtype=np.int8)
    nb = np.zeros(1, dtype=np.uint64)
    lib().oracle_surface_refraction(_p(rec), n, *[_p(r) for r in rows[:6]], _p(mu), _p(nb, _u64p),
                                    _p(flag, _i8p))
    return R, flag


def diffract(rec, rays, mu, L):
    R, rows = _rows(rays)
    n = R.shape[1]
    rec = np.ascontiguousarray(rec, dtype=np.float64)
    mu = np.ascontiguousarray(np.broadcast_to(mu, (n,)), dtype=np.float64)
    L = np.ascontiguousarray(np.broadcast_to(L, (n,)), dtype=np.float64)
    flag = np.ones(n, dtype=np.int8)
    nb = np.zeros(1, dtype=np.uint64)
    lib().oracle_surface_diffraction(_p(rec), n, *[_p(r) for r in rows[:6]], _p(mu), _p(L),
                                     _p(nb, _u64p), _p(flag, _i8p))
    return R, flag


def ccd_coordinates(X, Y, xsize, ysize, xpix, ypix):
    X = np.ascontiguousarray(X, dtype=np.float64)
    Y = np.ascontiguousarray(Y, dtype=np.float64)
    cx = np.zeros(len(X), dtype=np.uint32)
    cy = np.zeros(len(X), dtype=np.uint32)
    lib().oracle_ccd_coordinates(len(X), _p(X), _p(Y), xsize, ysize, xpix, ypix, _p(cx, _u32p), _p(cy, _u32p))
    return cx, cy


def grating_energy(lam, order, blaze, alpha, gamma, gr_const, rule_start=0.0, rule_stop=1.0):
    lam = np.ascontiguousarray(lam, dtype=np.float64)
    e = np.ones(len(lam))
    lib().oracle_grating_energy_distr(len(lam), _p(lam), order, blaze, alpha, gamma, gr_const,
                                      rule_start, rule_stop, _p(e))
    return e


# ---- Scheme::run's saved beam (rt_hip_trace_record, rt_hip_rearrange,
# rt_hip_beam_order semantics)

EXT_PER_SURF = 12
LOSS_DEAD = 255


class RecordResult:
    """rays (7 x n), flag (n), loss codes (n, uint8), extents (n_surf x 12),
    stats (uint64 counters), status"""

    def __init__(self, rays, flag, loss, extents, stats, status):
        self.rays, self.flag, self.loss, self.extents, self.stats = rays, flag, loss, extents, stats
        self.status = status


def trace_record(scheme, seed=0, n=None, beam=None, nthreads=0):
    """oracle_trace_record: the generated beam of `seed` traced with loss
    codes and the logged extents"""
    table = np.ascontiguousarray(scheme.table, dtype=np.float64)
    pool = np.ascontiguousarray(scheme.pool, dtype=np.float64)
    b = np.ascontiguousarray(scheme.beam.pack() if beam is None else beam, dtype=np.float64)
    n = scheme.beam.n_rays if n is None else n
    rays = np.zeros((7, n))
    flag = np.zeros(n, dtype=np.int8)
    loss = np.zeros(n, dtype=np.uint8)
    ext = np.zeros((scheme.n_surf, EXT_PER_SURF))
    stats = np.zeros(144, dtype=np.uint64)
    u8 = C.POINTER(C.c_uint8)
    rc = lib().oracle_trace_record(_p(table), scheme.n_surf, _p(pool), _p(b), seed, n,
                                   *[_p(rays[k]) for k in range(7)], _p(flag, _i8p), _p(stats, _u64p),
                                   _p(loss, u8), _p(ext), nthreads)
    return RecordResult(rays, flag, loss, ext, stats, rc)


def rearrange(loss):
    """oracle_rearrange: (perm, counts[256])"""
    loss = np.ascontiguousarray(loss, dtype=np.uint8)
    n = len(loss)
    perm = np.zeros(n, dtype=np.uint64)
    counts = np.zeros(256, dtype=np.uint64)
    rc = lib().oracle_rearrange(_p(loss, C.POINTER(C.c_uint8)), n, _p(perm, _u64p), _p(counts, _u64p))
    if rc != 0:
        raise RuntimeError(f"oracle_rearrange failed ({rc})")
    return perm, counts


def renorm_table(scheme):
    """How many more times Beam::Normalize (beam.cpp:527-571) touches a ray
    lost with each code: Scheme::run normalises every ray after each
    surface's transform and, on surfaces other than AUX, after its action
    (scheme.cpp:993,1040); a ray lost at surface s meets that surface's
    second call and every call of the surfaces after it."""
    types = [int(scheme.table[s * 56 + 1]) for s in range(scheme.n_surf)]
    table = np.zeros(256, dtype=np.uint8)
    for s in range(scheme.n_surf):
        later = sum(1 if t == 6 else 2 for t in types[s + 1:])
        for k in range(3):
            table[1 + 3 * s + k] = 1 + later
    return table


def beam_order(perm, loss, renorm, rays):
    """oracle_beam_order: the 7 x n arrays of Beam::Save in file order"""
    n = len(perm)
    out = np.zeros((7, n))
    rays = np.ascontiguousarray(rays, dtype=np.float64)
    u8 = C.POINTER(C.c_uint8)
    lib().oracle_beam_order(n, _p(np.ascontiguousarray(perm, dtype=np.uint64), _u64p),
                            _p(np.ascontiguousarray(loss, dtype=np.uint8), u8),
                            _p(np.ascontiguousarray(renorm, dtype=np.uint8), u8),
                            *[_p(rays[k]) for k in range(7)], _p(out))
    return out


def dat_bytes(n_rays, n_good, n_lambda, ordered):
    """Beam::Save's bytes (beam.cpp:636-715): three size_t, then the seven
    arrays over the good rays, then over the rest"""
    hdr = np.array([n_rays, n_good, n_lambda], dtype=np.uint64).tobytes()
    good = np.ascontiguousarray(ordered[:, :n_good]).tobytes()
    rest = np.ascontiguousarray(ordered[:, n_good:]).tobytes()
    return hdr + good + rest
