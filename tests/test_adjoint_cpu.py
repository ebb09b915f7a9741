"""The adjoint's derivative formulas (rt_amd/csrc/rt_adjoint.h, built for the
host) against central finite differences of the oracle's forward trace, no
GPU.  The rays' entry states come from the oracle's per-surface dump, so the
surface chain is reversed over the oracle's own forward states.

Objective J = sum over surviving rays of (u_r X_r + v_r Y_r) on the last
surface, u, v fixed random weights.  Bar: |adjoint - FD| <= 1e-6 |FD| + the
FD's own rounding noise (4 eps sum|terms| / 2h), with the ray set, the
vignetting and the Newton iteration counts unchanged between the three runs
(checked through the oracle's counters).
"""
import ctypes as C
import math
import os
import subprocess

import numpy as np
import pytest

from rt_amd import scheme as S
from rt_amd import schemes
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "native", "adjoint_host.cpp")
LIB = os.path.join(ROOT, "tests", "native", "libadjoint_host.so")
HDR = os.path.join(ROOT, "rt_amd", "csrc", "rt_adjoint.h")
_dp = C.POINTER(C.c_double)


@pytest.fixture(scope="module")
def host():
    if (not os.path.exists(LIB) or os.path.getmtime(LIB) < os.path.getmtime(SRC)
            or os.path.getmtime(LIB) < os.path.getmtime(HDR)):
        subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall",
                        "-I", os.path.join(ROOT, "rt_amd", "csrc"), "-I", os.path.join(ROOT, "include"),
                        "-o", LIB, SRC], check=True)
    L = C.CDLL(LIB)
    L.adj_host_surface.argtypes = [_dp, _dp, C.c_int, C.c_double, C.c_double, _dp, _dp]
    L.adj_host_slot.argtypes = [C.c_int]
    return L


def fn_value(ref, lam):
    return ref(lam) if isinstance(ref, S.Spline) else float(ref)


def ray_order(row, lam):
    if int(row[S.S_ORDMODE]) == S.RT_ORDER_FIXED:
        return row[S.S_ORDER]
    m = float(math.floor(row[S.S_BLAZEK] / lam + 0.5))
    return min(max(m, row[S.S_ORDMIN]), row[S.S_ORDMAX])


def chain_adjoint(host, sch, seed, u, v):
    """dJ/d(record field) for every surface and slot, reversing the chain per ray"""
    n_pg = host.adj_host_pg_n()
    t = np.asarray(sch.table).reshape(sch.n_surf, S.SURF_STRIDE)
    r = O.trace(sch, seed=seed, dump=True)
    beam = O.gen_rays(sch.beam.pack(), seed, 0, sch.beam.n_rays)
    G = np.zeros((sch.n_surf, n_pg))
    for i in np.nonzero(r.flag == 1)[0]:
        a = np.zeros(6)
        a[0], a[1] = u[i], v[i]
        lam = beam[6, i]
        for s in range(sch.n_surf - 1, -1, -1):
            entry = np.ascontiguousarray(beam[:6, i] if s == 0 else r.dump[s - 1][:6, i])
            row = np.ascontiguousarray(t[s])
            surf = sch.surfaces[s]
            mu = Lg = 0.0
            if surf.type in (S.RT_LENS, S.RT_GRATING):
                n1, n2 = fn_value(surf.n1, lam), fn_value(surf.n2, lam)
                mu = n1 / n2
                if surf.type == S.RT_GRATING:
                    Lg = ray_order(row, lam) * lam / n2 / row[S.S_GRCONST]
            gp = np.zeros(n_pg)
            rc = host.adj_host_surface(row.ctypes.data_as(_dp), entry.ctypes.data_as(_dp), 0, mu, Lg,
                                       a.ctypes.data_as(_dp), gp.ctypes.data_as(_dp))
            assert rc == 0, (i, s)
            G[s] += gp
    return G, r


def counters(st, n_surf):
    return np.array([st[S.ST_SURF + 4 * s + k] for s in range(n_surf) for k in (1, 2, 3)])


CASES = [(0, S.S_C), (2, S.S_C), (2, S.S_K), (5, S.S_C), (5, S.S_K), (5, S.S_A + 1), (5, S.S_A + 4),
         (6, S.S_C), (7, S.S_C), (7, S.S_CR), (8, S.S_A + 2), (9, S.S_C), (9, S.S_K), (10, S.S_C),
         (3, S.S_GRCONST), (4, S.S_GRCONST), (2, S.S_DIST + 0), (5, S.S_DIST + 1), (11, S.S_DIST + 0),
         (3, S.S_ANG + 0), (4, S.S_ANG + 1), (7, S.S_ANG + 0), (11, S.S_ANG + 1)]


def test_surface_adjoint_matches_finite_differences(host):
    sch = schemes.echelle(n_per_lambda=3, nlambda=64)
    seed = 3
    rng = np.random.default_rng(0)
    u, v = rng.normal(size=sch.beam.n_rays), rng.normal(size=sch.beam.n_rays)
    G, base = chain_adjoint(host, sch, seed, u, v)
    ok = base.flag == 1
    t0 = np.asarray(sch.table).reshape(sch.n_surf, S.SURF_STRIDE)
    checked = 0
    for s, f in CASES:
        x = t0[s, f]
        h = 1e-5 * abs(x) if x != 0.0 else 1e-6
        J, flags, cnt, mag = [], [], [], 0.0
        for sgn in (1, -1):
            tt = t0.copy()
            tt[s, f] = x + sgn * h
            if S.S_ANG <= f < S.S_ANG + 3:
                j = f - S.S_ANG
                tt[s, S.S_SIN + j], tt[s, S.S_COS + j] = math.sin(tt[s, f]), math.cos(tt[s, f])
            pert = S.Scheme.__new__(S.Scheme)
            pert.__dict__.update(sch.__dict__)
            pert.table = tt.reshape(-1)
            r = O.trace(pert, seed=seed)
            terms = (u * r.rays[0] + v * r.rays[1])[ok]
            J.append(float(np.sum(terms)))
            mag = max(mag, float(np.sum(np.abs(terms))))
            flags.append(r.flag)
            cnt.append(counters(r.stats, sch.n_surf))
        same = all(np.array_equal(fl, base.flag) for fl in flags) and \
            all(np.array_equal(c, counters(base.stats, sch.n_surf)) for c in cnt)
        if not same:
            continue
        fd = (J[0] - J[1]) / (2 * h)
        ad = G[s, host.adj_host_slot(f)]
        noise = 4 * np.finfo(float).eps * mag / (2 * h)
        assert abs(ad - fd) <= 1e-6 * abs(fd) + noise, (s, f, ad, fd, noise)
        checked += 1
    assert checked >= len(CASES) - 2
