"""Parity of the per-stage HIP kernels -- the engine calls behind the
rt_engine_func.h plugin API (beam_random ... compute_ccd_image) -- with the
oracle's per-stage functions, through the C ABI of librt_hip.so.

The chain test walks a whole scheme the way Scheme::run does
(scheme.cpp:983-1046): for every surface, transform, normalize, intersect,
mask, act, normalize, each stage on the rays still alive (the reference hands
the engine N_good rays after Beam::Rearrange).  Each stage gets the oracle's
output of the previous stage as input, so a difference is pinned on the
stage that made it.  Bars (written here): positions 1e-12 relative with the
scheme's length scale as floor, cosines 1e-12 with floor 1, flags and lost-ray
counts bit-exact, CCD pixel indices bit-exact except rays within 1e-9 px of a
pixel edge.  The chain's end state must equal the oracle's fused trace.
"""
import ctypes as C
import math

import numpy as np
import pytest
import torch

from rt_amd import scheme as S
from rt_amd import schemes
from oracle import oracle as O

pytestmark = pytest.mark.gpu

RTOL = 1e-12
_dp = C.POINTER(C.c_double)


class Host:
    """keeps host arrays alive while the engine reads them"""

    def __init__(self):
        self.keep = []

    def __call__(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        self.keep.append(a)
        return a.ctypes.data_as(_dp)


def dev(a, dtype=None):
    t = torch.from_numpy(np.ascontiguousarray(a)).cuda()
    return t if dtype is None else t.to(dtype)


def stream():
    return torch.cuda.current_stream().cuda_stream


def length_scale(sch):
    t = np.asarray(sch.table).reshape(sch.n_surf, S.SURF_STRIDE)
    return max(1.0, float(np.max(np.abs(t[:, S.S_DIST:S.S_DIST + 3]))))


def close(g, r, L, what, rows=6):
    g, r = np.asarray(g), np.asarray(r)
    if g.size == 0:
        return 0.0
    floor = np.array([L, L, L, 1.0, 1.0, 1.0, 0.0])[:rows, None]
    scale = np.maximum(np.maximum(np.abs(g), np.abs(r)), floor)
    e = float(np.max(np.abs(g - r) / scale))
    assert e <= RTOL, f"{what}: relative difference {e:.3e}"
    return e


def check(rc, hip, what):
    assert rc == 0, f"{what} returned {rc}: {hip.last_error()}"


# ---- one stage on the GPU ------------------------------------------------------

def g_transform(hip, row, R):
    h = Host()
    t = dev(R[:6])
    check(hip.lib().rt_hip_stage_transform(h(row[S.S_ANG:S.S_ANG + 3]), h(row[S.S_DIST:S.S_DIST + 3]),
                                           R.shape[1], *[t[k].data_ptr() for k in range(6)], stream()),
          hip, "rt_hip_stage_transform")
    return t.cpu().numpy()


def g_normalize(hip, R):
    t = dev(R[:6])
    check(hip.lib().rt_hip_stage_normalize(R.shape[1], t[3].data_ptr(), t[4].data_ptr(), t[5].data_ptr(),
                                           stream()), hip, "rt_hip_stage_normalize")
    return t.cpu().numpy()


def g_intersect(hip, row, R):
    h = Host()
    n = R.shape[1]
    t = dev(R[:6])
    f = torch.ones(n, dtype=torch.int8, device="cuda")
    cnt = torch.zeros(2, dtype=torch.int64, device="cuda")
    check(hip.lib().rt_hip_stage_intersect(h(row), n, *[t[k].data_ptr() for k in range(6)], f.data_ptr(),
                                           cnt.data_ptr(), cnt.data_ptr() + 8, stream()),
          hip, "rt_hip_stage_intersect")
    c = cnt.cpu().numpy()
    return t.cpu().numpy(), f.cpu().numpy(), int(c[0]), int(c[1])


def g_mask(hip, row, X, Y):
    h = Host()
    n = len(X)
    tx, ty = dev(X), dev(Y)
    f = torch.ones(n, dtype=torch.int8, device="cuda")
    cnt = torch.zeros(1, dtype=torch.int64, device="cuda")
    check(hip.lib().rt_hip_stage_mask(h(row), n, tx.data_ptr(), ty.data_ptr(), f.data_ptr(), cnt.data_ptr(),
                                      stream()), hip, "rt_hip_stage_mask")
    return f.cpu().numpy(), int(cnt.cpu().numpy()[0])


def g_action(hip, row, R, mu=None, Lv=None):
    h = Host()
    n = R.shape[1]
    t = dev(R[:6])
    P = [t[k].data_ptr() for k in range(6)]
    f = torch.ones(n, dtype=torch.int8, device="cuda")
    cnt = torch.zeros(1, dtype=torch.int64, device="cuda")
    typ = int(row[S.S_TYPE])
    L = hip.lib()
    if typ == S.RT_MIRROR:
        check(L.rt_hip_stage_reflect(h(row), n, *P, stream()), hip, "rt_hip_stage_reflect")
    elif typ == S.RT_LENS:
        tm = dev(mu)
        check(L.rt_hip_stage_refract(h(row), n, *P, tm.data_ptr(), f.data_ptr(), cnt.data_ptr(), stream()),
              hip, "rt_hip_stage_refract")
    else:
        tm = dev(mu) if mu is not None else None
        tl = dev(Lv)
        check(L.rt_hip_stage_diffract(h(row), n, *P, tm.data_ptr() if tm is not None else None, tl.data_ptr(),
                                      f.data_ptr(), cnt.data_ptr(), stream()), hip, "rt_hip_stage_diffract")
    return t.cpu().numpy(), f.cpu().numpy(), int(cnt.cpu().numpy()[0])


# ---- host-side per-ray indices, as the plugin computes them ---------------------

def fn_value(ref, lam):
    return ref(lam) if isinstance(ref, S.Spline) else float(ref)


def ray_order(row, lam):
    if int(row[S.S_ORDMODE]) == S.RT_ORDER_FIXED:
        return row[S.S_ORDER]
    m = float(math.floor(row[S.S_BLAZEK] / lam + 0.5))
    return min(max(m, row[S.S_ORDMIN]), row[S.S_ORDMAX])


def host_mu_L(surf, row, lam):
    mu = np.empty(len(lam))
    Lv = np.empty(len(lam))
    for i, l in enumerate(lam):
        n1 = fn_value(surf.n1, l)
        n2 = fn_value(surf.n2, l)
        mu[i] = n1 / n2
        Lv[i] = ray_order(row, l) * l / n2 / row[S.S_GRCONST]
    return mu, Lv


# ---- the chain -----------------------------------------------------------------

def beam_on_gpu(hip, sch, seed):
    h = Host()
    desc = h(sch.beam.pack())
    n = sch.beam.n_rays
    t = torch.zeros((7, n), dtype=torch.float64, device="cuda")
    P = [t[k].data_ptr() for k in range(7)]
    L = hip.lib()
    check(L.rt_hip_beam_positions(desc, seed, 0, n, P[0], P[1], P[2], stream()), hip, "rt_hip_beam_positions")
    check(L.rt_hip_beam_directions(desc, seed, 0, n, P[0], P[1], P[3], P[4], P[5], stream()), hip,
          "rt_hip_beam_directions")
    check(L.rt_hip_beam_wavelengths(desc, seed, 0, n, P[6], stream()), hip, "rt_hip_beam_wavelengths")
    return t.cpu().numpy()


def run_chain(hip, sch, seed, rays=None, flag=None, iters_exact=True):
    """Walk the scheme stage by stage on the GPU and the oracle (see the
    module docstring).  rays/flag: given input rays instead of the generated
    beam.  iters_exact=False leaves the Newton iteration sums unchecked (for
    rays whose Newton iterates wander chaotically until a give-up)."""
    Ls = length_scale(sch)
    if rays is None:
        n = sch.beam.n_rays
        print(f"[stage chain] {sch.name} n={n} seed={seed}", flush=True)
        ref = O.gen_rays(sch.beam.pack(), seed, 0, n)
        g = beam_on_gpu(hip, sch, seed)
        close(g, ref, Ls, "beam generation", rows=7)
        alive = np.ones(n, dtype=bool)
    else:
        n = rays.shape[1]
        print(f"[stage chain] {sch.name} n={n} (rays given)", flush=True)
        ref = np.array(rays, dtype=np.float64)
        alive = np.ones(n, dtype=bool) if flag is None else np.asarray(flag) != 0
    state = ref.copy()
    t = np.asarray(sch.table).reshape(sch.n_surf, S.SURF_STRIDE)
    for s, surf in enumerate(sch.surfaces):
        row = t[s]
        typ = int(row[S.S_TYPE])
        idx = np.nonzero(alive)[0]
        cur = state[:, idx].copy()
        tag = f"surface {s} ({surf.comment})"
        r = O.transform(row[S.S_ANG:S.S_ANG + 3], row[S.S_DIST:S.S_DIST + 3], cur[:6])
        close(g_transform(hip, row, cur), r, Ls, tag + " transform")
        cur[:6] = r
        r = O.normalize(cur[:6])
        close(g_normalize(hip, cur), r, Ls, tag + " normalize")
        cur[:6] = r
        keep = np.ones(len(idx), dtype=bool)
        if typ != S.RT_AUXTYPE:
            r, rf, rnb, rit = O.intersect(row, cur[:6])
            gR, gf, gnb, git = g_intersect(hip, row, cur)
            assert np.array_equal(gf, rf), f"{tag}: intersection flags differ"
            assert gnb == rnb, f"{tag}: N_bad {gnb} != {rnb}"
            assert git == rit or not iters_exact, f"{tag}: Newton iterations {git} != {rit}"
            close(gR[:, rf == 1], r[:, rf == 1], Ls, tag + " intersection")
            cur[:6] = r
            keep &= rf == 1
            k = np.nonzero(keep)[0]
            rm = O.mask(row, cur[0, k], cur[1, k])
            gm, gmb = g_mask(hip, row, cur[0, k], cur[1, k])
            assert np.array_equal(gm, rm), f"{tag}: mask flags differ"
            assert gmb == int(np.sum(rm == 0))
            keep[k[rm == 0]] = False
            k = np.nonzero(keep)[0]
            sub = cur[:, k]
            ok = np.ones(len(k), dtype=bool)
            if typ in (S.RT_MIRROR, S.RT_LENS, S.RT_GRATING) and len(k):
                mu, Lv = host_mu_L(surf, row, sub[6])
                if typ == S.RT_MIRROR:
                    r, rf = O.reflect(row, sub[:6]), np.ones(len(k), dtype=np.int8)
                    gR, gf, gnb = g_action(hip, row, sub)
                elif typ == S.RT_LENS:
                    r, rf = O.refract(row, sub[:6], mu)
                    gR, gf, gnb = g_action(hip, row, sub, mu=mu)
                else:
                    transparent = int(row[S.S_GRTYPE]) == S.RT_TRANSPARENT
                    r, rf = O.diffract(row, sub[:6], mu, Lv)
                    gR, gf, gnb = g_action(hip, row, sub, mu=mu if transparent else None, Lv=Lv)
                assert np.array_equal(gf, rf), f"{tag}: action flags differ"
                assert gnb == int(np.sum(rf == 0))
                close(gR[:, rf == 1], r[:, rf == 1], Ls, tag + " action")
                sub[:6] = r
                keep[k[rf == 0]] = False
                ok = rf == 1
            if ok.any():                           # Beam::Normalize after every non-Aux surface (:1040)
                rn = O.normalize(sub[:6, ok])
                close(g_normalize(hip, sub[:, ok]), rn, Ls, tag + " normalize after action")
                sub[:6, ok] = rn
            cur[:, k] = sub
        state[:, idx] = cur
        alive[idx[~keep]] = False
    return state, alive


@pytest.mark.parametrize("make,seed", [(lambda: schemes.czerny_turner(n_per_lambda=200, nlambda=5), 3),
                                       (lambda: schemes.echelle(n_per_lambda=16, nlambda=64), 5)])
def test_stage_chain_matches_oracle_and_fused_trace(hip, make, seed):
    sch = make()
    state, alive = run_chain(hip, sch, seed)
    fused = O.trace(sch, seed=seed)
    assert np.array_equal(alive, fused.flag == 1), "stage chain and fused trace lose different rays"
    assert np.array_equal(state[:, alive], fused.rays[:, alive]), "stage chain differs from the fused trace"
    assert alive.sum() > 0.5 * len(alive)


def test_ccd_stages(hip):
    sch = schemes.echelle(n_per_lambda=16, nlambda=64)
    res = O.trace(sch, seed=2)
    ok = res.flag == 1
    X, Y = res.rays[0, ok], res.rays[1, ok]
    det = sch.detector().ccd
    xs, ys = det["xdim"] * det["xpix"], det["ydim"] * det["ypix"]
    # stretch the spot past the chip edges on both axes (about a fifth of the rays off-chip)
    X = (X - X.mean()) * (1.2 * xs / max(1e-9, np.ptp(X))) + 0.05 * xs
    Y = (Y - Y.mean()) * (1.2 * ys / max(1e-9, np.ptp(Y))) - 0.05 * ys
    n = len(X)
    rx, ry = O.ccd_coordinates(X, Y, xs, ys, det["xpix"], det["ypix"])
    tx, ty = dev(X), dev(Y)
    cx = torch.zeros(n, dtype=torch.int32, device="cuda")
    cy = torch.zeros(n, dtype=torch.int32, device="cuda")
    L = hip.lib()
    check(L.rt_hip_ccd_coordinates(n, tx.data_ptr(), ty.data_ptr(), xs, ys, det["xpix"], det["ypix"],
                                   cx.data_ptr(), cy.data_ptr(), stream()), hip, "rt_hip_ccd_coordinates")
    gx = cx.cpu().numpy().view(np.uint32)
    gy = cy.cpu().numpy().view(np.uint32)
    vx = (X + xs / 2.0) / det["xpix"]
    vy = (Y + ys / 2.0) / det["ypix"]
    edge = lambda v: np.minimum(v - np.floor(v), np.ceil(v) - v) < 1e-9
    sure = ~(edge(vx) | edge(vy))
    assert np.array_equal(gx[sure], rx[sure]) and np.array_equal(gy[sure], ry[sure])
    # accumulate: hit counts and weighted image, off-chip rays counted
    xdim, ydim = int(det["xdim"]), int(det["ydim"])
    w = np.random.default_rng(0).uniform(0.5, 2.0, n)
    hist = torch.zeros(xdim * ydim, dtype=torch.int32, device="cuda")
    img = torch.zeros(xdim * ydim, dtype=torch.float64, device="cuda")
    off = torch.zeros(1, dtype=torch.int64, device="cuda")
    tw = dev(w)
    check(L.rt_hip_ccd_accumulate(n, cx.data_ptr(), cy.data_ptr(), tw.data_ptr(), xdim, ydim, img.data_ptr(),
                                  hist.data_ptr(), off.data_ptr(), stream()), hip, "rt_hip_ccd_accumulate")
    on = (gx < xdim) & (gy < ydim)
    pix = gy[on].astype(np.int64) * xdim + gx[on]
    ref_h = np.bincount(pix, minlength=xdim * ydim)
    ref_i = np.zeros(xdim * ydim)
    np.add.at(ref_i, pix, w[on])
    assert 0 < on.sum() < n
    assert np.array_equal(hist.cpu().numpy(), ref_h)
    assert int(off.cpu().numpy()[0]) == int((~on).sum())
    gi = img.cpu().numpy()
    assert np.all(np.abs(gi - ref_i) <= 1e-12 * np.maximum(1.0, np.abs(ref_i)))


def test_spline_scale_and_grating_energy(hip):
    L = hip.lib()
    sp = S.glass_spline("N-BK7")
    rec = np.concatenate([[len(sp.x)], sp.x, sp.y, sp.y2])
    xs = np.concatenate([np.linspace(0.30, 1.0, 5001), sp.x])          # knots and out-of-range points
    tx = dev(xs)
    ty = torch.zeros(len(xs), dtype=torch.float64, device="cuda")
    oob = torch.zeros(1, dtype=torch.int64, device="cuda")
    h = Host()
    check(L.rt_hip_spline_eval(h(rec), len(xs), tx.data_ptr(), ty.data_ptr(), oob.data_ptr(), stream()), hip,
          "rt_hip_spline_eval")
    gy = ty.cpu().numpy()
    ref = np.array([O.spline_eval(sp.x, sp.y, sp.y2, v) for v in xs], dtype=object)
    rv = np.array([v for v, e in ref])
    re = np.array([e for v, e in ref])
    assert int(oob.cpu().numpy()[0]) == int(np.sum(re != 0))
    good = re == 0
    assert np.all(np.abs(gy[good] - rv[good]) <= 1e-13 * np.abs(rv[good]))
    assert np.all(gy[~good] == 0.0)
    # y2 from the host recurrence equals the oracle's
    y2 = np.zeros(len(sp.x))
    check(L.rt_hip_spline_init(len(sp.x), h(sp.x), h(sp.y), y2.ctypes.data_as(_dp)), hip, "rt_hip_spline_init")
    assert np.array_equal(y2, O.spline_init(sp.x, sp.y))
    # surface_QE product
    v = np.linspace(1.0, 2.0, 777)
    f = np.linspace(0.1, 0.9, 777)
    tv, tf = dev(v), dev(f)
    check(L.rt_hip_scale(len(v), tv.data_ptr(), tf.data_ptr(), stream()), hip, "rt_hip_scale")
    assert np.array_equal(tv.cpu().numpy(), v * f)
    # grating blaze distribution (echelle, 31.6 gr/mm, blaze 63.4 deg, Littrow)
    lam = np.linspace(0.38, 0.90, 2048)
    d = 1000.0 / 31.6
    blaze = math.radians(63.4)
    for order in (70, 100, 140):
        ref_e = O.grating_energy(lam, order, blaze, blaze, 0.0, d, 0.0, 1.0)
        tl = dev(lam)
        te = torch.ones(len(lam), dtype=torch.float64, device="cuda")
        check(L.rt_hip_grating_energy(len(lam), tl.data_ptr(), order, blaze, blaze, 0.0, d, 0.0, 1.0,
                                      te.data_ptr(), stream()), hip, "rt_hip_grating_energy")
        ge = te.cpu().numpy()
        # where the order does not propagate (|sin beta| > 1) both give NaN, as the reference does
        nan = np.isnan(ref_e)
        assert np.array_equal(np.isnan(ge), nan), order
        assert nan.sum() < len(lam)
        ok = ~nan
        # The formula is ill-conditioned near grazing orders: asin(m lambda/d - sin alpha)
        # with an argument near +-1 amplifies a one-ulp difference between two math
        # libraries' sin/asin by 1/sqrt(1-x^2); perturbing sin(alpha) or the asin argument
        # by one ulp on the host moves these values by up to 2.9e-11 of the peak.  Bar:
        # 1e-10 of the distribution's peak, absolute.
        err = np.abs(ge[ok] - ref_e[ok])
        peak = float(np.max(np.abs(ref_e[ok])))
        assert float(np.max(err)) <= 1e-10 * peak, (order, float(np.max(err)), peak)


@pytest.mark.parametrize("btype,ang", [(S.RT_BEAM_PARALLEL, 0.0), (S.RT_BEAM_GAUSS, 2.35482), (S.RT_BEAM_GAUSS, 3600.0),
                                       (S.RT_BEAM_CONIC, 3.0)],
                         ids=["parallel", "gauss", "gauss-wide", "conic"])
def test_beam_angles_match_libm_within_ulps(hip, btype, ang):
    """The beam generators' sin/cos (disk angle, Box-Muller angle, conic
    azimuth) use a bounded-range reduction and polynomial of their own; the
    positions and direction cosines agree with the oracle's (C libm) within a
    few ulps of their scale: |dX|, |dY| <= 1e-15 x outer radius, |dc| <= 1e-15.
    The Gaussian cases take both forms of 1/sqrt(1 + t^2): the series of a
    narrow beam (FWHM 2.35482 arcsec, tan sigma ~ 5e-6) and the reciprocal
    square root of a wide one (FWHM 1 degree, tan sigma ~ 7e-3)."""
    sch = schemes.czerny_turner(40000, 5)
    sch.beam.type = btype
    p = list(sch.beam.params) + [0.0] * (4 - len(sch.beam.params))
    p[3] = ang
    sch.beam.params = tuple(p)
    print(f"[beam angles] type {btype}", flush=True)
    ref = O.gen_rays(sch.beam.pack(), 11, 0, sch.beam.n_rays)
    g = beam_on_gpu(hip, sch, 11)
    R = max(float(sch.beam.params[2]), 1.0)
    assert float(np.max(np.abs(g[0:2] - ref[0:2]))) <= 1e-15 * R
    assert float(np.max(np.abs(g[3:6] - ref[3:6]))) <= 1e-15
    assert np.array_equal(g[6], ref[6])
