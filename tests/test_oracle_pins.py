"""Pins of the CPU oracle (oracle/rt_oracle.c) against published vectors,
libraries and closed forms.  The reference (amirych/rt) holds no tests,
fixtures or golden data and its engine cannot be built here, so every
oracle function is pinned by something independent of its own text:

  Philox4x32-10      Random123 known-answer vectors; rocRAND's generator
  u01 mapping        exact end points
  spline             scipy CubicSpline(bc_type='natural'); exact linear data
  beam generation    distribution tests (KS) of the reference's formulas
  transform          rotation matrices composed in numpy
  intersections      textbook ray-sphere, sag formula, brentq roots
  masks              hand-placed edge rays
  reflection         law of reflection about the analytic normal
  refraction         Snell tangential continuity, TIR threshold
  gratings           grating equation, Littrow, zero order = mirror/lens
  CCD mapping        brute-force truncation
  stage order        fused trace == composition of the per-stage functions
"""
import math

import numpy as np
import pytest
from scipy import interpolate, optimize, stats

from oracle import oracle as O
from rt_amd import scheme as S

GOLDEN = __file__.rsplit("/", 1)[0] + "/golden/"


def rec(**kw):
    """packed record of one surface"""
    t, _ = S.pack_surfaces([S.Surface(**kw)])
    return t[:S.SURF_STRIDE]


def rays_from(P, C):
    P = np.atleast_2d(P).T if np.ndim(P) == 1 else np.asarray(P).T
    C = np.atleast_2d(C).T if np.ndim(C) == 1 else np.asarray(C).T
    C = C / np.linalg.norm(C, axis=0)
    return np.vstack([P, C, np.full((1, P.shape[1]), 0.5)])


# ---------------------------------------------------------------- Philox

def _read_vectors(path):
    out = []
    for line in open(path):
        if line.startswith("#") or not line.strip():
            continue
        out.append([int(w, 16) for w in line.split()])
    return out


def test_philox_published_kat():
    for v in _read_vectors(GOLDEN + "philox_kat.txt"):
        assert list(O.philox(v[:4], v[4:6])) == v[6:10]


def test_philox_matches_rocrand():
    vs = _read_vectors(GOLDEN + "philox_rocrand.txt")
    assert len(vs) == 64
    for v in vs:
        assert list(O.philox(v[:4], v[4:6])) == v[6:10]


def test_u01_end_points():
    L = O.lib()
    assert L.oracle_u01(0, 0) == 2.0 ** -53
    assert L.oracle_u01(0xFFFFFFFF, 0xFFFFFFFF) == 1.0
    assert L.oracle_u01(0, 2048) == 2.0 * 2.0 ** -53
    assert L.oracle_u01(0x80000000, 0) == 0.5 + 2.0 ** -53


# ---------------------------------------------------------------- splines

def test_spline_matches_scipy_natural():
    rng = np.random.default_rng(1)
    for K in (4, 7, 64, 200):
        x = np.sort(rng.uniform(0.3, 1.0, K))
        y = np.sin(5 * x) + rng.normal(0, 0.1, K)
        y2 = O.spline_init(x, y)
        cs = interpolate.CubicSpline(x, y, bc_type="natural")
        assert np.allclose(y2[:-1], 2.0 * cs.c[1], rtol=1e-9, atol=1e-9 * np.max(np.abs(y2)))
        assert y2[0] == 0.0 and y2[-1] == 0.0
        # the package's own table builder uses the same recurrence, bit for bit
        assert np.array_equal(S.natural_spline_y2(x, y), y2)
        for xv in rng.uniform(x[0], x[-1], 50):
            v, err = O.spline_eval(x, y, y2, xv)
            assert err == 0
            assert abs(v - cs(xv)) <= 1e-12 * max(1.0, abs(v))


def test_spline_exact_cases():
    x = np.array([0.0, 0.5, 1.25, 2.0, 3.0])
    y = 3.0 * x - 1.0                                   # linear data: y2 == 0
    y2 = O.spline_init(x, y)
    assert np.all(np.abs(y2) < 1e-14)
    for xv in (0.1, 0.77, 1.9, 2.999):
        assert abs(O.spline_eval(x, y, y2, xv)[0] - (3 * xv - 1)) < 1e-14
    for k in range(5):                                  # knots are returned exactly
        assert O.spline_eval(x, y, y2, x[k]) == (y[k], 0)
    assert O.spline_eval(x, y, y2, -1e-12)[1] == 1     # out of range (tabulated_function.cpp:179)
    assert O.spline_eval(x, y, y2, 3.0 + 1e-12)[1] == 1


# ---------------------------------------------------------------- beams

def _beam(**kw):
    return S.Beam(**kw).pack()


def test_disk_positions_follow_reference_radius_law():
    # random_circle_kernel (beam.cu:141): R = (R2-R1) sqrt(u1) + R1, phi = PI2 u2
    b = _beam(type=S.RT_BEAM_PARALLEL, params=(20000, 10.0, 20.0), center=(1.0, -2.0, 3.0),
              range_distr=S.RT_RANGE_MONO, range=(0.5, 0.6))
    R = O.gen_rays(b, 5, 0, 20000)
    dx, dy = R[0] - 1.0, R[1] + 2.0
    r = np.hypot(dx, dy)
    assert np.all(R[2] == 3.0)
    assert r.min() >= 10.0 - 1e-12 and r.max() <= 20.0 + 1e-12
    assert stats.kstest(((r - 10.0) / 10.0) ** 2, "uniform").pvalue > 1e-3
    phi = np.mod(np.arctan2(dy, dx), 2 * np.pi) / (2 * np.pi)
    assert stats.kstest(phi, "uniform").pvalue > 1e-3
    assert np.all(R[5] == 1.0) and np.all(R[3] == 0.0) and np.all(R[4] == 0.0)
    assert np.all(R[6] == 0.55)                         # monochromatic: (r1+r0)/2 (beam.cpp:436)


def test_rectangle_positions_uniform():
    b = _beam(type=S.RT_BEAM_PARALLEL, shape=S.RT_RECTANGLE, params=(20000, 0.0, 40.0, 2.0),
              range_distr=S.RT_RANGE_MONO)
    R = O.gen_rays(b, 2, 0, 20000)
    assert np.all(np.abs(R[0]) <= 20.0) and np.all(np.abs(R[1]) <= 10.0)
    assert stats.kstest(R[0] / 40.0 + 0.5, "uniform").pvalue > 1e-3
    assert stats.kstest(R[1] / 20.0 + 0.5, "uniform").pvalue > 1e-3


def test_gauss_directions():
    # beam.cu:1099 sigma = FWHM/2.35482 in rad; cX, cY ~ N(0, tan sigma) composed (:952-991)
    fwhm = 2.35482 * 3600.0                             # sigma = 1 deg
    b = _beam(type=S.RT_BEAM_GAUSS, params=(40000, 0.0, 5.0, fwhm), range_distr=S.RT_RANGE_MONO)
    R = O.gen_rays(b, 9, 0, 40000)
    c = R[3:6]
    assert np.allclose(np.sum(c * c, axis=0), 1.0, atol=4e-16)
    tX = -c[1] / np.hypot(c[0], c[2])
    tY = -c[0] / c[2]
    sig = math.tan(fwhm / 2.35482 * S.RT_PI / 180 / 3600)
    for t in (tX, tY):
        assert abs(np.std(t) / sig - 1.0) < 0.02
        assert stats.kstest(t / sig, "norm").pvalue > 1e-3
    assert abs(np.corrcoef(tX, tY)[0, 1]) < 0.03


def test_conic_directions():
    b = _beam(type=S.RT_BEAM_CONIC, params=(500, 0.0, 5.0, 30.0), range_distr=S.RT_RANGE_MONO)
    R = O.gen_rays(b, 1, 0, 500)
    half = 30.0 * S.RT_PI / 180.0 / 2.0              # beam.cpp:383 macro expansion
    assert np.all(R[5] == math.cos(half))
    az = np.arctan2(R[4], R[3])
    assert np.allclose(np.cos(az - np.arctan2(R[1], R[0])), 1.0, atol=1e-12)
    assert np.allclose(np.hypot(R[3], R[4]), math.sin(half), atol=1e-15)


def test_wavelength_grids():
    b = _beam(params=(3, 0.0, 1.0), range=(0.4, 0.7), nlambda=5, composite=True)
    R = O.gen_rays(b, 0, 0, 15)
    grid = 0.4 + np.arange(5) * ((0.7 - 0.4) / 4)      # beam.cu:726 step, beam.cpp:470-484 tiling
    assert np.array_equal(R[6], np.tile(grid, 3))
    b = _beam(params=(6, 0.0, 1.0), range=(0.4, 0.7), range_distr=S.RT_RANGE_RANDOM)
    lam = O.gen_rays(b, 3, 0, 6)[6]
    assert np.all((lam > 0.4) & (lam <= 0.7))


# ---------------------------------------------------------------- transforms

def _rot(axis, a):
    c, s = math.cos(a), math.sin(a)
    if axis == "Y":       # (X, Z) -> (X c - Z s, X s + Z c)   beam.cu:340-346, :492-495
        return np.array([[c, 0, -s], [0, 1, 0], [s, 0, c]])
    if axis == "X":       # (Y, Z)
        return np.array([[1, 0, 0], [0, c, -s], [0, s, c]])
    return np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]])


def test_transform_matches_rotation_matrices():
    rng = np.random.default_rng(4)
    P = rng.normal(0, 10, (20, 3))
    Cv = rng.normal(0, 1, (20, 3))
    R0 = rays_from(P, Cv)
    ang = np.array([0.3, -0.7, 1.1])
    d = np.array([1.5, -2.0, 30.0])
    R = O.transform(ang, d, R0)
    M = _rot("Z", ang[2]) @ _rot("X", ang[1]) @ _rot("Y", ang[0])
    assert np.allclose(R[0:3], M @ (R0[0:3] - d[:, None]), rtol=0, atol=1e-13)
    assert np.allclose(R[3:6], M @ R0[3:6], rtol=0, atol=1e-15)
    # quarter turn about Y maps the X axis onto +Z
    e = O.transform([math.pi / 2, 0, 0], [0, 0, 0], rays_from([[1.0, 0, 0]], [[1.0, 0, 0]]))
    assert np.allclose(e[:6, 0], [0, 0, 1, 0, 0, 1], atol=1e-15)


def test_normalize_unit_and_direction():
    rng = np.random.default_rng(5)
    R = np.vstack([np.zeros((3, 100)), rng.normal(0, 3, (3, 100)), np.ones((1, 100))])
    N = O.normalize(R)
    assert np.allclose(np.sum(N[3:6] ** 2, axis=0), 1.0, atol=3e-16)
    assert np.allclose(N[3:6] * np.linalg.norm(R[3:6], axis=0), R[3:6], rtol=1e-15)


# ---------------------------------------------------------------- intersections

def test_plane_intersection():
    r = rec(cls=S.RT_PLANE, type=S.RT_STOP, size=(100, 0))
    R0 = rays_from([[1, 2, -5], [0, 0, 5]], [[0.1, 0.2, 1.0], [0, 0, 1.0]])
    R, flag, nbad, _ = O.intersect(r, R0)
    s0 = 5.0 / R0[5, 0]
    assert np.allclose(R[:3, 0], [1 + R0[3, 0] * s0, 2 + R0[4, 0] * s0, 0.0], atol=1e-14)
    assert list(flag) == [1, 0] and nbad == 1          # s < 0 behind the ray: lost, unmoved
    assert np.array_equal(R[:, 1], R0[:, 1])


def _sphere_hit(p, d, Rad):
    """textbook ray-sphere: vertex at origin, center (0,0,R); root nearest the vertex"""
    ctr = np.array([0, 0, Rad])
    oc = p - ctr
    b = np.dot(oc, d)
    disc = b * b - (np.dot(oc, oc) - Rad * Rad)
    ts = (-b - math.sqrt(disc), -b + math.sqrt(disc))
    return min((p + t * d for t in ts), key=lambda q: abs(q[2]))


@pytest.mark.parametrize("Rad", [-800.0, 250.0])
def test_sphere_intersection_textbook(Rad):
    rng = np.random.default_rng(6)
    r = rec(cls=S.RT_CONIC, type=S.RT_MIRROR, params=(1.0 / Rad, 0.0, 0), size=(100, 0))
    P = np.column_stack([rng.uniform(-40, 40, 50), rng.uniform(-40, 40, 50), np.full(50, -30.0)])
    Cv = np.column_stack([rng.normal(0, 0.1, 50), rng.normal(0, 0.1, 50), np.ones(50)])
    R0 = rays_from(P, Cv)
    R, flag, _, _ = O.intersect(r, R0)
    assert np.all(flag == 1)
    for i in range(50):
        q = _sphere_hit(R0[:3, i], R0[3:6, i], Rad)
        assert np.allclose(R[:3, i], q, rtol=0, atol=1e-11)


def _sag(r2, c, k, a):
    return c * r2 / (1 + np.sqrt(1 - (1 + k) * c * c * r2)) + sum(ai * r2 ** (i + 1) for i, ai in enumerate(a))


@pytest.mark.parametrize("c,k,a", [(-1 / 2000.0, -1.0, []),                 # paraboloid, closed form
                                   (1 / 200.0, -0.3, [0, -2e-9, 1e-13, -1e-17, 1e-21]),  # Newton
                                   (0.0, 0.0, [1e-4, 2e-8]),                  # c = 0 asphere (D6)
                                   (-1 / 150.0, 0.0, [0, 3e-9, -2e-13])])
def test_conic_axis_parallel_hits_the_sag(c, k, a):
    rng = np.random.default_rng(7)
    r = rec(cls=S.RT_CONIC, type=S.RT_LENS, params=[c, k, len(a)] + list(a), size=(100, 0))
    P = np.column_stack([rng.uniform(-50, 50, 40), rng.uniform(-50, 50, 40), np.full(40, -20.0)])
    R, flag, _, _ = O.intersect(r, rays_from(P, np.tile([0, 0, 1.0], (40, 1))))
    assert np.all(flag == 1)
    r2 = P[:, 0] ** 2 + P[:, 1] ** 2
    assert np.allclose(R[2], _sag(r2, c, k, a), rtol=1e-13, atol=1e-13)
    assert np.array_equal(R[0], P[:, 0]) and np.array_equal(R[1], P[:, 1])


def test_asphere_tilted_rays_converge_to_the_root():
    rng = np.random.default_rng(8)
    c, k, a = 1 / 200.0, -0.3, [0, -2e-9, 1e-13, -1e-17, 1e-21]
    r = rec(cls=S.RT_CONIC, type=S.RT_LENS, params=[c, k, 5] + a, size=(100, 0))
    P = np.column_stack([rng.uniform(-40, 40, 30), rng.uniform(-40, 40, 30), np.full(30, -15.0)])
    Cv = np.column_stack([rng.normal(0, 0.15, 30), rng.normal(0, 0.15, 30), np.ones(30)])
    R0 = rays_from(P, Cv)
    R, flag, _, iters = O.intersect(r, R0)
    assert np.all(flag == 1) and iters > 30
    for i in range(30):
        p, d = R0[:3, i], R0[3:6, i]
        F = lambda t: (p[2] + t * d[2]) - _sag((p[0] + t * d[0]) ** 2 + (p[1] + t * d[1]) ** 2, c, k, a)
        t = optimize.brentq(F, 0.0, 60.0, xtol=1e-14)
        # the reference returns the iterate before the last Newton step (surface.cu:319)
        assert np.linalg.norm(R[:3, i] - (p + t * d)) < 1e-7


def _toric_F(X, Y, Z, c, k, cR, a):
    f = _sag(Y * Y, c, k, a)
    return Z - f - 0.5 * cR * (X * X + Z * Z - f * f)          # surface.cpp:58-60


def test_toric_axis_parallel_closed_form():
    rng = np.random.default_rng(9)
    c, k, cR = 1 / 300.0, 0.0, 1 / 310.0
    r = rec(cls=S.RT_TORIC, type=S.RT_LENS, params=(c, k, 0, cR), size=(100, 0))
    P = np.column_stack([rng.uniform(-40, 40, 40), rng.uniform(-40, 40, 40), np.full(40, -10.0)])
    R, flag, _, _ = O.intersect(r, rays_from(P, np.tile([0, 0, 1.0], (40, 1))))
    assert np.all(flag == 1)
    f = _sag(P[:, 1] ** 2, c, k, [])
    # -0.5 cR Z^2 + Z - (f + 0.5 cR (X^2 - f^2)) = 0, root nearest 0
    q = f + 0.5 * cR * (P[:, 0] ** 2 - f * f)
    Z = (1 - np.sqrt(1 - 2 * cR * q)) / cR
    assert np.allclose(R[2], Z, rtol=0, atol=1e-7)
    assert np.all(np.abs(_toric_F(R[0], R[1], R[2], c, k, cR, [])) < 1e-7)


# ---------------------------------------------------------------- normals, actions

def _num_normal(F, X, Y, Z, h=1e-6):
    g = np.array([(F(X + h, Y, Z) - F(X - h, Y, Z)), (F(X, Y + h, Z) - F(X, Y - h, Z)),
                  (F(X, Y, Z + h) - F(X, Y, Z - h))]) / (2 * h)
    return g / np.linalg.norm(g)


def _surface_points(kind):
    rng = np.random.default_rng(10)
    if kind == "conic":
        c, k, a = 1 / 200.0, -0.3, [0, -2e-9, 1e-13]
        r = rec(cls=S.RT_CONIC, type=S.RT_MIRROR, params=[c, k, 3] + a, size=(100, 0))
        X, Y = rng.uniform(-40, 40, 20), rng.uniform(-40, 40, 20)
        Z = _sag(X * X + Y * Y, c, k, a)
        F = lambda x, y, z: z - _sag(x * x + y * y, c, k, a)
    else:
        c, k, cR, a = 1 / 300.0, -0.2, 1 / 250.0, [0, 4e-9]
        r = rec(cls=S.RT_TORIC, type=S.RT_MIRROR, params=[c, k, 2] + a + [cR], size=(100, 0))
        X, Y = rng.uniform(-30, 30, 20), rng.uniform(-30, 30, 20)
        Z = np.array([optimize.brentq(lambda z: _toric_F(x, y, z, c, k, cR, a), -40, 40, xtol=1e-15)
                      for x, y in zip(X, Y)])
        F = lambda x, y, z: _toric_F(x, y, z, c, k, cR, a)
    return r, X, Y, Z, F


@pytest.mark.parametrize("kind", ["conic", "toric"])
def test_reflection_law_about_the_normal(kind):
    r, X, Y, Z, F = _surface_points(kind)
    rng = np.random.default_rng(11)
    Cv = np.column_stack([rng.normal(0, 0.2, 20), rng.normal(0, 0.2, 20), np.ones(20)])
    R0 = rays_from(np.column_stack([X, Y, Z]), Cv)
    Rr = O.reflect(r, R0)
    for i in range(20):
        n = _num_normal(F, X[i], Y[i], Z[i])
        c = R0[3:6, i]
        assert np.allclose(Rr[3:6, i], c - 2 * np.dot(c, n) * n, atol=1e-8)
        assert abs(np.linalg.norm(Rr[3:6, i]) - 1) < 1e-14


@pytest.mark.parametrize("kind", ["plane", "conic", "toric"])
def test_refraction_snell(kind):
    if kind == "plane":
        r = rec(cls=S.RT_PLANE, type=S.RT_LENS, size=(100, 0))
        X = Y = Z = np.zeros(20)
        n = np.tile([0, 0, 1.0], (20, 1))
    else:
        r, X, Y, Z, F = _surface_points(kind)
        r = r.copy()
        r[S.S_TYPE] = S.RT_LENS
        n = np.array([_num_normal(F, X[i], Y[i], Z[i]) for i in range(20)])
        n = n * np.sign(n[:, 2:3])
    rng = np.random.default_rng(12)
    Cv = np.column_stack([rng.normal(0, 0.3, 20), rng.normal(0, 0.3, 20), np.ones(20)])
    R0 = rays_from(np.column_stack([X, Y, Z]), Cv)
    mu = 1.0 / 1.5
    Rr, flag = O.refract(r, R0, mu)
    assert np.all(flag == 1)
    for i in range(20):
        c, cp = R0[3:6, i], Rr[3:6, i]
        # n1 (c x n) = n2 (c' x n): tangential continuity
        assert np.allclose(np.cross(cp, n[i]), mu * np.cross(c, n[i]), atol=1e-8)
        assert abs(np.linalg.norm(cp) - 1) < 1e-14 and np.dot(cp, n[i]) > 0
    # total internal reflection beyond the critical angle (surface.cu:973)
    crit = math.asin(1.0 / 1.5)
    Rt = rays_from([[0, 0, 0], [0, 0, 0]], [[math.sin(crit + 1e-3), 0, math.cos(crit + 1e-3)],
                                            [math.sin(crit - 1e-3), 0, math.cos(crit - 1e-3)]])
    _, f = O.refract(rec(cls=S.RT_PLANE, type=S.RT_LENS, size=(1, 0)), Rt, 1.5)
    assert list(f) == [0, 1]


def test_plane_reflective_grating_equation_and_littrow():
    r = rec(cls=S.RT_PLANE, type=S.RT_GRATING, size=(100, 100), grating_type=S.RT_REFLECTIVE)
    rng = np.random.default_rng(13)
    Cv = np.column_stack([rng.uniform(-0.5, 0.5, 20), rng.uniform(-0.3, 0.3, 20), np.ones(20)])
    R0 = rays_from(np.zeros((20, 3)), Cv)
    L = 0.3
    Rd, flag = O.diffract(r, R0, 1.0, L)
    assert np.all(flag == 1)
    assert np.allclose(Rd[3], R0[3] - L, atol=1e-15) and np.array_equal(Rd[4], R0[4])
    assert np.all(Rd[5] < 0) and np.allclose(np.sum(Rd[3:6] ** 2, axis=0), 1, atol=1e-15)
    # Littrow: L = 2 sin(alpha) sends the ray back on itself
    a = math.radians(63.4)
    Rl = rays_from([[0, 0, 0]], [[math.sin(a), 0, math.cos(a)]])
    Rd, _ = O.diffract(r, Rl, 1.0, 2 * math.sin(a))
    assert np.allclose(Rd[3:6, 0], -Rl[3:6, 0], atol=1e-15)
    # zero order is the mirror
    Rd, _ = O.diffract(r, R0, 1.0, 0.0)
    assert np.array_equal(Rd[3:6], O.reflect(rec(cls=S.RT_PLANE, type=S.RT_MIRROR, size=(1, 0)), R0)[3:6]) or \
        np.allclose(Rd[3:6], O.reflect(rec(cls=S.RT_PLANE, type=S.RT_MIRROR, size=(1, 0)), R0)[3:6], atol=1e-16)
    # evanescent order (surface.cu:1302)
    _, f = O.diffract(r, rays_from([[0, 0, 0]], [[0, 0, 1.0]]), 1.0, 2.1)
    assert list(f) == [0]


@pytest.mark.parametrize("kind", ["conic", "toric"])
@pytest.mark.parametrize("refl", [True, False])
def test_curved_grating_tangential_jump(kind, refl):
    """c' = mu c - L p + Gamma N: the groove-normal direction p is the unit
    vector perpendicular to N in the plane of N and the X axis, so the
    tangential jump is -L p (pins the u, v, w of surface.cu:1336-1343 and the
    D5 fix of the toric groove vector)."""
    r, X, Y, Z, F = _surface_points(kind)
    r = r.copy()
    r[S.S_TYPE] = S.RT_GRATING
    r[S.S_GRTYPE] = S.RT_REFLECTIVE if refl else S.RT_TRANSPARENT
    rng = np.random.default_rng(14)
    Cv = np.column_stack([rng.normal(0, 0.2, 20), rng.normal(0, 0.2, 20), np.ones(20)])
    R0 = rays_from(np.column_stack([X, Y, Z]), Cv)
    mu, L = (1.0, 0.25) if refl else (0.8, 0.2)
    Rd, flag = O.diffract(r, R0, mu, L)
    assert np.all(flag == 1)
    for i in range(20):
        n = _num_normal(F, X[i], Y[i], Z[i])
        ex = np.array([1.0, 0, 0])
        p = ex - np.dot(ex, n) * n
        p /= np.linalg.norm(p)
        m = mu if not refl else 1.0
        jump = Rd[3:6, i] - m * R0[3:6, i] + L * p
        assert np.linalg.norm(np.cross(jump, n)) < 1e-7
        assert abs(np.linalg.norm(Rd[3:6, i]) - 1) < 1e-14


def test_zero_order_transmission_is_refraction():
    for cls, params in ((S.RT_PLANE, (0, 0, 0)), (S.RT_CONIC, (1 / 200.0, -0.3, 0))):
        g = rec(cls=cls, type=S.RT_GRATING, params=params, size=(100, 0), grating_type=S.RT_TRANSPARENT)
        l = rec(cls=cls, type=S.RT_LENS, params=params, size=(100, 0))
        rng = np.random.default_rng(15)
        P = np.column_stack([rng.uniform(-20, 20, 10), rng.uniform(-20, 20, 10), np.zeros(10)])
        R0 = rays_from(P, np.column_stack([rng.normal(0, .2, 10), rng.normal(0, .2, 10), np.ones(10)]))
        Rg, fg = O.diffract(g, R0, 0.7, 0.0)
        Rl, fl = O.refract(l, R0, 0.7)
        assert np.array_equal(fg, fl)
        assert np.allclose(Rg[3:6], Rl[3:6], rtol=0, atol=2e-16)


# ---------------------------------------------------------------- masks, CCD

def test_mask_edges():
    circ = rec(cls=S.RT_PLANE, type=S.RT_STOP, shape=S.RT_CIRCLE, size=(5.0, 0), center=(1.0, 0.0))
    f = O.mask(circ, [6.0, 6.0 + 1e-12, 1.0, -4.0], [0.0, 0.0, 5.0, 0.0])
    assert list(f) == [1, 0, 1, 1]                     # ray_R2 > R2 is lost (surface.cpp:292)
    rect = rec(cls=S.RT_PLANE, type=S.RT_STOP, shape=S.RT_RECTANGLE, size=(4.0, 2.0), center=(0.0, 0.0))
    f = O.mask(rect, [2.0, -2.0, 2.0 + 1e-12, 0.0, 0.0], [1.0, -1.0, 0.0, 1.0 + 1e-12, 0.0])
    assert list(f) == [1, 1, 0, 0, 1]                  # strict outside test (surface.cpp:302-303)
    shade = rec(cls=S.RT_PLANE, type=S.RT_SHADE, shape=S.RT_CIRCLE, size=(5.0, 0), center=(0.0, 0.0))
    f = O.mask(shade, [5.0, 5.0 + 1e-12, 0.0], [0.0, 0.0, 0.0])
    assert list(f) == [0, 1, 0]                        # shade blocks r <= R (surface.cpp:676)
    shr = rec(cls=S.RT_PLANE, type=S.RT_SHADE, shape=S.RT_RECTANGLE, size=(4.0, 2.0))
    f = O.mask(shr, [2.0, 2.0 + 1e-12], [1.0, 0.0])
    assert list(f) == [0, 1]


def test_ccd_coordinates_brute_force():
    rng = np.random.default_rng(16)
    X = rng.uniform(-8, 8, 1000)
    Y = rng.uniform(-8, 8, 1000)
    cx, cy = O.ccd_coordinates(X, Y, 13.824, 13.824, 0.0135, 0.0135)
    for i in range(1000):
        vx, vy = (X[i] + 13.824 / 2.0) / 0.0135, (Y[i] + 13.824 / 2.0) / 0.0135
        assert cx[i] == (int(vx) if vx >= 0 else 0) and cy[i] == (int(vy) if vy >= 0 else 0)


def test_grating_energy_sums_to_one_over_orders():
    lam = np.array([0.5])
    d, alpha, blaze, gamma = 1000 / 31.6, math.radians(63.4), math.radians(63.4), 0.0
    omin = int(d / 0.5 * (math.sin(alpha) - 1) * math.cos(gamma))
    omax = int(d / 0.5 * (math.sin(alpha) + 1) * math.cos(gamma))
    tot = sum(O.grating_energy(lam, m, blaze, alpha, gamma, d)[0] for m in range(omin, omax + 1))
    assert abs(tot - 1.0) < 1e-12
    e = [O.grating_energy(lam, m, blaze, alpha, gamma, d)[0] for m in range(omin, omax + 1)]
    best = omin + int(np.argmax(e))
    assert abs(best - 2 * d * math.sin(blaze) / 0.5) < 1.0   # peak at the blaze order


def _facet_model_energy(lam, order, blaze, alpha, gamma, d, rule):
    """Scalar blaze efficiency of a sawtooth grating from the facet geometry,
    built from vectors in the dispersion plane (an independent statement of
    the model grating_energy_distr follows, surface.cu:1988-2019,2098-2153):

    * unit vectors from the grating toward the source, s_in = (sin a, cos a),
      and along the diffracted light, s_out = (sin b, cos b), with the
      grating equation sin a + sin b = m lam / (d cos gamma) (Littrow: b = a);
    * the facet tilted by the blaze angle: normal n_f = (sin c, cos c), facet
      direction f = (cos c, -sin c);
    * the illuminated facet length l: the part of the facet whose projection
      across the diffracted beam equals one period's, l = rule d |x . s_perp| /
      |f . s_perp| with s_perp perpendicular to s_out;
    * the phase difference across it, U = pi l f . (s_in + s_out) / lam (half
      the path difference in waves: zero at the facet's specular direction);
    * the obliquity factor (s_out . z)(s_in . n_f) / (s_out . n_f);
    * I = (obliquity sin(U)/U)^2, normalised over the propagating orders
      (|sin b| <= 1)."""
    c = blaze
    s_in = np.array([math.sin(alpha), math.cos(alpha)])
    n_f = np.array([math.sin(c), math.cos(c)])
    f = np.array([math.cos(c), -math.sin(c)])
    z = np.array([0.0, 1.0])
    x = np.array([1.0, 0.0])

    def intensity(m):
        sb = m * lam / (d * math.cos(gamma)) - math.sin(alpha)
        cb = math.sqrt(max(0.0, 1.0 - sb * sb))
        s_out = np.array([sb, cb])
        s_perp = np.array([cb, -sb])
        length = rule * d * abs(x @ s_perp) / abs(f @ s_perp)
        U = math.pi * length * (f @ (s_in + s_out)) / lam
        sinc = math.sin(U) / U if U != 0.0 else 1.0
        obl = (s_out @ z) * (s_in @ n_f) / (s_out @ n_f)
        return (obl * sinc) ** 2

    lo = math.ceil(d * math.cos(gamma) * (math.sin(alpha) - 1.0) / lam - 1e-12)
    hi = math.floor(d * math.cos(gamma) * (math.sin(alpha) + 1.0) / lam + 1e-12)
    norm = sum(intensity(m) for m in range(lo, hi + 1))
    return intensity(order) / norm


@pytest.mark.parametrize("d,blaze_deg,alpha_deg,gamma_deg,rule", [
    (1000 / 31.6, 63.4, 63.4, 0.0, 1.0),        # R2 echelle in Littrow
    (1000 / 31.6, 63.4, 61.0, 0.0, 0.85),       # off Littrow, part of each facet ruled
    (1000 / 600.0, 17.5, 25.0, 0.0, 1.0),       # first-order grating
    (1000 / 300.0, 8.6, 12.0, 5.0, 0.7),        # out-of-plane incidence (gamma)
])
def test_grating_energy_matches_the_facet_model(d, blaze_deg, alpha_deg, gamma_deg, rule):
    """grating_energy_distr against the facet model above for every
    propagating order at several wavelengths.  The reference evaluates sin(U)/U
    by the polynomial 1 - 0.16605 U^2 + 0.00761 U^4 for |U| < pi/2
    (surface.cu:2001-2004), within 1.2e-4 of sin(U)/U there, so the bar is
    1e-3 relative (plus 1e-9 absolute for orders near a zero of the sinc).
    A wrong factor in U (rule fraction, pi, lambda, grating constant), a sine
    and cosine exchanged, or a wrong obliquity term fails it."""
    blaze, alpha, gamma = (math.radians(v) for v in (blaze_deg, alpha_deg, gamma_deg))
    for lam in (0.41, 0.55, 0.73):
        lo = math.ceil(d * math.cos(gamma) * (math.sin(alpha) - 1.0) / lam - 1e-12)
        hi = math.floor(d * math.cos(gamma) * (math.sin(alpha) + 1.0) / lam + 1e-12)
        for m in range(lo, hi + 1):
            got = O.grating_energy(np.array([lam]), m, blaze, alpha, gamma, d, 0.0, rule)[0]
            want = _facet_model_energy(lam, m, blaze, alpha, gamma, d, rule)
            assert abs(got - want) <= 1e-3 * want + 1e-9, (lam, m, got, want)


def test_grating_energy_littrow_blaze_known_answers():
    """In Littrow at the blaze wavelength of order m (lambda_b = 2 d sin c / m)
    the diffracted order m leaves along the facet normal: U = 0 and its
    intensity is the obliquity term alone, cos^2 c; the neighbouring orders
    m +- 1 lie near zeros of the sinc (U ~ -+pi(1 - O(angle change))), so they
    carry little energy; adjacent orders trade the energy between their blaze
    wavelengths."""
    d, c = 1000 / 31.6, math.radians(63.4)
    m = 90
    lam_b = 2 * d * math.sin(c) / m
    e = {k: O.grating_energy(np.array([lam_b]), k, c, c, 0.0, d)[0] for k in (m - 1, m, m + 1)}
    lo = math.ceil(d * (math.sin(c) - 1.0) / lam_b - 1e-12)
    hi = math.floor(d * (math.sin(c) + 1.0) / lam_b + 1e-12)
    # unnormalised order-m intensity cos^2 c: the other orders' share follows from the
    # normalisation, so compare ratios
    assert e[m] > 0.8
    assert e[m - 1] < 0.02 * e[m] and e[m + 1] < 0.02 * e[m]
    want_m = _facet_model_energy(lam_b, m, c, c, 0.0, d, 1.0)
    assert abs(e[m] - want_m) <= 1e-3 * want_m
    # exact U = 0 for order m: the intensity ratio of m to an order far off the
    # blaze, with the obliquity cos^2 c for m, fixes the normalisation
    far = lo + 3
    sb = far * lam_b / d - math.sin(c)
    b = math.asin(sb)
    U = math.pi * d * math.cos(b) / math.cos(c - b) * (math.sin(c - c) + math.sin(c - b)) / lam_b
    sinc = math.sin(U) / U if abs(U) >= math.pi / 2 else (0.00761 * U * U - 0.16605) * U * U + 1
    i_far = (math.cos(b) * math.cos(c - c) / math.cos(c - b) * sinc) ** 2
    e_far = O.grating_energy(np.array([lam_b]), far, c, c, 0.0, d)[0]
    assert abs((e[m] / e_far) - (math.cos(c) ** 2 / i_far)) <= 1e-9 * (e[m] / e_far)
    # adjacent orders trade places between their blaze wavelengths: at the
    # wavelength where their angles straddle the blaze angle symmetrically
    # (sin b_m + sin b_(m+1) = 2 sin c) both carry a large, similar share (the
    # facet's obliquity and illuminated length make the two sides unequal by
    # ~10%), and each dominates at its own blaze wavelength
    lam_x = 4 * d * math.sin(c) / (2 * m + 1)
    em = O.grating_energy(np.array([lam_x]), m, c, c, 0.0, d)[0]
    em1 = O.grating_energy(np.array([lam_x]), m + 1, c, c, 0.0, d)[0]
    assert min(em, em1) > 0.3 and abs(em - em1) <= 0.15 * max(em, em1), (em, em1)
    lam_b1 = 2 * d * math.sin(c) / (m + 1)
    assert O.grating_energy(np.array([lam_b1]), m + 1, c, c, 0.0, d)[0] > \
        50 * O.grating_energy(np.array([lam_b1]), m, c, c, 0.0, d)[0]
    assert hi > lo


def test_flux_weight_is_photons_from_codata():
    """The CCD flux weight (compute_ccd_image, ccd.cu:227,286-287,360):
    photon rate = spectrum x lambda / (h c) x QE x pixel area, with the
    spectrum in erg/s/cm^2/um-like units, lambda in um, the pixel area in cm^2
    (mm^2 x 1e-2).  h and c from CODATA (scipy.constants) give hc in erg um to
    within 1e-6 of the reference's constant; for a monochromatic beam with a
    constant QE and spectrum the frame's sum is N_hits x that weight, and
    each lit pixel holds an integer multiple of it.  Pixels of unequal sides
    catch a squared side, a scaled hc or a dropped 1e-2."""
    import dataclasses

    from scipy import constants as K

    from rt_amd import schemes
    hc = K.h * K.c * 1e7 * 1e6                          # J m -> erg um
    assert abs(hc / 1.98644560233e-12 - 1.0) < 1e-6
    base = schemes.czerny_turner(n_per_lambda=3000, nlambda=1)
    surfs = list(base.surfaces)
    xpix, ypix, qe, spec = 0.0135, 0.02, 0.8, 2.5
    surfs[-1] = dataclasses.replace(surfs[-1], size=(1024 * xpix, 700 * ypix),
                                    ccd=dict(xdim=1024, ydim=700, xpix=xpix, ypix=ypix), qe=qe, spectrum=spec)
    beam = dataclasses.replace(base.beam, range_distr=S.RT_RANGE_MONO, range=(0.54, 0.56), composite=False,
                               nlambda=1)
    sch = S.Scheme(surfs, beam, name="ct-mono")
    res = O.trace(sch, seed=4, hist=True, flux=True)
    lam = (0.56 + 0.54) / 2.0
    w = spec * lam / hc * qe * (xpix * ypix * 1e-2)
    hits = int(res.stats[S.ST_CCD_HITS])
    assert hits > 2000
    assert abs(res.flux.sum() - hits * w) <= 1e-6 * hits * w
    lit = res.hist > 0
    assert np.allclose(res.flux[lit], res.hist[lit] * w, rtol=1e-6, atol=0)
    assert np.all(res.flux[~lit] == 0.0)


# ---------------------------------------------------------------- stage order

def test_fused_trace_is_the_stage_composition():
    """oracle_trace == transform, normalize, intersect, mask, action, normalize per
    surface (scheme.cpp:983-1046) composed from the per-stage functions"""
    from rt_amd import schemes
    sch = schemes.czerny_turner(n_per_lambda=40, nlambda=5)
    res = O.trace(sch, seed=3)
    R = O.gen_rays(sch.beam.pack(), 3, 0, sch.beam.n_rays)
    alive = np.ones(R.shape[1], dtype=bool)
    t = sch.table.reshape(sch.n_surf, S.SURF_STRIDE)
    for s in range(sch.n_surf):
        p = t[s]
        idx = np.nonzero(alive)[0]
        sub = O.transform(p[S.S_ANG:S.S_ANG + 3], p[S.S_DIST:S.S_DIST + 3], R[:, idx])
        sub = O.normalize(sub)
        sub, f, _, _ = O.intersect(p, sub)
        ok = f == 1
        ok &= O.mask(p, sub[0], sub[1]) == 1
        if p[S.S_TYPE] == S.RT_MIRROR:
            sub[:, ok] = O.reflect(p, sub[:, ok])
        elif p[S.S_TYPE] == S.RT_GRATING:
            L = p[S.S_ORDER] * sub[6, ok] / 1.0 / p[S.S_GRCONST]
            d, fd = O.diffract(p, sub[:, ok], 1.0, L)
            sub[:, ok] = d
            ok[np.nonzero(ok)[0][fd == 0]] = False
        sub[:, ok] = O.normalize(sub[:, ok])
        R[:, idx] = sub
        alive[idx[~ok]] = False
    assert np.array_equal(alive.astype(np.int8), res.flag)
    assert np.array_equal(R[:, alive], res.rays[:, alive])
