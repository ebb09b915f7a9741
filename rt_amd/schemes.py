"""The frozen workload schemes.

The reference ships no scheme XML, glass or QE tables (its test strategy is
empty), so the two spectrographs its benchmark configurations name are
defined here, in the reference's own scheme vocabulary (surface class, type,
shape, params "c k N a1..aN [cR]", size, center, distance, angles in
degrees; scheme.cpp:406-664), with geometry frozen as literal numbers.

* czerny_turner(): 4-surface Czerny-Turner, FP64.  Collimated 50 mm disk
  beam, 1e5 rays per wavelength x 5 wavelengths over 400-700 nm, seed 0;
  spherical collimator R = -1000 mm, plane reflective grating 600 gr/mm,
  order +1, tilted 10 deg; spherical camera R = -800 mm; flat 1024^2 CCD,
  13.5 um pixels.
* echelle(): 12-surface echelle, FP64.  100 mm beam, Gaussian angular
  spread sigma = 1 arcsec (FWHM 2.35482 arcsec), 4883 rays per wavelength x
  2048 wavelengths over 380-900 nm (10,000,384 rays); feed paraboloid,
  slit, off-axis paraboloid collimator (k = -1); R2 echelle 31.6 gr/mm,
  blaze 63.4 deg, per-ray order nearest the blaze (63..149 over 380-900 nm);
  transmissive cross-disperser 300 gr/mm; camera of six refractive surfaces
  (three singlets of CaF2, N-BK7 and fused silica from 3-term Sellmeier
  64-knot tables) with even-asphere terms to r^10 and one toric surface;
  4096^2 CCD with 15 um pixels and a 200-knot QE spline.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

from .scheme import (Beam, Scheme, Surface, aim, glass_spline, qe_spline,
                     RT_PLANE, RT_CONIC, RT_TORIC, RT_MIRROR, RT_LENS, RT_GRATING,
                     RT_STOP, RT_DETECTOR, RT_CIRCLE, RT_RECTANGLE, RT_REFLECTIVE,
                     RT_TRANSPARENT, RT_ORDER_BLAZE, RT_BEAM_PARALLEL, RT_BEAM_GAUSS,
                     RT_RANGE_UNIFORM)


def _unit(v):
    v = np.asarray(v, dtype=np.float64)
    return v / np.linalg.norm(v)


def _rot(angles_deg, v):
    """Direction `v` of the previous frame in a frame rotated by (Y, X, Z) degrees."""
    ay, ax, az = (math.radians(a) for a in angles_deg)
    x, y, z = v
    x, z = x * math.cos(ay) - z * math.sin(ay), x * math.sin(ay) + z * math.cos(ay)
    y, z = y * math.cos(ax) - z * math.sin(ax), y * math.sin(ax) + z * math.cos(ax)
    x, y = x * math.cos(az) - y * math.sin(az), x * math.sin(az) + y * math.cos(az)
    return np.array([x, y, z])


def _place(point, direction, d, tilt_y=0.0, tilt_x=0.0, rot_z=0.0):
    """distance/angles putting the next vertex d mm from `point` along `direction`."""
    u = _unit(direction)
    dist, ang = aim(u, d, tilt_y, tilt_x)
    dist = tuple(float(p + q) for p, q in zip(point, dist))
    return dist, (ang[0], ang[1], rot_z)


def czerny_turner(n_per_lambda: int = 100000, nlambda: int = 5) -> Scheme:
    beam = Beam(type=RT_BEAM_PARALLEL, shape=RT_CIRCLE, range_distr=RT_RANGE_UNIFORM,
                params=(n_per_lambda, 0.0, 25.0), center=(0.0, 0.0, 0.0),
                range=(0.400, 0.700), nlambda=nlambda, composite=True)
    s = []
    # collimator mirror, 5 deg off-axis
    m1_ang = (5.0, 0.0, 0.0)
    s.append(Surface(RT_CONIC, RT_MIRROR, RT_CIRCLE, params=(-1.0 / 1000.0, 0.0, 0),
                     size=(35.0, 0.0), distance=(0.0, 0.0, 100.0), angles=m1_ang,
                     comment="collimator mirror R=-1000"))
    chief = _rot(m1_ang, (0.0, 0.0, 1.0))
    chief = np.array([chief[0], chief[1], -chief[2]])           # reflection at the vertex
    # plane grating 50 mm before the collimator focus, tilted 10 deg
    g_dist, g_ang = _place((0, 0, 0), chief, 450.0, tilt_y=-10.0)
    d_um = 1000.0 / 600.0
    s.append(Surface(RT_PLANE, RT_GRATING, RT_RECTANGLE, size=(60.0, 60.0),
                     distance=g_dist, angles=g_ang, grating_type=RT_REFLECTIVE,
                     order=1, gr_const=d_um, comment="grating 600 gr/mm, order +1"))
    c = _rot((-10.0, 0.0, 0.0), (0.0, 0.0, 1.0))
    L = 1.0 * 0.55 / d_um
    cx = c[0] - L
    chief = np.array([cx, c[1], -math.sqrt(1.0 - cx * cx - c[1] * c[1])])
    # camera mirror: the grating focus (50 mm) is imaged 1:1 at 2f
    m2_dist, m2_ang = _place((0, 0, 0), chief, 850.0, tilt_y=5.0)
    s.append(Surface(RT_CONIC, RT_MIRROR, RT_CIRCLE, params=(-1.0 / 800.0, 0.0, 0),
                     size=(160.0, 0.0), distance=m2_dist, angles=m2_ang,
                     comment="camera mirror R=-800"))
    c = _rot((5.0, 0.0, 0.0), (0.0, 0.0, 1.0))
    chief = np.array([c[0], c[1], -c[2]])
    ccd_dist, ccd_ang = _place((0, 0, 0), chief, CT_CCD_DISTANCE)
    ccd = dict(xdim=1024, ydim=1024, xpix=0.0135, ypix=0.0135)
    s.append(Surface(RT_PLANE, RT_DETECTOR, RT_RECTANGLE, size=(1024 * 0.0135, 1024 * 0.0135),
                     distance=ccd_dist, angles=ccd_ang, ccd=ccd, comment="CCD 1024x1024 13.5um"))
    return Scheme(s, beam, name="czerny_turner")


CT_CCD_DISTANCE = 800.0


# ---- echelle ----------------------------------------------------------------

ECH_F_FEED = 1000.0        # feed paraboloid focal length (mm)
ECH_H_FEED = 150.0         # its off-axis distance (mm)
ECH_F_COLL = 1000.0        # collimator paraboloid focal length (mm)
ECH_H_COLL = 150.0
ECH_D_ECH = 1000.0 / 31.6  # echelle groove spacing (um)
ECH_BLAZE = 63.4           # deg
ECH_D_XD = 1000.0 / 300.0  # cross-disperser groove spacing (um)
ECH_LAMBDA_C = 0.60        # wavelength that defines the camera axis (um)
ECH_CAM_DIST = 200.0       # cross-disperser -> first camera vertex (mm)
ECH_CCD_DIST = 52.0        # last camera vertex -> CCD (mm), set by best focus
ECH_SLIT = (0.024, 1.0)    # slit width x height (mm)


def echelle(n_per_lambda: int = 4883, nlambda: int = 2048) -> Scheme:
    beam = Beam(type=RT_BEAM_GAUSS, shape=RT_CIRCLE, range_distr=RT_RANGE_UNIFORM,
                params=(n_per_lambda, 0.0, 50.0, 2.35482), center=(0.0, 0.0, 0.0),
                range=(0.380, 0.900), nlambda=nlambda, composite=True)
    caf2 = glass_spline("CAF2")
    bk7 = glass_spline("N-BK7")
    fsil = glass_spline("F_SILICA")
    qe = qe_spline()
    s = []

    # S1 feed paraboloid: collimated beam -> focus, used off-axis by ECH_H_FEED
    c1 = -1.0 / (2.0 * ECH_F_FEED)
    s.append(Surface(RT_CONIC, RT_MIRROR, RT_CIRCLE, params=(c1, -1.0, 0),
                     size=(56.0, 0.0), center=(ECH_H_FEED, 0.0),
                     distance=(-ECH_H_FEED, 0.0, 100.0), comment="feed OAP f=1000"))
    hit = np.array([ECH_H_FEED, 0.0, c1 * ECH_H_FEED ** 2 / 2.0])
    focus = np.array([0.0, 0.0, -ECH_F_FEED])
    chief = focus - hit
    # S2 slit at the focus, normal to the chief ray
    d2, a2 = _place(hit, chief, float(np.linalg.norm(chief)))
    s.append(Surface(RT_PLANE, RT_STOP, RT_RECTANGLE, size=ECH_SLIT,
                     distance=d2, angles=a2, comment="slit"))
    # S3 collimator paraboloid with its focus on the slit; chief ray (+Z in the
    # slit frame) meets it ECH_H_COLL off axis
    f2, h2 = ECH_F_COLL, ECH_H_COLL
    c3 = -1.0 / (2.0 * f2)
    u = _unit([h2, 0.0, f2 - h2 * h2 / (4.0 * f2)])
    a = math.atan2(-u[0], u[2])
    d3 = (f2 * math.sin(a), 0.0, f2 * math.cos(a))
    s.append(Surface(RT_CONIC, RT_MIRROR, RT_CIRCLE, params=(c3, -1.0, 0),
                     size=(60.0, 0.0), center=(h2, 0.0),
                     distance=d3, angles=(math.degrees(a), 0.0, 0.0),
                     comment="collimator OAP f=1000, k=-1"))
    hit3 = np.array([h2, 0.0, c3 * h2 * h2 / 2.0])
    # S4 echelle in Littrow, grooves along Y, incidence = blaze
    d4, a4 = _place(hit3, (0.0, 0.0, -1.0), 600.0, tilt_y=-ECH_BLAZE)
    blaze_k = 2.0 * ECH_D_ECH * math.sin(math.radians(ECH_BLAZE))
    s.append(Surface(RT_PLANE, RT_GRATING, RT_RECTANGLE, size=(250.0, 110.0),
                     distance=d4, angles=a4, grating_type=RT_REFLECTIVE,
                     order=int(round(blaze_k / ECH_LAMBDA_C)), gr_const=ECH_D_ECH,
                     order_mode=RT_ORDER_BLAZE, blaze_k=blaze_k, order_range=(1, 400),
                     comment="R2 echelle 31.6 gr/mm blaze 63.4"))
    c = _rot((-ECH_BLAZE, 0.0, 0.0), (0.0, 0.0, 1.0))
    m = math.floor(blaze_k / ECH_LAMBDA_C + 0.5)
    L = m * ECH_LAMBDA_C / ECH_D_ECH
    cx = c[0] - L
    chief = np.array([cx, c[1], -math.sqrt(1.0 - cx * cx - c[1] * c[1])])
    # S5 cross-disperser, grooves along the echelle dispersion (rotated 90 deg about Z)
    d5, a5 = _place((0, 0, 0), chief, 400.0, rot_z=90.0)
    s.append(Surface(RT_PLANE, RT_GRATING, RT_RECTANGLE, size=(130.0, 130.0),
                     distance=d5, angles=a5, grating_type=RT_TRANSPARENT, order=1,
                     gr_const=ECH_D_XD, n1=1.0, n2=1.0, comment="cross-disperser 300 gr/mm"))
    c = np.array([0.0, 0.0, 1.0])
    L = ECH_LAMBDA_C / ECH_D_XD
    cx = c[0] - L
    chief = np.array([cx, c[1], math.sqrt(1.0 - cx * cx - c[1] * c[1])])
    # camera: three singlets, axis along the cross-dispersed chief ray
    # (paraxial EFL ~280 mm at 0.6 um; L3 is a negative field flattener)
    d6, a6 = _place((0, 0, 0), chief, ECH_CAM_DIST)
    s.append(Surface(RT_CONIC, RT_LENS, RT_CIRCLE,
                     params=(1.0 / 200.0, -0.3, 5, 0.0, -2.0e-9, 1.0e-13, -1.0e-17, 1.0e-21),
                     size=(80.0, 0.0), distance=d6, angles=a6, n1=1.0, n2=caf2,
                     comment="L1 front, CaF2, even asphere to r^10"))
    s.append(Surface(RT_CONIC, RT_LENS, RT_CIRCLE, params=(-1.0 / 1200.0, 0.0, 0),
                     size=(80.0, 0.0), distance=(0.0, 0.0, 32.0), n1=caf2, n2=1.0,
                     comment="L1 back"))
    s.append(Surface(RT_TORIC, RT_LENS, RT_CIRCLE, params=(1.0 / 300.0, 0.0, 0, 1.0 / 310.0),
                     size=(78.0, 0.0), distance=(0.0, 0.0, 5.0), n1=1.0, n2=bk7,
                     comment="L2 front, N-BK7, toric"))
    s.append(Surface(RT_CONIC, RT_LENS, RT_CIRCLE,
                     params=(-1.0 / 1500.0, 0.0, 5, 0.0, 1.5e-9, -5.0e-14, 2.0e-18, -1.0e-22),
                     size=(78.0, 0.0), distance=(0.0, 0.0, 24.0), n1=bk7, n2=1.0,
                     comment="L2 back, even asphere to r^10"))
    s.append(Surface(RT_CONIC, RT_LENS, RT_CIRCLE,
                     params=(-1.0 / 150.0, 0.0, 5, 0.0, 3.0e-9, -2.0e-13, 0.0, 0.0),
                     size=(60.0, 0.0), distance=(0.0, 0.0, 150.0), n1=1.0, n2=fsil,
                     comment="L3 front, fused silica field flattener, even asphere"))
    s.append(Surface(RT_CONIC, RT_LENS, RT_CIRCLE, params=(1.0 / 400.0, 0.0, 0),
                     size=(60.0, 0.0), distance=(0.0, 0.0, 6.0), n1=fsil, n2=1.0,
                     comment="L3 back"))
    ccd = dict(xdim=4096, ydim=4096, xpix=0.015, ypix=0.015)
    s.append(Surface(RT_PLANE, RT_DETECTOR, RT_RECTANGLE, size=(4096 * 0.015, 4096 * 0.015),
                     distance=(0.0, 0.0, ECH_CCD_DIST), ccd=ccd, qe=qe, spectrum=1.0,
                     comment="CCD 4096x4096 15um"))
    return Scheme(s, beam, name="echelle")


def echelle_fixed_order(n_per_lambda: int = 4883, nlambda: int = 2048) -> Scheme:
    """The echelle with its echelle grating in one fixed order, as the
    reference's scheme file states a grating (a list of fixed orders,
    surface.cpp:520-590), the beam inside that order's free spectral range,
    and a detector spectrum table, so the scheme can be written as the
    reference's XML and run by Scheme::run."""
    from .scheme import RT_ORDER_FIXED, Spline
    base = echelle(n_per_lambda=n_per_lambda, nlambda=nlambda)
    surfs = list(base.surfaces)
    g = surfs[3]
    m = int(round(g.blaze_k / 0.6))
    surfs[3] = dataclasses.replace(g, order_mode=RT_ORDER_FIXED, order=m)
    surfs[-1] = dataclasses.replace(surfs[-1], spectrum=Spline(np.linspace(0.55, 0.65, 11),
                                                                1.0 + 0.5 * np.linspace(0, 1, 11) ** 2))
    beam = dataclasses.replace(base.beam, range=(0.597, 0.603))
    return Scheme(surfs, beam, name="echelle-fixed-order")


def tolerance_instances(base: Scheme, n_inst: int, seed: int = 0, decenter_mm: float = 0.020,
                        tilt_arcsec: float = 30.0, curvature_rel: float = 1e-4,
                        min_rays: int = 10_000, max_rays: int = 1_000_000, first: int = None):
    """Monte-Carlo tolerance instances of `base`: every surface from `first`
    on decentred by N(0, decenter_mm) in x and y, tilted by N(0, tilt_arcsec)
    about y and x, curvatures (c and, for torics, cR) scaled by
    1 + N(0, curvature_rel); ray counts log-uniform in [min_rays, max_rays].
    `first` defaults to the surface after the last stop (the slit of the
    echelle: the optics that only illuminate the slit are not toleranced).
    Returns (schemes, counts); the instances share the base's tables (spline
    objects), so they pack to one spline pool."""
    rng = np.random.default_rng(seed)
    if first is None:
        stops = [i for i, s in enumerate(base.surfaces) if s.type == RT_STOP]
        first = stops[-1] + 1 if stops else 0
    out, counts = [], []
    for _ in range(n_inst):
        surfs = list(base.surfaces[:first])
        for srf in base.surfaces[first:]:
            dx, dy = rng.normal(0.0, decenter_mm, 2)
            ty, tx = rng.normal(0.0, tilt_arcsec / 3600.0, 2)
            dist = (srf.distance[0] + dx, srf.distance[1] + dy, srf.distance[2])
            ang = (srf.angles[0] + ty, srf.angles[1] + tx, srf.angles[2])
            p = list(srf.params)
            if srf.cls in (RT_CONIC, RT_TORIC) and p[0] != 0.0:
                p[0] *= 1.0 + rng.normal(0.0, curvature_rel)
            if srf.cls == RT_TORIC:
                n = int(p[2])
                p[3 + n] *= 1.0 + rng.normal(0.0, curvature_rel)
            surfs.append(dataclasses.replace(srf, distance=dist, angles=ang, params=tuple(p)))
        out.append(Scheme(surfs, base.beam, name=f"{base.name}-tol"))
        counts.append(int(round(math.exp(rng.uniform(math.log(min_rays), math.log(max_rays))))))
    return out, counts
