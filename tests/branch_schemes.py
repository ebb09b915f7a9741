"""Schemes that drive every optical branch of the engine, for the parity tests.

branches(): a conic beam with random wavelengths (0.35-1.0 um) through
  a circular Shade (central obstruction), a tilted plane mirror, a BK7 block
  (plane lens in, plane lens out at 40 deg: the rays above the critical angle
  undergo total internal reflection), a concave aspheric reflective grating
  in order 2 (the long wavelengths are evanescent), a toric transmissive
  grating, a conic transmissive grating into fused silica, a toric reflective
  grating inside the glass, a plane exit face, a rectangular Shade and a CCD.
  Every surface is placed on the chief ray (0.6 um) traced by the oracle, with
  its axis along that ray, so the beam stays on the optics.

newton_giveups(): one aspheric surface and hand-placed rays: rays that meet
  it normally, a ray tangent to it at its first iterate (|dF| < eps: the
  Newton loop gives up) and a ray whose line has no root on the surface (the
  Newton iterates wander until the 1e5 iteration cap).
"""
import dataclasses

import numpy as np

from rt_amd import scheme as S
from oracle import oracle as O

LAM_CHIEF = 0.6


def _chief_after(surfs, beam):
    """direction of the chief ray (on axis, LAM_CHIEF) after the last surface,
    in that surface's frame, traced by the oracle"""
    # shades would stop the on-axis chief ray: probe with open stops in their place
    probe = [dataclasses.replace(x, type=S.RT_STOP, shape=S.RT_CIRCLE, size=(1e6, 0.0))
             if x.type == S.RT_SHADE else x for x in surfs]
    sch = S.Scheme(probe, beam, name="probe")
    rays = np.array([[0.0], [0.0], [0.0], [0.0], [0.0], [1.0], [LAM_CHIEF]])
    r = O.trace(sch, rays=rays, flag=np.ones(1, dtype=np.int8), dump=True)
    assert r.flag[0] == 1, "the chief ray is lost while placing the surfaces"
    return r.rays[3:6, 0].copy()


def _next(surfs, beam, d, tilt_y=0.0, tilt_x=0.0, rot_z=0.0):
    """distance/angles of a surface d mm along the current chief ray"""
    u = _chief_after(surfs, beam)
    dist, ang = S.aim(u, d, tilt_y, tilt_x)
    return dist, (ang[0], ang[1], rot_z)


def branches(n=20000):
    beam = S.Beam(type=S.RT_BEAM_CONIC, shape=S.RT_CIRCLE, range_distr=S.RT_RANGE_RANDOM,
                  params=(n, 0.0, 4.0, 6.0), center=(0.0, 0.0, 0.0), range=(0.35, 1.0), nlambda=1)
    bk7 = S.glass_spline("N-BK7", 0.3, 1.05)
    fsil = S.glass_spline("F_SILICA", 0.3, 1.05)
    s = []
    s.append(S.Surface(S.RT_PLANE, S.RT_SHADE, S.RT_CIRCLE, size=(0.8, 0.0), distance=(0, 0, 10.0),
                       comment="circular shade (central obstruction)"))
    s.append(S.Surface(S.RT_PLANE, S.RT_MIRROR, S.RT_CIRCLE, size=(30.0, 0.0), distance=(0, 0, 30.0),
                       angles=(20.0, 0.0, 0.0), comment="plane mirror, tilted 20 deg"))
    d, a = _next(s, beam, 40.0)
    s.append(S.Surface(S.RT_PLANE, S.RT_LENS, S.RT_CIRCLE, size=(30.0, 0.0), distance=d, angles=a,
                       n1=1.0, n2=bk7, comment="BK7 block, entrance face (plane lens)"))
    s.append(S.Surface(S.RT_PLANE, S.RT_LENS, S.RT_CIRCLE, size=(40.0, 0.0), distance=(0, 0, 15.0),
                       angles=(40.0, 0.0, 0.0), n1=bk7, n2=1.0,
                       comment="BK7 block, exit face at 40 deg (partial TIR)"))
    d, a = _next(s, beam, 60.0, tilt_y=-8.0)
    s.append(S.Surface(S.RT_CONIC, S.RT_GRATING, S.RT_CIRCLE, params=(-1 / 300.0, -0.4, 2, 1e-6, -2e-10),
                       size=(40.0, 0.0), distance=d, angles=a, grating_type=S.RT_REFLECTIVE, order=2,
                       gr_const=1.4, comment="concave aspheric reflective grating, order 2 (evanescent tail)"))
    d, a = _next(s, beam, 80.0, rot_z=20.0)
    s.append(S.Surface(S.RT_TORIC, S.RT_GRATING, S.RT_CIRCLE, params=(1 / 500.0, 0.0, 0, 1 / 400.0),
                       size=(60.0, 0.0), distance=d, angles=a, grating_type=S.RT_TRANSPARENT, order=1,
                       gr_const=10.0 / 3.0, n1=1.0, n2=1.0, comment="toric transmissive grating"))
    d, a = _next(s, beam, 50.0)
    s.append(S.Surface(S.RT_CONIC, S.RT_GRATING, S.RT_CIRCLE, params=(1 / 250.0, 0.0, 0),
                       size=(70.0, 0.0), distance=d, angles=a, grating_type=S.RT_TRANSPARENT, order=-1,
                       gr_const=5.0, n1=1.0, n2=fsil, comment="conic transmissive grating into fused silica"))
    s.append(S.Surface(S.RT_TORIC, S.RT_GRATING, S.RT_CIRCLE, params=(-1 / 300.0, 0.1, 1, 2e-8, -1 / 280.0),
                       size=(80.0, 0.0), distance=(0, 0, 20.0), angles=(6.0, 2.0, 0.0),
                       grating_type=S.RT_REFLECTIVE, order=1, gr_const=2.5, n1=fsil, n2=fsil,
                       comment="toric reflective grating inside the glass"))
    d, a = _next(s, beam, 15.0)
    s.append(S.Surface(S.RT_PLANE, S.RT_LENS, S.RT_CIRCLE, size=(90.0, 0.0), distance=d, angles=a,
                       n1=fsil, n2=1.0, comment="exit face (plane lens)"))
    d, a = _next(s, beam, 30.0)
    s.append(S.Surface(S.RT_PLANE, S.RT_SHADE, S.RT_RECTANGLE, size=(3.0, 2.0), center=(0.4, -0.2),
                       distance=d, angles=a, comment="rectangular shade"))
    s.append(S.Surface(S.RT_PLANE, S.RT_DETECTOR, S.RT_RECTANGLE, size=(1024 * 0.05, 1024 * 0.05),
                       distance=(0, 0, 40.0), ccd=dict(xdim=1024, ydim=1024, xpix=0.05, ypix=0.05),
                       qe=S.qe_spline(0.3, 1.05), comment="CCD 1024x1024 50um"))
    return S.Scheme(s, beam, name="branches")


def newton_giveups():
    """(scheme, rays, flag): one aspheric mirror at z = 0 (a paraboloid, k = -1,
    with N = 1 so that it takes the Newton intersection; no square root of a
    negative number anywhere) and rays given in its frame."""
    c, k, a1 = 1.0 / 50.0, -1.0, 0.0
    surf = S.Surface(S.RT_CONIC, S.RT_MIRROR, S.RT_CIRCLE, params=(c, k, 1, a1), size=(1e9, 0.0),
                     comment="aspheric mirror")
    beam = S.Beam(type=S.RT_BEAM_PARALLEL, params=(1, 0.0, 1.0), range=(0.6, 0.6), nlambda=1)
    sch = S.Scheme([surf], beam, name="newton-giveups")
    rays = []
    # ordinary rays from z = -50 toward +z
    for x in (-5.0, 0.0, 3.0, 7.5):
        rays.append([x, 1.0, -50.0, 0.01, 0.0, 1.0])
    # tangent at the first iterate: the ray starts on the vertex plane (Z = 0,
    # so s0 = 0) at X0 = 10 with cZ = E X0 cX, E = c for the paraboloid, which
    # makes dF = cZ - E X0 cX vanish to rounding
    x0 = 10.0
    rays.append([x0, 0.0, 0.0, 1.0, 0.0, c * x0])
    # no root: along this ray F(s) = s cZ - c (X0 + s cX)^2 / 2 is a downward
    # parabola whose maximum (about -23 mm) stays below zero, so Newton's
    # iterates never converge; every step is at least ~50 mm (|F/dF| is
    # bounded below for a parabola without real roots), so the loop ends at
    # the 1e5 iteration cap, or where an iterate lands within ~5e-6 mm of the
    # parabola's vertex (|dF| < 1e-7, about 1% of such runs)
    rays.append([-100.0, 0.0, 0.0, -0.97, 0.0, 0.2431])
    R = np.array(rays, dtype=np.float64).T
    n = R.shape[1]
    nrm = np.sqrt(R[3] ** 2 + R[4] ** 2 + R[5] ** 2)
    R[3:6] /= nrm
    R = np.vstack([R, np.full((1, n), 0.6)])
    return sch, np.ascontiguousarray(R), np.ones(n, dtype=np.int8)


def with_beam(sch, **changes):
    return S.Scheme(list(sch.surfaces), dataclasses.replace(sch.beam, **changes), name=sch.name)
