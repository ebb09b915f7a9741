"""Differentiable spot loss over a scheme's parameters (adjoint on the GPU).

A loss on the spots of the surviving rays on the last surface, per ray group
(group of ray r = r mod n_groups; for a composite beam of nlambda wavelengths
with n_groups = nlambda that is one group per wavelength):

    loss = sum_g [ rms_g + w * ((cx_g - tx_g)^2 + (cy_g - ty_g)^2) ]

with rms_g the RMS spot radius, (cx_g, cy_g) the centroid and (tx_g, ty_g)
a target (a dispersion line).  Its gradient with respect to the parameters
comes from rt_hip_adjoint: dloss/dX_r = ax_g + b_g X_r with b_g = 1/(n_g rms_g),
ax_g = -b_g cx_g + 2 w (cx_g - tx_g)/n_g (and likewise for Y), which is exact
for the loss above at a fixed set of surviving rays.

Parameters are (surface index, field) pairs of the packed table
(rt_amd.scheme S_C, S_K, S_A + i, S_CR, S_DIST + j, S_ANG + j, S_GRCONST) in
the units of the scheme description: angles in degrees, distances in mm,
grating constants in um; ParamSpec.density(s) is a groove density in lines per
um (1 / grating constant).  SpotLoss.value_and_grad() runs one forward trace
(rays stored in HBM, spots reduced per group on the device) and one adjoint
launch; SchemeLoss is the torch.autograd.Function on top.  SpotFunction
differentiates any host function of the group spots the same way (seeds from
its partial derivatives, spot_seeds), e.g. the linear dispersion
(linear_dispersion: least-squares slope of the centroids against wavelength).
"""
from __future__ import annotations

import dataclasses
import math
from typing import List, Sequence, Tuple

import numpy as np

from . import lib
from . import scheme as S


@dataclasses.dataclass
class Param:
    surface: int
    field: int
    kind: str = "native"        # "native", or "density" (1 / grating constant)


class ParamSpec:
    def __init__(self, params: Sequence[Param]):
        self.params = list(params)

    def __len__(self):
        return len(self.params)

    @staticmethod
    def density(surface: int) -> Param:
        return Param(surface, S.S_GRCONST, "density")

    def values(self, sch: "S.Scheme") -> np.ndarray:
        return np.array([_get(sch.surfaces[p.surface], p) for p in self.params])

    def apply(self, sch: "S.Scheme", values) -> "S.Scheme":
        surfs = list(sch.surfaces)
        for p, v in zip(self.params, values):
            surfs[p.surface] = _set(surfs[p.surface], p, float(v))
        return S.Scheme(surfs, sch.beam, name=sch.name)

    def chain(self, sch: "S.Scheme", grad_fields: np.ndarray) -> np.ndarray:
        """gradient with respect to the packed-table fields -> parameter units"""
        out = np.array(grad_fields, dtype=np.float64)
        for j, p in enumerate(self.params):
            if S.S_ANG <= p.field < S.S_ANG + 3:
                out[j] *= S.DEG2RAD                       # record angle = deg * pi/180
            elif p.kind == "density":
                d = sch.surfaces[p.surface].gr_const
                out[j] *= -d * d                          # d = 1 / density
        return out


def _npoly(srf) -> int:
    return int(srf.params[2]) if len(srf.params) > 2 else 0


def _get(srf, p: Param) -> float:
    f = p.field
    if f == S.S_C:
        return float(srf.params[0])
    if f == S.S_K:
        return float(srf.params[1])
    if S.S_A <= f < S.S_A + S.RT_MAX_POLY:
        return float(srf.params[3 + f - S.S_A])
    if f == S.S_CR:
        return float(srf.params[3 + _npoly(srf)])
    if S.S_DIST <= f < S.S_DIST + 3:
        return float(srf.distance[f - S.S_DIST])
    if S.S_ANG <= f < S.S_ANG + 3:
        return float(srf.angles[f - S.S_ANG])
    if f == S.S_GRCONST:
        return 1.0 / srf.gr_const if p.kind == "density" else float(srf.gr_const)
    raise ValueError(f"field {f} is not a differentiable parameter")


def _set(srf, p: Param, v: float):
    f = p.field
    if f in (S.S_C, S.S_K) or S.S_A <= f < S.S_A + S.RT_MAX_POLY or f == S.S_CR:
        pr = list(srf.params)
        idx = 0 if f == S.S_C else 1 if f == S.S_K else (3 + _npoly(srf)) if f == S.S_CR else 3 + f - S.S_A
        pr[idx] = v
        return dataclasses.replace(srf, params=tuple(pr))
    if S.S_DIST <= f < S.S_DIST + 3:
        d = list(srf.distance)
        d[f - S.S_DIST] = v
        return dataclasses.replace(srf, distance=tuple(d))
    if S.S_ANG <= f < S.S_ANG + 3:
        a = list(srf.angles)
        a[f - S.S_ANG] = v
        return dataclasses.replace(srf, angles=tuple(a))
    if f == S.S_GRCONST:
        d = 1.0 / v if p.kind == "density" else v
        return dataclasses.replace(srf, gr_const=d)
    raise ValueError(f"field {f} is not a differentiable parameter")


def group_spots(rays, flag, n_groups: int):
    """per group (ray r in group r mod n_groups) of the alive rays: count,
    centroid x, y and RMS radius (CUDA float64 tensors), from the fixed-order
    group reduction of the engine (rt_hip_group_spots)"""
    sp = lib.group_spots(rays, flag, n_groups)
    return sp[:, 0], sp[:, 1], sp[:, 2], sp[:, 5]


class SpotLoss:
    """loss(values) and its gradient for `spec` parameters of `base`."""

    def __init__(self, base: "S.Scheme", spec: ParamSpec, *, n_groups: int, target=None, weight: float = 1.0,
                 seed: int = 0):
        self.base, self.spec, self.n_groups, self.weight, self.seed = base, spec, n_groups, weight, seed
        self.target = None if target is None else np.asarray(target, dtype=np.float64).reshape(n_groups, 2)
        self.ps = np.array([p.surface for p in spec.params], dtype=np.int32)
        self.pf = np.array([p.field for p in spec.params], dtype=np.int32)

    def _forward(self, values):
        sch = self.spec.apply(self.base, values)
        ds = lib.DeviceScheme(sch)
        # the trace forms the group spots itself (rt_hip_trace_spots): no rays stored
        sp = ds.trace_spots(self.n_groups, seed=self.seed)
        return sch, ds, sp

    def spots(self, values):
        _, ds, sp = self._forward(values)
        ds.close()
        return sp[:, 0], sp[:, 1], sp[:, 2], sp[:, 5]

    def value_and_grad(self, values, grad=True):
        import torch
        sch, ds, sp = self._forward(values)
        tgt = None if self.target is None else torch.from_numpy(self.target).cuda().contiguous()
        # loss and adjoint seeds on the device (rt_hip_spot_loss): per group with
        # live rays rms + w |centroid - target|^2; dloss/dX_r = ax + b X_r with
        # b = 1/(n rms), ax = -b cx + 2 w (cx - tx)/n (likewise for y)
        loss_t, seeds = lib.spot_loss(sp, tgt, self.weight)
        loss = float(loss_t.item())
        if not grad:
            ds.close()
            return loss, None
        g = ds.adjoint(seeds, self.ps, self.pf, seed=self.seed)
        torch.cuda.synchronize()
        ds.close()
        return loss, self.spec.chain(sch, g.cpu().numpy())


def spot_seeds(cnt, cx, cy, rms, d_cx, d_cy, d_rms):
    """Adjoint seeds (n_groups x 3 float64: ax, ay, b) of a scalar function of
    the group spots with partial derivatives d_cx, d_cy, d_rms per group:
    with cx = mean X, rms = sqrt(var X + var Y) over a group's n live rays,
    dcx/dX_r = 1/n and drms/dX_r = (X_r - cx)/(n rms), so
    dF/dX_r = ax + b X_r with b = d_rms/(n rms), ax = d_cx/n - b cx (and
    likewise for Y).  Groups without live rays get zero seeds; a group of zero
    radius takes no radius term."""
    cnt, cx, cy, rms = (np.asarray(t, dtype=np.float64) for t in (cnt, cx, cy, rms))
    d_cx, d_cy, d_rms = (np.broadcast_to(np.asarray(t, dtype=np.float64), cnt.shape) for t in (d_cx, d_cy, d_rms))
    live = cnt > 0
    n = np.where(live, cnt, 1.0)
    b = np.where(live & (rms > 0), d_rms / (n * np.where(rms > 0, rms, 1.0)), 0.0)
    ax = np.where(live, d_cx / n - b * cx, 0.0)
    ay = np.where(live, d_cy / n - b * cy, 0.0)
    return np.ascontiguousarray(np.stack([ax, ay, b], axis=1))


class SpotFunction:
    """value and gradient of F(spots) for `spec` parameters of `base`, F any
    scalar function of the per-group spots of the surviving rays, evaluated
    on the host: f(cnt, cx, cy, rms) -> (F, dF/dcx, dF/dcy, dF/drms) (numpy
    arrays over the groups).  One forward trace (rays stored on the device,
    spots reduced per group on the device), the seeds of spot_seeds, one
    adjoint launch."""

    def __init__(self, base: "S.Scheme", spec: ParamSpec, *, n_groups: int, f, seed: int = 0):
        self.loss = SpotLoss(base, spec, n_groups=n_groups, seed=seed)
        self.f = f

    def value_and_grad(self, values, grad=True):
        import torch
        L = self.loss
        sch, ds, sp = L._forward(values)
        h = sp.cpu().numpy()
        cnt, cx, cy, rms = h[:, 0], h[:, 1], h[:, 2], h[:, 5]
        val, d_cx, d_cy, d_rms = self.f(cnt, cx, cy, rms)
        if not grad:
            ds.close()
            return float(val), None
        seeds = torch.from_numpy(spot_seeds(cnt, cx, cy, rms, d_cx, d_cy, d_rms)).cuda()
        g = ds.adjoint(seeds, L.ps, L.pf, seed=L.seed)
        torch.cuda.synchronize()
        ds.close()
        return float(val), L.spec.chain(sch, g.cpu().numpy())


def linear_dispersion(lam, axis: int = 0):
    """F for SpotFunction: the linear dispersion, the least-squares slope of
    the centroid coordinate (axis 0: x, 1: y) against the groups' wavelengths
    `lam` over the groups with live rays (mm per um), and its derivatives:
    D = sum_g (l_g - l_mean)(c_g - c_mean) / sum_g (l_g - l_mean)^2, so
    dD/dc_g = (l_g - l_mean) / sum_k (l_k - l_mean)^2."""
    lam = np.asarray(lam, dtype=np.float64)

    def f(cnt, cx, cy, rms):
        live = cnt > 0
        c = cx if axis == 0 else cy
        lm = lam[live].mean()
        dl = np.where(live, lam - lm, 0.0)
        sll = float(np.sum(dl * dl))
        cm = c[live].mean()
        D = float(np.sum(dl * np.where(live, c - cm, 0.0))) / sll
        dD = dl / sll
        zero = np.zeros_like(c)
        return (D, dD, zero, zero) if axis == 0 else (D, zero, dD, zero)

    return f


def scheme_loss_function():
    """torch.autograd.Function computing SpotLoss on a CPU float64 parameter
    tensor: SchemeLoss.apply(values, spot_loss) -> scalar tensor."""
    import torch

    class SchemeLoss(torch.autograd.Function):
        @staticmethod
        def forward(ctx, values, spot_loss):
            v = values.detach().cpu().numpy().astype(np.float64)
            loss, grad = spot_loss.value_and_grad(v)
            ctx.save_for_backward(torch.from_numpy(grad).to(values.dtype))
            return torch.tensor(loss, dtype=values.dtype)

        @staticmethod
        def backward(ctx, g):
            (grad,) = ctx.saved_tensors
            return g * grad, None

    return SchemeLoss


def echelle_design_params(sch: "S.Scheme") -> ParamSpec:
    """The 48 design parameters of the echelle: curvatures (8), conic
    constants (5), aspheric a2..a5 (r^4..r^10) of the three aspheres (12), the
    toric's cR (1), tilts about Y and X of six surfaces (12), decentres x, y of
    four (8), groove densities of the two gratings (2)."""
    P = []
    for s in (0, 2, 5, 6, 7, 8, 9, 10):
        P.append(Param(s, S.S_C))
    for s in (0, 2, 5, 8, 9):
        P.append(Param(s, S.S_K))
    for s in (5, 8, 9):
        for i in range(2, 6):
            P.append(Param(s, S.S_A + i - 1))
    P.append(Param(7, S.S_CR))
    for s in (3, 4, 5, 7, 9, 11):
        P.append(Param(s, S.S_ANG + 0))
        P.append(Param(s, S.S_ANG + 1))
    for s in (5, 7, 9, 11):
        P.append(Param(s, S.S_DIST + 0))
        P.append(Param(s, S.S_DIST + 1))
    P.append(ParamSpec.density(3))
    P.append(ParamSpec.density(4))
    assert len(P) == 48
    return ParamSpec(P)
