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
launch; SchemeLoss is the torch.autograd.Function on top.
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
        import torch
        sch = self.spec.apply(self.base, values)
        ds = lib.DeviceScheme(sch)
        n = sch.beam.n_rays
        rays = torch.zeros((7, n), dtype=torch.float64, device="cuda")
        flag = torch.zeros(n, dtype=torch.int8, device="cuda")
        ds.trace(rays, flag, seed=self.seed, flags=S.RT_TRACE_GENERATE | S.RT_TRACE_STORE)
        sp = lib.group_spots(rays, flag, self.n_groups)
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
