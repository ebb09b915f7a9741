"""Runs a scheme through the reference's own host code.

oracle/_ref/ref_pipeline_{oracle,hip} are built by oracle/ref/build_ref.sh
from the reference's common-lib sources (Beam, Surface and its subclasses,
TabulatedFunction, AsciiFile) and a small driver that replays the
per-surface loop of Scheme::run (common-lib/scheme.cpp:983-1046), linked
against an engine plugin: the oracle's (librt_engine_oracle.so) or the HIP
product (rt_amd/librt_engine_hip.so).  This module writes the driver's text
description of an rt_amd Scheme, runs it, and reads back what it wrote:
Beam::Save's file (beam.cpp:636-715), the CCD image and the efficiency
curves.
"""
from __future__ import annotations

import os
import subprocess

import numpy as np

from rt_amd import scheme as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_DIR = os.path.join(ROOT, "oracle", "_ref")
EXE = {
    "oracle": os.path.join(REF_DIR, "ref_pipeline_oracle"),
    "hip": os.path.join(REF_DIR, "ref_pipeline_hip"),
}


def available(kind: str) -> bool:
    return os.access(EXE[kind], os.X_OK)


def _num(v) -> str:
    return "%.17g" % float(v)


def _fn(v, dirpath: str, tag: str) -> str:
    """'v:<value>' for a constant, 'f:<file>' for a tabulated function
    (an ASCII file of 'x y' lines, read by TabulatedFunction's file
    constructor, tabulated_function.cpp:60-115)."""
    if isinstance(v, S.Spline):
        path = os.path.join(dirpath, tag + ".txt")
        with open(path, "w") as f:
            for x, y in zip(v.x, v.y):
                f.write("%s %s\n" % (_num(x), _num(y)))
        return "f:" + path
    return "v:" + _num(v)


def write_scheme(scheme: S.Scheme, dirpath: str, rules=None) -> str:
    """Writes `scheme` as the driver's text description into dirpath and
    returns its path.  rules: {surface index: (blaze, alpha, gamma)} in
    degrees for Grating::SetRuleParams.  The detector's qe and spectrum
    become compute_ccd_image's QE and incident spectrum."""
    rules = rules or {}
    b = scheme.beam
    lines = []
    lines.append(" ".join(["beam", str(b.type), str(b.shape), str(b.range_distr), str(int(b.composite)),
                           str(int(b.nlambda)), str(len(b.params))]
                          + [_num(p) for p in b.params] + [_num(c) for c in b.center]
                          + [_num(r) for r in b.range]))
    for i, s in enumerate(scheme.surfaces):
        if s.type == S.RT_GRATING and s.order_mode != S.RT_ORDER_FIXED:
            raise ValueError("the reference's Grating takes a fixed diffraction order")
        blaze, alpha, gamma = rules.get(i, (0.0, 0.0, 0.0))
        ccd = s.ccd or {}
        xdim, ydim = int(ccd.get("xdim", 0)), int(ccd.get("ydim", 0))
        xpix, ypix = ccd.get("xpix", 0.0), ccd.get("ypix", 0.0)
        xsize, ysize = ccd.get("xsize", xdim * xpix), ccd.get("ysize", ydim * ypix)
        qe = _fn(s.qe, dirpath, "qe%d" % i)
        sp = _fn(s.spectrum, dirpath, "spec%d" % i)
        fields = ["surface", str(s.cls), str(s.type), str(s.shape), str(len(s.params))]
        fields += [_num(p) for p in s.params]
        fields += [_num(v) for v in s.size] + [_num(v) for v in s.center]
        fields += [_num(v) for v in s.distance] + [_num(v) for v in s.angles]
        fields += [_fn(s.n1, dirpath, "n1_%d" % i), _fn(s.n2, dirpath, "n2_%d" % i)]
        fields += [str(s.grating_type), str(int(s.order)), _num(s.gr_const), _num(blaze), _num(alpha), _num(gamma)]
        fields += [str(xdim), str(ydim), _num(xpix), _num(ypix), _num(xsize), _num(ysize), qe, sp]
        lines.append(" ".join(fields))
    path = os.path.join(dirpath, "scheme.txt")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
    return path


class Result:
    """What one driver run wrote: the beam (7 x n: X Y Z cX cY cZ lambda, good
    rays first), the CCD image (ydim x xdim, or None) and the efficiency
    curves (one row of 64 per grating/detector)."""

    def __init__(self, n, ngood, nlambda, rays, image, qe, stdout):
        self.n, self.ngood, self.nlambda = n, ngood, nlambda
        self.rays, self.image, self.qe, self.stdout = rays, image, qe, stdout

    @property
    def good(self) -> np.ndarray:
        return self.rays[:, :self.ngood]


def read_dat(path: str):
    """Beam::Save's layout: Nrays, N_good, N_lambda (size_t), then X Y Z cX
    cY cZ lambda of the good rays, then the same of the rest."""
    with open(path, "rb") as f:
        hdr = np.frombuffer(f.read(24), dtype=np.uint64)
        n, ngood, nl = int(hdr[0]), int(hdr[1]), int(hdr[2])
        body = np.frombuffer(f.read(), dtype=np.float64)
    rays = np.empty((7, n))
    rays[:, :ngood] = body[:7 * ngood].reshape(7, ngood)
    rays[:, ngood:] = body[7 * ngood:7 * n].reshape(7, n - ngood)
    return n, ngood, nl, rays


def run(kind: str, scheme: S.Scheme, dirpath: str, seed: int = 0, rules=None,
        timeout: int = 120, device: int = 0) -> Result:
    """Runs the driver linked against the `kind` plugin ('oracle' or 'hip')
    in a child process.  The plugin's beam generators draw the rays of the
    fused trace for `seed` (RT_ENGINE_SEED)."""
    os.makedirs(dirpath, exist_ok=True)
    path = write_scheme(scheme, dirpath, rules=rules)
    prefix = os.path.join(dirpath, "out")
    env = dict(os.environ, RT_ENGINE_SEED=str(seed), RT_ENGINE_DEVICE=str(device))
    p = subprocess.run([EXE[kind], path, prefix], env=env, capture_output=True, text=True, timeout=timeout)
    if p.returncode != 0:
        raise RuntimeError("%s driver failed (%d): %s %s" % (kind, p.returncode, p.stdout[-2000:], p.stderr[-2000:]))
    n, ngood, nl, rays = read_dat(prefix + ".dat")
    image = None
    det = scheme.surfaces[-1]
    if os.path.exists(prefix + ".ccd"):
        image = np.fromfile(prefix + ".ccd", dtype=np.float64).reshape(int(det.ccd["ydim"]), int(det.ccd["xdim"]))
    qe = np.fromfile(prefix + ".qe", dtype=np.float64).reshape(-1, 64) if os.path.getsize(prefix + ".qe") else None
    return Result(n, ngood, nl, rays, image, qe, p.stdout)


def canonical(rays: np.ndarray) -> np.ndarray:
    """Rays sorted by (lambda, X, Y): Beam::Rearrange (beam.cpp:574-599)
    swaps bad rays to the end, so the order of the good ones depends on
    where each ray was lost; the sets are compared in this order."""
    order = np.lexsort((rays[1], rays[0], rays[6]))
    return rays[:, order]
