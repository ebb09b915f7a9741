"""The reference's own host code drives the engine plugin libraries.

oracle/_ref/ref_pipeline_* link the reference's common-lib (Beam, Surface,
Mirror, Lens, Grating, Stop, Shade, TabulatedFunction, AsciiFile; compiled
unchanged from its sources by oracle/ref/build_ref.sh) with a driver that
replays Scheme::run's per-surface loop (common-lib/scheme.cpp:983-1046),
saves the beam (Beam::Save), builds a CCD frame the way ccd/ccd_frame.cpp:
184-195 does and calls Surface/Grating::ApplyQE (surface.cpp:331-340,
614-636).  Against the oracle's engine plugin it pins the oracle's fused
trace to the reference's own composition of the stages (host-side
Beam::Normalize, Surface::ApplyConstrains, Shade::ApplyConstrains and
Beam::Rearrange included); against the HIP plugin (librt_engine_hip.so) it
is the drop-in check: the product under the reference's unchanged host code
gives the oracle's result.

The binaries exist only where the reference tree was present when the test
infrastructure was built; the tests skip without them.
"""
import dataclasses

import numpy as np
import pytest

import refpipe as R
from rt_amd import scheme as S
from rt_amd import schemes

RTOL = 1e-12          # positions, cosines: relative to max(1, |value|)


def echelle_fixed_order(n_per_lambda=1000, nlambda=16):
    """The echelle in one fixed order (schemes.echelle_fixed_order) and the
    efficiency rules of its gratings."""
    rules = {3: (63.4, 63.4, 0.0), 4: (12.0, 0.0, 0.0)}
    return schemes.echelle_fixed_order(n_per_lambda=n_per_lambda, nlambda=nlambda), rules


def mixed(n=20000):
    """Conic rectangular beam with random wavelengths through a central
    obstruction (Shade), an aspheric N-F2 doublet pair, a tilted toric mirror,
    an off-centre rectangular stop and a CCD: every ray-loss path of the
    host code (intersection, circular and rectangular masks, Shade)."""
    beam = S.Beam(type=S.RT_BEAM_CONIC, shape=S.RT_RECTANGLE, range_distr=S.RT_RANGE_RANDOM,
                  params=(n, 0.0, 20.0, 0.5, 3.0), center=(0.5, -0.3, 0.0), range=(0.45, 0.85), nlambda=1)
    nf2 = S.glass_spline("N-F2", 0.4, 0.9)
    s = [S.Surface(S.RT_PLANE, S.RT_SHADE, S.RT_CIRCLE, size=(4.0, 0.0), distance=(0, 0, 20.0)),
         S.Surface(S.RT_CONIC, S.RT_LENS, S.RT_CIRCLE, params=(1 / 60., 0.0, 0), size=(22.0, 0.0),
                   distance=(0, 0, 10.0), n1=1.0, n2=nf2),
         S.Surface(S.RT_CONIC, S.RT_LENS, S.RT_CIRCLE, params=(-1 / 30., -0.5, 2, 0.0, 1e-6), size=(22.0, 0.0),
                   distance=(0, 0, 8.0), n1=nf2, n2=1.0),
         S.Surface(S.RT_TORIC, S.RT_MIRROR, S.RT_CIRCLE, params=(-1 / 400., 0.2, 2, 1e-7, -2e-11, -1 / 450.),
                   size=(30.0, 0.0), distance=(0, 0, 100.0), angles=(4.0, 1.0, 0.0)),
         S.Surface(S.RT_PLANE, S.RT_STOP, S.RT_RECTANGLE, size=(50.0, 46.0), center=(0.3, 0.0),
                   distance=(0, 0, -40.0)),
         S.Surface(S.RT_PLANE, S.RT_DETECTOR, S.RT_RECTANGLE, size=(2048 * 0.03, 2048 * 0.03),
                   distance=(0, 0, -30.0), ccd=dict(xdim=2048, ydim=2048, xpix=0.03, ypix=0.03),
                   qe=S.qe_spline())]
    return S.Scheme(s, beam, name="mixed"), {}


def czerny_turner():
    sc = schemes.czerny_turner(n_per_lambda=2000, nlambda=5)
    return sc, {1: (8.6, 10.0, 0.0)}


CASES = {"czerny_turner": (czerny_turner, 7), "echelle_fixed_order": (echelle_fixed_order, 11),
         "mixed": (mixed, 5)}


def rel_err(a, b):
    return float(np.max(np.abs(a - b) / np.maximum(1.0, np.abs(b)))) if a.size else 0.0


def assert_same_beam(got, want, what):
    """Same good-ray set: count, wavelengths exactly, positions and cosines
    within RTOL, compared in (lambda, X, Y) order."""
    assert got.shape == want.shape, (what, got.shape, want.shape)
    a, b = R.canonical(got), R.canonical(want)
    assert np.array_equal(a[6], b[6]), what
    err = rel_err(a[:6], b[:6])
    assert err <= RTOL, (what, err)


def expected_qe(sc, rules, lam):
    """ApplyQE rows: surface_QE (spec *= QE[lambda]) then, for gratings,
    grating_energy_distr with |order| and the rule angles converted by the
    deg2rad macro as SetRuleParams does (surface.cpp:595-636)."""
    from oracle import oracle as O
    rows = []
    for i, s in enumerate(sc.surfaces):
        if s.type != S.RT_GRATING and not (s.type == S.RT_DETECTOR and isinstance(s.qe, S.Spline)):
            continue
        q = np.array([s.qe(v) for v in lam]) if isinstance(s.qe, S.Spline) else np.full(len(lam), float(s.qe))
        if s.type == S.RT_GRATING:
            blaze, alpha, gamma = (v * S.RT_PI / 180.0 for v in rules.get(i, (0.0, 0.0, 0.0)))
            q = q * O.grating_energy(lam, abs(int(s.order)), blaze, alpha, gamma, s.gr_const, 0.0, 1.0)
        rows.append(q)
    return np.array(rows)


@pytest.mark.skipif(not R.available("oracle"), reason="oracle/_ref not built (needs the reference tree)")
@pytest.mark.parametrize("case", sorted(CASES))
def test_reference_host_loop_equals_fused_oracle(case, tmp_path):
    from oracle import oracle as O
    make, seed = CASES[case]
    sc, rules = make()
    res = R.run("oracle", sc, str(tmp_path), seed=seed, rules=rules)
    tr = O.trace(sc, seed=seed, flux=True)
    good = tr.flag == 1
    assert res.n == sc.beam.n_rays
    assert res.ngood == int(good.sum()) > 0
    assert res.ngood < res.n or case == "czerny_turner"
    assert_same_beam(res.good, tr.rays[:, good], case)
    # rays the host code dropped carry flag 0; the beam's lost rays are the fused trace's lost rays
    lost_ref = R.canonical(res.rays[:, res.ngood:])
    assert np.array_equal(np.sort(lost_ref[6]), np.sort(tr.rays[6, ~good]))
    # the CCD frame (ccd_coordinates + compute_ccd_image on the saved beam)
    img = res.image
    flux = tr.flux.reshape(img.shape)
    assert np.array_equal(img > 0, flux > 0)
    assert np.all(np.abs(img - flux) <= RTOL * np.abs(flux))
    # efficiency rows
    lam = sc.beam.range[0] + (sc.beam.range[1] - sc.beam.range[0]) * np.arange(64) / 63.0
    want = expected_qe(sc, rules, lam)
    if len(want):
        assert res.qe.shape == want.shape
        nan = np.isnan(want)
        assert np.array_equal(np.isnan(res.qe), nan)
        assert np.all(np.abs(res.qe[~nan] - want[~nan]) <= 1e-13 * np.maximum(1.0, np.abs(want[~nan])))


@pytest.mark.gpu
@pytest.mark.skipif(not (R.available("oracle") and R.available("hip")),
                    reason="oracle/_ref not built (needs the reference tree)")
@pytest.mark.parametrize("case", sorted(CASES))
def test_reference_host_loop_on_hip_plugin(case, tmp_path):
    """Drop-in: the reference's host code on librt_engine_hip.so gives what it
    gives on the oracle plugin."""
    make, seed = CASES[case]
    sc, rules = make()
    print("ref pipeline", case, "seed", seed, "rays", sc.beam.n_rays, flush=True)
    want = R.run("oracle", sc, str(tmp_path / "oracle"), seed=seed, rules=rules)
    got = R.run("hip", sc, str(tmp_path / "hip"), seed=seed, rules=rules)
    assert "Architecture: gfx950" in got.stdout, got.stdout
    assert got.n == want.n and got.nlambda == want.nlambda
    assert got.ngood == want.ngood
    assert_same_beam(got.good, want.good, case)
    flux = want.image
    assert np.array_equal(got.image > 0, flux > 0)
    assert np.all(np.abs(got.image - flux) <= RTOL * np.abs(flux))
    if want.qe is not None:
        nan = np.isnan(want.qe)
        assert np.array_equal(np.isnan(got.qe), nan)
        # grating_energy_distr near grazing orders amplifies one-ulp differences
        # of sin/asin between math libraries: 1e-10 of each row's peak
        for g, w, m in zip(got.qe, want.qe, nan):
            ok = ~m
            if ok.any():
                assert float(np.max(np.abs(g[ok] - w[ok]))) <= 1e-10 * float(np.max(np.abs(w[ok])))
