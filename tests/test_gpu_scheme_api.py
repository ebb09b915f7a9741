"""The Scheme path on the GPU against the oracle (librt_scheme_hip.so).

A scheme written in the reference's XML format is run by Scheme::run on the
engine (rt_scheme_run_file, and the reference's own unchanged
ray-tracing/main.cpp built on rt_amd's Scheme), and its outputs are checked
against the oracle:
- the input beam file: the oracle's beam for the first seed (wavelengths bit
  for bit, positions and cosines within 1e-12 relative);
- every result file: the oracle's saved beam (its two-pointer Rearrange
  replay and renormalised lost rays, pinned byte for byte to the reference's
  Beam::Save by test_saved_beam) for the next seed: header exact, rays in the
  same order (wavelengths bit for bit), states within 1e-12 relative (length
  scale / 1 as floors), except the positions of rays lost to a Newton
  give-up;
- the log: the good-ray count after every stage equals the oracle's; the
  printed extents equal the oracle's to the 6 digits printed;
- a range file with per-range orders gives one result file per range, each
  the oracle's for its range and seed;
- the reference's main.cpp on rt_amd writes the same bytes as
  rt_scheme_run_file; the reference's losses_comp.cpp reads the result
  files and its per-wavelength throughput equals the oracle's count;
- ComputeQE's file equals the product of the surfaces' QE tables and the
  gratings' blaze functions computed by the oracle (1e-10 of the peak:
  asin near +-1 amplifies one-ulp differences between math libraries);
- ccd_frame's FITS image equals the oracle's flux image of the result
  file's good rays (pixel map bit for bit, flux within 1e-12 relative: FP64
  atomics add in any order); its error paths return ccd_frame's codes.
"""
import os
import re
import subprocess

import numpy as np
import pytest

import refpipe as R
from rt_amd import scheme_xml as X
import test_saved_beam as SB
from rt_amd import scheme as S
from rt_amd import schemes
from oracle import oracle as O
from test_scheme_api import SCHEMES, read_fits

pytestmark = pytest.mark.gpu

RTOL = 1e-12
MAIN = os.path.join(R.REF_DIR, "ray_tracing_hip")
LOSSES = os.path.join(R.REF_DIR, "losses_comp_hip")


def with_range(sc, r):
    import dataclasses
    return S.Scheme(list(sc.surfaces), dataclasses.replace(sc.beam, range=tuple(r[:2])), name=sc.name)


def run_scheme(hip, sc, d, **kw):
    path = X.write_scheme(sc, d, **kw)
    lst, log = os.path.join(d, "list.txt"), os.path.join(d, "log.txt")
    rc = hip.scheme_run_file(path, lst, log)
    assert rc == 0, open(log).read()[-2000:] if os.path.exists(log) else rc
    files = [l.strip() for l in open(lst) if l.strip()]
    return path, files, open(log).read()


def check_saved(path, sc, seed, what):
    """a result file against the oracle's saved beam of `seed`"""
    n, ng, nl, rays = R.read_dat(path)
    rec, perm, counts, want = SB.saved_oracle(sc, seed)
    assert (n, ng, nl) == (sc.beam.n_rays, int(counts[0]), max(int(sc.beam.nlambda), 1)), what
    assert np.array_equal(rays[6], want[6]), what + ": ray order or wavelengths differ"
    skip = np.isin(rec.loss[perm.astype(np.int64)], list(SB.newton_giveup_codes(sc)))
    L = SB.length_scale(sc)
    for k in range(6):
        assert SB.rel(rays[k][~skip], want[k][~skip], L if k < 3 else 1.0) <= RTOL, (what, k)
    return rec, counts


def log_counts(log):
    return [int(v) for v in re.findall(r"number of good rays: (\d+)\)", log)]


def oracle_counts(sc, counts):
    good, out = sc.beam.n_rays, []
    for s, surf in enumerate(sc.surfaces):
        if surf.type == S.RT_AUXTYPE:
            continue
        for k in range(3):
            good -= int(counts[1 + 3 * s + k])
            out.append(good)
    return out


def log_extents(log):
    vals = []
    for m in re.finditer(r"Z-cosins range: \[([^,\]]+),([^\]]+)\]|x = \[([^,]+),([^\]]+)\], y = \[([^,]+),([^\]]+)\]",
                         log):
        vals += [float(v) for v in m.groups() if v is not None]
    return np.array(vals)


def oracle_log_extents(sc, ext):
    vals = []
    for s, surf in enumerate(sc.surfaces):
        e = ext[s]
        vals += [e[0], e[1]]
        if surf.type != S.RT_AUXTYPE:
            vals += list(e[2:12])
    return np.array(vals)


@pytest.mark.parametrize("case", sorted(SCHEMES))
def test_scheme_run_matches_oracle(hip, case, tmp_path, monkeypatch):
    monkeypatch.delenv("RT_SEED", raising=False)
    sc = SCHEMES[case]()
    print(f"[scheme run] {case} n={sc.beam.n_rays}", flush=True)
    d = str(tmp_path)
    _, files, log = run_scheme(hip, sc, d)
    assert files == [os.path.join(d, "result.dat")]
    # input beam: seed 0, all rays good
    n, ng, nl, inp = R.read_dat(os.path.join(d, "result_input_beam.dat"))
    assert (n, ng) == (sc.beam.n_rays, sc.beam.n_rays)
    want = O.gen_rays(sc.beam.pack(), 0, 0, n)
    assert np.array_equal(inp[6], want[6])
    assert SB.rel(inp[:3], want[:3], SB.length_scale(sc)) <= RTOL and SB.rel(inp[3:6], want[3:6], 1.0) <= RTOL
    # result: seed 1
    rec, counts = check_saved(files[0], sc, 1, case)
    assert log_counts(log) == oracle_counts(sc, counts)
    got, ref = log_extents(log), oracle_log_extents(sc, rec.extents)
    assert got.shape == ref.shape
    assert np.allclose(got, ref, rtol=1e-5, atol=1e-9, equal_nan=True)
    assert "Finished!" in log and "RAY-TRACING ENGINE INFO" in log


def test_range_file_one_result_per_range(hip, tmp_path, monkeypatch):
    monkeypatch.delenv("RT_SEED", raising=False)
    sc = SCHEMES["czerny_turner"]()
    ranges = [(0.40, 0.55, 1), (0.55, 0.70, 1)]
    d = str(tmp_path)
    _, files, log = run_scheme(hip, sc, d, ranges=ranges)
    assert [os.path.basename(f) for f in files] == ["result_range_(0.4-0.55).dat", "result_range_(0.55-0.7).dat"]
    for i, (f, r) in enumerate(zip(files, ranges)):
        check_saved(f, with_range(sc, r), 1 + i, f)
    assert log.count("*** Current range:") == 2


@pytest.mark.skipif(not os.access(MAIN, os.X_OK), reason="oracle/_ref not built (needs the reference tree)")
def test_reference_main_and_losses_comp_on_rt_amd(hip, tmp_path, monkeypatch):
    monkeypatch.delenv("RT_SEED", raising=False)
    sc = SCHEMES["echelle_fixed_order"]()
    a, b = str(tmp_path / "api"), str(tmp_path / "main")
    monkeypatch.setenv("RT_HIP_VERBOSE", "1")
    print(f"[scheme main] jit {hip.set_jit(True)} table {hip.set_ltab(True)}", flush=True)
    _, files, _ = run_scheme(hip, sc, a)
    os.makedirs(b)
    path = X.write_scheme(sc, b)
    p = subprocess.run([MAIN, path, os.path.join(b, "list.txt"), os.path.join(b, "log.txt")], capture_output=True,
                       text=True, timeout=300)
    print("[scheme main] main's stderr:", p.stderr[-1500:], flush=True)
    assert p.returncode == 0, p.stderr[-2000:]
    assert "Loaded %d surfaces." % sc.n_surf in p.stdout
    for name in ("result_input_beam.dat", "result.dat"):
        ra, rb = R.read_dat(os.path.join(a, name)), R.read_dat(os.path.join(b, name))
        assert ra[:3] == rb[:3], name
        assert np.array_equal(ra[3][6], rb[3][6]), name            # same rays in the same order
        same = open(os.path.join(a, name), "rb").read() == open(os.path.join(b, name), "rb").read()
        diff = np.flatnonzero((ra[3] != rb[3]).any(axis=0))
        print(f"[scheme main] {name}: bit-identical {same}; rays that differ {diff.size}: {diff[:10]}", flush=True)
        L = SB.length_scale(sc)
        for k in range(6):
            assert SB.rel(ra[3][k], rb[3][k], L if k < 3 else 1.0) <= RTOL, (name, k)
    # losses_comp over the result file: per-wavelength fraction of good rays
    out = os.path.join(b, "losses.bin")
    lst = os.path.join(b, "list.txt")
    p = subprocess.run([LOSSES, lst, out], capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stderr[-2000:]
    raw = open(out, "rb").read()
    nl = int(np.frombuffer(raw[:8], dtype=np.uint64)[0])
    ll = np.frombuffer(raw[8:8 + 8 * nl], dtype=np.float64)
    ratio = np.frombuffer(raw[8 + 8 * nl:8 + 16 * nl], dtype=np.float64)
    n, ng, _, rays = R.read_dat(os.path.join(b, "result.dat"))
    allsorted = np.sort(rays[6])
    per = n // nl
    assert np.array_equal(ll, allsorted[::per][:nl])
    good = rays[6][:ng]
    want = np.array([np.count_nonzero(good == v) / per for v in ll])
    assert np.array_equal(ratio, want)


def test_compute_qe_matches_oracle(hip, tmp_path):
    sc = SCHEMES["czerny_turner"]()
    g = [i for i, s in enumerate(sc.surfaces) if s.type == S.RT_GRATING][0]
    qe_tab = S.Spline(np.linspace(0.3, 0.9, 40), 0.6 + 0.3 * np.sin(np.linspace(0, 3, 40)))
    qe = {0: dict(qe_curve=0.93), g: dict(alpha=8.6, gamma=1.5, blaze_angle=8.6, rule_params=(0.05, 0.9),
                                          qe_curve=qe_tab),
          sc.n_surf - 1: dict(qe_curve=qe_tab)}
    for i, s in enumerate(sc.surfaces):
        if s.type in (S.RT_MIRROR, S.RT_LENS, S.RT_DETECTOR) and i not in qe:
            qe[i] = dict(qe_curve=0.97)
    d = str(tmp_path)
    path = X.write_scheme(sc, d, qe=qe)
    out = os.path.join(d, "qe.bin")
    assert hip.scheme_compute_qe(path, 101, out) == 0
    raw = open(out, "rb").read()
    N = int(np.frombuffer(raw[:8], dtype=np.uint64)[0])
    assert N == 101 and len(raw) == 8 + 16 * N
    lam = np.frombuffer(raw[8:8 + 8 * N], dtype=np.float64)
    got = np.frombuffer(raw[8 + 8 * N:], dtype=np.float64)
    r0, r1 = sc.beam.range
    assert np.array_equal(lam, r0 + (r1 - r0) / (N - 1) * np.arange(N))
    want = np.ones(N)
    for i, s in enumerate(sc.surfaces):
        # ComputeQE keeps mirrors, lenses, gratings and detectors (scheme.cpp:1172-1219)
        if i not in qe or s.type not in (S.RT_MIRROR, S.RT_LENS, S.RT_GRATING, S.RT_DETECTOR):
            continue
        q = qe[i]["qe_curve"]
        want = want * (np.array([q(v) for v in lam]) if isinstance(q, S.Spline) else q)
        if s.type == S.RT_GRATING:
            k = S.RT_PI / 180.0
            want = want * O.grating_energy(lam, abs(int(s.order)), 8.6 * k, 8.6 * k, 1.5 * k, s.gr_const, 0.05, 0.9)
    assert np.max(np.abs(got - want)) <= 1e-10 * np.max(np.abs(want))


def test_ccd_frame_matches_oracle(hip, tmp_path, monkeypatch):
    monkeypatch.delenv("RT_SEED", raising=False)
    sc = SCHEMES["mixed"]()
    d = str(tmp_path)
    _, files, _ = run_scheme(hip, sc, d)
    det = [s for s in sc.surfaces if s.type == S.RT_DETECTOR][0]
    c = det.ccd
    ccd_spec = os.path.join(d, "ccd.txt")
    with open(ccd_spec, "w") as f:
        f.write("# CCD\n%r %r\n%d %d\n%r %r\n" % (c["xdim"] * c["xpix"], c["ydim"] * c["ypix"], c["xdim"], c["ydim"],
                                                   c["xpix"], c["ypix"]))
    qx = np.linspace(0.4, 0.9, 57)
    qy = 0.5 + 0.4 * np.cos(qx * 7.0) ** 2
    qe_file = os.path.join(d, "qe.bin")
    with open(qe_file, "wb") as f:
        f.write(np.uint64(len(qx)).tobytes() + qx.tobytes() + qy.tobytes())
    spec = S.Spline(np.linspace(0.4, 0.9, 11), 1.0 + np.linspace(0, 1, 11) ** 2)
    spec_file = os.path.join(d, "spec.txt")
    X.write_table(spec_file, spec)
    fits = os.path.join(d, "frame.fits")
    rc = hip.ccd_frame(os.path.join(d, "list.txt"), qe_file, spec_file, ccd_spec, fits)
    assert rc == 0
    cards, img = read_fits(fits)
    assert img.shape == (c["ydim"], c["xdim"])
    assert cards["QE_FILE"].strip("'").strip() == qe_file[:68]
    # the oracle: pixels by ccd_coordinates, weight spec(l) l / hc QE(l) area
    n, ng, _, rays = R.read_dat(files[0])
    X_, Y_, L_ = rays[0][:ng], rays[1][:ng], rays[6][:ng]
    cx, cy = O.ccd_coordinates(X_, Y_, c["xdim"] * c["xpix"], c["ydim"] * c["ypix"], c["xpix"], c["ypix"])
    on = (cx < c["xdim"]) & (cy < c["ydim"])
    y2 = O.spline_init(qx, qy)
    q = np.array([O.spline_eval(qx, qy, y2, v)[0] for v in L_[on]])
    s = np.array([spec(v) for v in L_[on]]) * L_[on]
    w = s / 1.98644560233e-12 * q * (c["xpix"] * c["ypix"] * 1.0e-2)
    want = np.zeros(c["xdim"] * c["ydim"])
    np.add.at(want, cy[on].astype(np.int64) * c["xdim"] + cx[on], w)
    want = want.reshape(img.shape)
    assert np.array_equal(img != 0, want != 0)
    assert np.max(np.abs(img - want) / np.maximum(np.abs(want), 1e-300)) <= RTOL
    # error paths: a bad CCD specification, a missing list
    bad = os.path.join(d, "bad_ccd.txt")
    open(bad, "w").write("1 2\nthree\n")
    assert hip.ccd_frame(os.path.join(d, "list.txt"), qe_file, spec_file, bad, fits) == 2
    assert hip.ccd_frame(os.path.join(d, "none.txt"), qe_file, spec_file, ccd_spec, fits) == 5
