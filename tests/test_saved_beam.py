"""Scheme::run's saved beam: the Rearrange order, the lost rays' state and the
logged extents.

CPU (pins of the oracle):
- the oracle's saved beam (oracle_trace_record -> oracle_rearrange, the
  reference's two-pointer loop restated -> oracle_beam_order with the lost
  rays renormalised as Beam::Normalize leaves them) is byte for byte the file
  the reference's own host code writes with Beam::Save (oracle/_ref driver:
  common-lib beam.cpp, surface.cpp compiled unchanged, Scheme::run's stage
  loop, on the oracle's engine plugin), on schemes that lose rays at every
  kind of stage; a stable partition in place of the two-pointer loop, or no
  renormalisation of the lost rays, changes those bytes (planted mistakes);
- the logged extents equal those of a stage-by-stage composition of the
  oracle's per-stage functions (min / max of cZ after the first normalise,
  X / Y after the intersection and after the mask, cZ after the action).

GPU (parity of the product with the oracle): rt_hip_trace_record's loss
codes equal the oracle's bit for bit, its extents within 1e-12 relative (the
extremes of rays whose states agree to 1e-12); rt_hip_rearrange's order
equals oracle_rearrange's exactly (same codes in, integer permutation out);
rt_hip_beam_order's arrays equal oracle_beam_order's in file order with
wavelengths bit-exact and positions / cosines within 1e-12 relative (floor:
the scheme's length scale for positions, 1 for cosines), except the
positions of rays lost to a Newton give-up, which stop at a diverged iterate
(their order and wavelength still match); at the echelle's full size
(10,000,384 rays) the order is checked against the oracle's and the saved
good rays against the oracle's state.
"""
import os

import numpy as np
import pytest

import refpipe as R
from rt_amd import scheme as S
from rt_amd import schemes
from oracle import oracle as O

import branch_schemes as B
from test_ref_pipeline import CASES as REF_CASES

RTOL = 1e-12


def cases():
    c = dict(REF_CASES)
    c["branches"] = (lambda: (B.branches(6000), {}), 13)
    return c


CASES = cases()


def saved_oracle(sc, seed):
    rec = O.trace_record(sc, seed=seed)
    perm, counts = O.rearrange(rec.loss)
    ordered = O.beam_order(perm, rec.loss, O.renorm_table(sc), rec.rays)
    return rec, perm, counts, ordered


@pytest.mark.skipif(not R.available("oracle"), reason="oracle/_ref not built (needs the reference tree)")
@pytest.mark.parametrize("case", sorted(CASES))
def test_oracle_saved_beam_is_the_reference_file(case, tmp_path):
    make, seed = CASES[case]
    sc, rules = make()
    R.run("oracle", sc, str(tmp_path), seed=seed, rules=rules)
    ref = open(os.path.join(str(tmp_path), "out.dat"), "rb").read()
    rec, perm, counts, ordered = saved_oracle(sc, seed)
    n, nl = sc.beam.n_rays, max(int(sc.beam.nlambda), 1)
    mine = O.dat_bytes(n, int(counts[0]), nl, ordered)
    assert len(mine) == len(ref)
    assert mine == ref, "the oracle's saved beam differs from Beam::Save's file"
    lost = n - int(counts[0])
    if lost:
        # planted mistakes: a stable partition, or lost rays left unnormalised
        good = np.flatnonzero(rec.loss == 0)
        bad = np.flatnonzero(rec.loss != 0)
        stable = np.concatenate([good, bad]).astype(np.uint64)
        assert not np.array_equal(stable, perm)
        assert O.dat_bytes(n, int(counts[0]), nl, O.beam_order(stable, rec.loss, O.renorm_table(sc), rec.rays)) != ref
        assert O.dat_bytes(n, int(counts[0]), nl, O.beam_order(perm, rec.loss, np.zeros(256, np.uint8),
                                                               rec.rays)) != ref


def staged_extents(sc, seed):
    """the 12 logged extents per surface from the oracle's per-stage
    functions, composed as Scheme::run composes its stages"""
    n = sc.beam.n_rays
    rays = O.gen_rays(sc.beam.pack(), seed, 0, n)
    alive = np.ones(n, dtype=bool)
    t = np.asarray(sc.table).reshape(sc.n_surf, S.SURF_STRIDE)
    ext = np.full((sc.n_surf, O.EXT_PER_SURF), np.nan)

    def mm(row, q, v):
        if v.size:
            row[2 * q], row[2 * q + 1] = v.min(), v.max()

    def fn(f, lam):
        return np.array([f(v) for v in lam]) if isinstance(f, S.Spline) else np.full(lam.shape, float(f))

    for s, surf in enumerate(sc.surfaces):
        rec = np.ascontiguousarray(t[s])
        idx = np.flatnonzero(alive)
        sub = O.transform(rec[S.S_ANG:S.S_ANG + 3], rec[S.S_DIST:S.S_DIST + 3], rays[:, idx])
        sub = O.normalize(sub)
        mm(ext[s], 0, sub[5])
        if surf.type == S.RT_AUXTYPE:
            rays[:, idx] = sub
            continue
        sub, f, _, _ = O.intersect(rec, sub)
        ok = f == 1
        mm(ext[s], 1, sub[0][ok])
        mm(ext[s], 2, sub[1][ok])
        m = ok & (O.mask(rec, sub[0], sub[1]) == 1)
        mm(ext[s], 3, sub[0][m])
        mm(ext[s], 4, sub[1][m])
        a = m.copy()
        w = np.flatnonzero(m)
        part = sub[:, w]
        if surf.type == S.RT_MIRROR:
            part = O.reflect(rec, part)
        elif surf.type in (S.RT_LENS, S.RT_GRATING):
            n2 = fn(surf.n2, part[6])
            mu = fn(surf.n1, part[6]) / n2
            if surf.type == S.RT_LENS:
                part, fa = O.refract(rec, part, mu)
            else:
                part, fa = O.diffract(rec, part, mu, surf.order * part[6] / n2 / surf.gr_const)
            a[w] = fa == 1
        sub[:, w] = part
        w = np.flatnonzero(a)
        sub[:, w] = O.normalize(sub[:, w])
        mm(ext[s], 5, sub[5][a])
        rays[:, idx] = sub
        keep = np.zeros(n, dtype=bool)
        keep[idx[a]] = True
        alive = keep
    return ext


def test_oracle_extents_equal_staged_composition():
    sc, _ = REF_CASES["mixed"][0](n=3000)
    rec = O.trace_record(sc, seed=5)
    want = staged_extents(sc, 5)
    assert np.array_equal(np.isnan(rec.extents), np.isnan(want))
    ok = ~np.isnan(want)
    assert np.array_equal(rec.extents[ok], want[ok])


def test_renorm_table_counts_later_normalisations():
    sc, _ = REF_CASES["mixed"][0](n=100)
    ren = O.renorm_table(sc)
    ns = sc.n_surf
    # last surface: its own second normalisation only; each earlier surface adds two
    for s in range(ns):
        for k in range(3):
            assert ren[1 + 3 * s + k] == 1 + 2 * (ns - 1 - s)
    assert ren[0] == 0 and ren[255] == 0


# ---------------------------------------------------------------- GPU parity

def length_scale(sc):
    t = np.asarray(sc.table).reshape(sc.n_surf, S.SURF_STRIDE)
    return max(1.0, float(np.max(np.abs(t[:, S.S_DIST:S.S_DIST + 3]))))


def rel(a, b, floor):
    return float(np.max(np.abs(a - b) / np.maximum(np.maximum(np.abs(a), np.abs(b)), floor))) if a.size else 0.0


def gpu_saved(hip, sc, seed):
    import torch
    n = sc.beam.n_rays
    ds = hip.DeviceScheme(sc)
    rays = torch.zeros((7, n), dtype=torch.float64, device="cuda")
    flag = torch.zeros(n, dtype=torch.int8, device="cuda")
    loss = torch.zeros(n, dtype=torch.uint8, device="cuda")
    ext = torch.zeros((sc.n_surf, hip.EXT_PER_SURF), dtype=torch.float64, device="cuda")
    ds.trace_record(rays, flag, loss, seed=seed, extents=ext)
    perm, counts = hip.rearrange(loss)
    out = hip.beam_order(perm, loss, O.renorm_table(sc), rays)
    torch.cuda.synchronize()
    res = (rays.cpu().numpy(), flag.cpu().numpy(), loss.cpu().numpy(), ext.cpu().numpy(),
           perm.cpu().numpy().astype(np.uint64), counts, out.cpu().numpy())
    ds.close()
    return res


def newton_giveup_codes(sc):
    codes = set()
    for s, surf in enumerate(sc.surfaces):
        if surf.cls in (S.RT_CONIC, S.RT_TORIC) and (surf.cls == S.RT_TORIC or int(surf.params[2]) > 0):
            codes.add(1 + 3 * s)
    return codes


@pytest.mark.gpu
@pytest.mark.parametrize("case", sorted(CASES))
def test_gpu_saved_beam_matches_oracle(hip, case):
    make, seed = CASES[case]
    sc, _ = make()
    print(f"[saved beam] {case} n={sc.beam.n_rays} seed={seed}", flush=True)
    rays, flag, loss, ext, perm, counts, out = gpu_saved(hip, sc, seed)
    rec, operm, ocounts, oout = saved_oracle(sc, seed)
    assert np.array_equal(loss, rec.loss), "loss codes differ from the oracle"
    assert np.array_equal(flag, rec.flag)
    assert np.array_equal(counts, ocounts)
    assert np.array_equal(perm, operm), "Rearrange order differs from the oracle's two-pointer replay"
    assert np.array_equal(np.isnan(ext), np.isnan(rec.extents))
    okx = ~np.isnan(ext)
    L = length_scale(sc)
    floor = np.array([1.0, 1.0] + [L] * 8 + [1.0, 1.0])[None, :].repeat(sc.n_surf, 0)
    assert rel(ext[okx], rec.extents[okx], floor[okx]) <= RTOL
    # file order: wavelengths exact, positions / cosines within RTOL
    assert np.array_equal(out[6], oout[6])
    skip = np.isin(loss[perm.astype(np.int64)], list(newton_giveup_codes(sc)))
    keep = ~skip
    for k in range(6):
        fl = L if k < 3 else 1.0
        assert rel(out[k][keep], oout[k][keep], fl) <= RTOL, (case, k)


@pytest.mark.gpu
def test_gpu_saved_beam_full_echelle(hip):
    """The echelle workload at full size (4883 rays x 2048 wavelengths, its
    echelle in the per-ray blaze order): order equal to the oracle's, the
    saved good rays equal to the oracle's state."""
    import torch
    sc = schemes.echelle()
    seed = 21
    n = sc.beam.n_rays
    print(f"[saved beam] full echelle n={n}", flush=True)
    rays, flag, loss, ext, perm, counts, out = gpu_saved(hip, sc, seed)
    rec = O.trace_record(sc, seed=seed, nthreads=16)
    assert np.array_equal(loss, rec.loss)
    operm, ocounts = O.rearrange(rec.loss)
    assert np.array_equal(counts, ocounts)
    assert np.array_equal(perm, operm)
    ng = int(counts[0])
    assert ng > 0.5 * n
    g = perm[:ng].astype(np.int64)
    assert np.array_equal(out[6][:ng], rec.rays[6][g])
    L = length_scale(sc)
    for k in range(6):
        assert rel(out[k][:ng], rec.rays[k][g], L if k < 3 else 1.0) <= RTOL
