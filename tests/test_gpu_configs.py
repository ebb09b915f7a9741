"""Every benchmark configuration at its stated size on the GPU, against the oracle.

C1  Czerny-Turner, 1e5 rays per wavelength x 5 wavelengths (5e5 rays), seed 0:
    every ray against the oracle.
C2  the 12-surface echelle at 4883 x 2048 = 10,000,384 rays: every ray's flag
    and state, the whole 4096^2 hit-count frame, the flux frame and the
    counters against the oracle; rays within 1e-9 px of a pixel edge are
    counted (RT_ST_CCD_EDGE) and printed.
C3  1024 tolerance instances, ray counts log-uniform in [1e4, 1e6] (~2.2e8
    rays) in one batched launch: oracle parity on sampled instances (first,
    last, smallest, largest, a partial last tile, three more), conservation of
    the counters and every instance's spot count against its stored flags.
C5  256 tolerance instances, counts log-uniform in [1e4, 5e5], the design-loop
    step (batched trace, spots, loss, batched adjoint, descent on the
    compensators) captured in a hipGraph: 5 replays equal 5 eager steps of an
    identical batch bit for bit, and the loss does not increase.
(C4, the differentiable scheme at 1e6 rays with its finite-difference check,
is in test_gpu_adjoint.py.)

Bars (written here): positions within 1e-12 relative with the scheme's
length scale as floor, cosines 1e-12, wavelengths bit-exact; flags, integer
hit counts and counters bit-exact, except that a ray whose pixel differs
between the two traces must lie within 1e-6 px of a pixel edge (such rays
are counted and reported); flux per pixel within 1e-12 relative (FP64
atomics add in any order).
"""
import math

import numpy as np
import pytest
import torch

from rt_amd import diff
from rt_amd import scheme as S
from rt_amd import schemes
from oracle import oracle as O

pytestmark = pytest.mark.gpu

RTOL = 1e-12
HOST_THREADS = 16


def length_scale(sch):
    t = np.asarray(sch.table).reshape(sch.n_surf, S.SURF_STRIDE)
    return max(1.0, float(np.max(np.abs(t[:, S.S_DIST:S.S_DIST + 3]))))


def state_err(g, r, ok, L):
    """largest relative error over the alive rays, row by row (7, n)"""
    worst = 0.0
    for k in range(7):
        a, b = g[k][ok], r[k][ok]
        if a.size == 0:
            continue
        floor = L if k < 3 else (1.0 if k < 6 else 0.0)
        scale = np.maximum(np.maximum(np.abs(a), np.abs(b)), floor)
        worst = max(worst, float(np.max(np.abs(a - b) / scale)))
    return worst


def pixels(sch, X, Y):
    """pixel index of each ray (ccd_coordinates, ccd.cu:30-45) and its
    distance in pixels from the nearest pixel edge, on the host"""
    det = sch.detector()
    c = det.ccd
    vx = (X + c["xdim"] * c["xpix"] / 2.0) / c["xpix"]
    vy = (Y + c["ydim"] * c["ypix"] / 2.0) / c["ypix"]
    px = np.where(vx >= 0, np.floor(vx), -1).astype(np.int64)
    py = np.where(vy >= 0, np.floor(vy), -1).astype(np.int64)
    on = (px >= 0) & (px < c["xdim"]) & (py >= 0) & (py < c["ydim"])
    fx, fy = vx - np.floor(vx), vy - np.floor(vy)
    edge = np.minimum(np.minimum(fx, 1 - fx), np.minimum(fy, 1 - fy))
    return np.where(on, py * c["xdim"] + px, -1), edge


def check_frames(sch, g_rays, g_ok, r_rays, r_ok, g_hist, r_hist, label):
    """hit frames equal except for rays whose pixel differs between the two
    traces; each of those must lie within 1e-6 px of an edge"""
    pg, eg = pixels(sch, g_rays[0][g_ok], g_rays[1][g_ok])
    pr, er = pixels(sch, r_rays[0][r_ok], r_rays[1][r_ok])
    assert pg.shape == pr.shape
    moved = pg != pr
    n_moved = int(moved.sum())
    if n_moved:
        assert np.all(np.minimum(eg[moved], er[moved]) < 1e-6), "a ray away from a pixel edge changed pixel"
    diff_frame = g_hist.astype(np.int64) - r_hist.astype(np.int64)
    npix = len(g_hist)
    expect = np.zeros(npix, dtype=np.int64)
    if n_moved:
        a, b = pg[moved], pr[moved]
        np.add.at(expect, a[a >= 0], 1)
        np.add.at(expect, b[b >= 0], -1)
    assert np.array_equal(diff_frame, expect), "hit frames differ beyond the rays that changed pixel"
    print(f"[{label}] rays that changed pixel between GPU and oracle: {n_moved}", flush=True)
    return n_moved


def full_trace(hip, sch, seed, flux=True):
    n = sch.beam.n_rays
    det = sch.detector()
    npix = det.ccd["ydim"] * det.ccd["xdim"]
    ds = hip.DeviceScheme(sch)
    R = torch.zeros((7, n), dtype=torch.float64, device="cuda")
    F = torch.zeros(n, dtype=torch.int8, device="cuda")
    H = torch.zeros(npix, dtype=torch.int32, device="cuda")
    X = torch.zeros(npix, dtype=torch.float64, device="cuda") if flux else None
    st = torch.zeros(S.STATS_LEN, dtype=torch.int64, device="cuda")
    flags = S.RT_TRACE_GENERATE | S.RT_TRACE_STORE | S.RT_TRACE_HIST | (S.RT_TRACE_FLUX if flux else 0)
    print(f"[config] GPU trace {sch.name} n={n} seed={seed}", flush=True)
    ds.trace(R, F, seed=seed, flags=flags, hist=H, flux=X, stats=st)
    torch.cuda.synchronize()
    out = dict(rays=R.cpu().numpy(), flag=F.cpu().numpy(), hist=H.cpu().numpy().view(np.uint32),
               stats=st.cpu().numpy().view(np.uint64), flux=X.cpu().numpy() if flux else None)
    ds.close()
    del R, F, H, X
    return out


def compare_full(sch, g, ref, label):
    L = length_scale(sch)
    assert np.array_equal(g["flag"], ref.flag), f"{int(np.sum(g['flag'] != ref.flag))} flags differ"
    ok = ref.flag == 1
    err = state_err(g["rays"], ref.rays, ok, L)
    assert err <= RTOL, err
    assert np.array_equal(g["rays"][6].view(np.uint64), ref.rays[6].view(np.uint64)), "wavelengths differ"
    moved = check_frames(sch, g["rays"], ok, ref.rays, ok, g["hist"], ref.hist, label)
    gs, rs = g["stats"].astype(np.int64), ref.stats.astype(np.int64)
    # the CCD counters may differ only by the rays that changed pixel
    for k in range(S.STATS_LEN):
        if k in (S.ST_CCD_HITS, S.ST_CCD_OFFCHIP, S.ST_CCD_EDGE):
            assert abs(gs[k] - rs[k]) <= moved, (k, gs[k], rs[k])
        else:
            assert gs[k] == rs[k], (k, gs[k], rs[k])
    if g["flux"] is not None:
        nz = (ref.flux != 0) | (g["flux"] != 0)
        if moved == 0:
            rel = np.abs(g["flux"][nz] - ref.flux[nz]) / np.abs(ref.flux[nz])
            assert float(np.max(rel)) < 1e-12
    print(f"[{label}] {sch.beam.n_rays} rays: flags bit-exact, worst state error {err:.2e} (L = {L:.0f} mm), "
          f"{int(ok.sum())} alive, {int(gs[S.ST_CCD_HITS])} CCD hits, "
          f"{int(gs[S.ST_CCD_EDGE])} within 1e-9 px of a pixel edge (RT_ST_CCD_EDGE)", flush=True)


def test_c1_czerny_turner_full_size(hip):
    sch = schemes.czerny_turner()                   # 1e5 x 5 = 5e5 rays
    assert sch.beam.n_rays == 500_000
    g = full_trace(hip, sch, seed=0)
    ref = O.trace(sch, seed=0, hist=True, flux=True, nthreads=HOST_THREADS)
    compare_full(sch, g, ref, "C1")
    assert ref.stats[S.ST_CCD_HITS] > 0.9 * sch.beam.n_rays


def test_c2_echelle_full_size(hip):
    sch = schemes.echelle()                         # 4883 x 2048 = 10,000,384 rays
    assert sch.beam.n_rays == 10_000_384
    g = full_trace(hip, sch, seed=0)
    print("[C2] oracle trace of the whole beam", flush=True)
    ref = O.trace(sch, seed=0, hist=True, flux=True, nthreads=HOST_THREADS)
    compare_full(sch, g, ref, "C2")
    assert int(ref.stats[S.ST_GOOD]) > 0.95 * sch.beam.n_rays


# ---- C3 ---------------------------------------------------------------------

def _spot_np(x, y):
    n = len(x)
    if n == 0:
        return np.zeros(6)
    cx, cy = x.mean(), y.mean()
    vx, vy = np.mean((x - cx) ** 2), np.mean((y - cy) ** 2)
    return np.array([n, cx, cy, vx, vy, math.sqrt(vx + vy)])


def test_c3_batched_forward_full_size(hip):
    base = schemes.echelle()
    inst, counts = schemes.tolerance_instances(base, 1024, seed=1, min_rays=10_000, max_rays=1_000_000)
    counts = np.array(counts, dtype=np.int64)
    total = int(counts.sum())
    print(f"[C3] 1024 instances, {total} rays", flush=True)
    assert total > 1.5e8
    db = hip.DeviceBatch(inst, counts)
    R = torch.zeros((7, total), dtype=torch.float64, device="cuda")
    F = torch.zeros(total, dtype=torch.int8, device="cuda")
    spots = torch.zeros((len(inst), hip.SPOT_STRIDE), dtype=torch.float64, device="cuda")
    st = torch.zeros(S.STATS_LEN, dtype=torch.int64, device="cuda")
    db.trace(seed=100, flags=S.RT_TRACE_GENERATE | S.RT_TRACE_STORE, rays=R, flag=F, spots=spots, stats=st)
    torch.cuda.synchronize()
    sp = spots.cpu().numpy()
    stn = st.cpu().numpy().view(np.uint64).astype(np.int64)
    start = np.concatenate([[0], np.cumsum(counts)])
    # conservation over the whole batch
    alive_total = int(F.to(torch.int64).sum().item())
    assert int(stn[S.ST_RAYS]) == total and int(stn[S.ST_GOOD]) == alive_total
    for s in range(base.n_surf - 1):
        entering, lost_i, lost_a = (int(stn[S.ST_SURF + 4 * s + k]) for k in (0, 2, 3))
        assert int(stn[S.ST_SURF + 4 * (s + 1)]) == entering - lost_i - lost_a, s
    # every instance's spot count against its stored flags
    cs = torch.cumsum(F.to(torch.int64), 0).cpu().numpy()
    per_inst = np.diff(np.concatenate([[0], cs[start[1:] - 1]]))
    assert np.array_equal(sp[:, 0].astype(np.int64), per_inst), "spot counts differ from the stored flags"
    assert np.all(np.isfinite(sp))
    # oracle parity on sampled instances
    rng = np.random.default_rng(5)
    partial = int(np.flatnonzero(counts % 256 != 0)[0])
    sample = sorted({0, len(inst) - 1, int(np.argmin(counts)), int(np.argmax(counts)), partial,
                     *rng.choice(len(inst), 3, replace=False).tolist()})
    L = length_scale(base)
    for i in sample:
        a, b = int(start[i]), int(start[i + 1])
        print(f"[C3] instance {i}: {b - a} rays against the oracle", flush=True)
        ref = O.trace(inst[i], seed=100 + i, n=b - a, nthreads=HOST_THREADS)
        g_rays = R[:, a:b].cpu().numpy()
        g_flag = F[a:b].cpu().numpy()
        assert np.array_equal(g_flag, ref.flag), i
        ok = ref.flag == 1
        err = state_err(g_rays, ref.rays, ok, L)
        assert err <= RTOL, (i, err)
        want = _spot_np(ref.rays[0][ok], ref.rays[1][ok])
        got = sp[i]
        assert got[0] == want[0], i
        assert np.all(np.abs(got[1:3] - want[1:3]) <= 1e-12 * L), (i, got, want)
        assert np.all(np.abs(got[3:] - want[3:]) <= 1e-9 * np.abs(want[3:]) + 1e-24), (i, got, want)
    print(f"[C3] sampled instances {sample} equal the oracle; spot counts of all 1024 equal their flags",
          flush=True)
    db.close()


# ---- C5 ---------------------------------------------------------------------

def _design_params(base):
    spec = diff.echelle_design_params(base)
    ps = [p.surface for p in spec.params]
    pf = [p.field for p in spec.params]
    ccd = base.n_surf - 1
    for f in (S.S_DIST + 2, S.S_ANG + 0, S.S_ANG + 1):
        if (ccd, f) not in set(zip(ps, pf)):
            ps.append(ccd)
            pf.append(f)
    return np.array(ps, dtype=np.int32), np.array(pf, dtype=np.int32)


def test_c5_design_loop_full_size(hip):
    G = 64
    base = schemes.echelle(n_per_lambda=4883, nlambda=G)
    inst, counts = schemes.tolerance_instances(base, 256, seed=2, min_rays=10_000, max_rays=500_000)
    ps, pf = _design_params(base)
    ccd = base.n_surf - 1
    comp = [j for j, (s, f) in enumerate(zip(ps, pf)) if s == ccd and f in (S.S_DIST + 2, S.S_ANG, S.S_ANG + 1)]
    n, total = len(inst), int(sum(counts))
    print(f"[C5] 256 instances, {total} rays, {len(ps)} parameters", flush=True)

    def buffers():
        return dict(rays=torch.zeros((7, total), dtype=torch.float64, device="cuda"),
                    flag=torch.zeros(total, dtype=torch.int8, device="cuda"),
                    spots=torch.zeros((n, G, hip.SPOT_STRIDE), dtype=torch.float64, device="cuda"),
                    seeds=torch.zeros((n * G, 3), dtype=torch.float64, device="cuda"),
                    loss=torch.zeros(1, dtype=torch.float64, device="cuda"),
                    grad=torch.zeros((n, len(ps)), dtype=torch.float64, device="cuda"))

    lr = torch.zeros(len(ps), dtype=torch.float64, device="cuda")

    def step(db, b):
        db.trace(seed=9, flags=S.RT_TRACE_GENERATE | S.RT_TRACE_STORE, rays=b["rays"], flag=b["flag"])
        db.group_spots(b["rays"], b["flag"], G, out=b["spots"])
        hip.spot_loss(b["spots"].view(n * G, hip.SPOT_STRIDE), seeds=b["seeds"], loss=b["loss"])
        db.adjoint(b["seeds"].view(n, G, 3), ps, pf, seed=9, grad=b["grad"])
        db.descend(ps, pf, b["grad"], lr)

    db_e, db_g = hip.DeviceBatch(inst, counts), hip.DeviceBatch(inst, counts)
    be, bg = buffers(), buffers()
    # step sizes as the design benchmark takes them: 2 um of focus and 2 arcsec of
    # tilt for the largest gradient of the batch (measured with lr = 0)
    step(db_e, be)
    torch.cuda.synchronize()
    g = be["grad"].abs().max(dim=0).values.cpu().numpy()
    lr_h = np.zeros(len(ps))
    for j in comp:
        size = 2e-3 if pf[j] == S.S_DIST + 2 else 2.0 / 3600.0 * S.DEG2RAD
        lr_h[j] = size / max(g[j], 1e-300)
    lr.copy_(torch.tensor(lr_h))
    eager = []
    for k in range(5):
        step(db_e, be)
        torch.cuda.synchronize()
        eager.append((float(be["loss"].item()), be["grad"].clone()))
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    lr_keep = lr.clone()
    lr.zero_()
    with torch.cuda.stream(s):
        step(db_g, bg)                  # the same lr = 0 first step as the eager batch
    torch.cuda.synchronize()
    lr.copy_(lr_keep)
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph, stream=s, capture_error_mode="thread_local"):
        step(db_g, bg)
    # capture records the step without running it: the tables are as after the first step
    got = []
    for k in range(5):
        with torch.cuda.stream(s):
            graph.replay()
        torch.cuda.synchronize()
        got.append((float(bg["loss"].item()), bg["grad"].clone()))
        print(f"[C5] replay {k}: loss {got[-1][0]:.9f} (eager {eager[k][0]:.9f})", flush=True)
    for k in range(5):
        assert got[k][0] == eager[k][0], (k, got[k][0], eager[k][0])
        assert torch.equal(got[k][1], eager[k][1]), k
    losses = [l for l, _ in got]
    assert all(b <= a for a, b in zip(losses, losses[1:])), losses
    assert losses[-1] < losses[0]
    db_e.close()
    db_g.close()
