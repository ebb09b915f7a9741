"""Batched adjoint (rt_hip_batch_adjoint) and batched group spots.

Bars (written here): the spots of every (instance, group) equal, bit for bit,
those of rt_hip_group_spots on that instance's rays alone (same summation
order); the gradient of every instance equals the single-scheme adjoint of
that instance (its own beam seed + i, its own group seeds) within 1e-11 of
the instance's largest gradient component (the two sum the same per-ray terms
in different orders); the generic kernel agrees with the generated one to the
same bar; the ray counter equals the batch's ray total.
"""
import numpy as np
import pytest
import torch

from rt_amd import diff
from rt_amd import scheme as S
from rt_amd import schemes

pytestmark = pytest.mark.gpu

G = 16


def setup_batch(hip, n_inst=4, seed=3):
    base = schemes.echelle(n_per_lambda=8, nlambda=G)
    inst, counts = schemes.tolerance_instances(base, n_inst, seed=seed, min_rays=2000, max_rays=12000)
    counts[1] = 777                       # a partial tile, groups of uneven size
    spec = diff.echelle_design_params(base)
    ps = np.array([p.surface for p in spec.params], dtype=np.int32)
    pf = np.array([p.field for p in spec.params], dtype=np.int32)
    db = hip.DeviceBatch(inst, counts)
    return base, inst, counts, ps, pf, db


def test_batch_group_spots_equal_per_instance(hip):
    print("[batch spots]", flush=True)
    base, inst, counts, ps, pf, db = setup_batch(hip)
    total = int(sum(counts))
    rays = torch.zeros((7, total), dtype=torch.float64, device="cuda")
    flag = torch.zeros(total, dtype=torch.int8, device="cuda")
    db.trace(seed=7, flags=S.RT_TRACE_GENERATE | S.RT_TRACE_STORE, rays=rays, flag=flag)
    sp = db.group_spots(rays, flag, G).cpu().numpy()
    start = np.concatenate([[0], np.cumsum(counts)])
    for i in range(len(inst)):
        sl = slice(int(start[i]), int(start[i + 1]))
        one = hip.group_spots(rays[:, sl].contiguous(), flag[sl].contiguous(), G).cpu().numpy()
        assert np.array_equal(sp[i], one), i
    db.close()


def test_batch_trace_group_spots_equal_stored_rays(hip):
    """rt_hip_batch_trace_group_spots (spots per instance and group formed per
    tile inside the batched trace) equals the batched trace with the rays
    stored followed by rt_hip_batch_group_spots: counts exact, centroids within
    1e-12 of their size, variances and RMS radii within 1e-10; the counters
    are the same."""
    print("[batch trace spots]", flush=True)
    base, inst, counts, ps, pf, db = setup_batch(hip)
    total = int(sum(counts))
    rays = torch.zeros((7, total), dtype=torch.float64, device="cuda")
    flag = torch.zeros(total, dtype=torch.int8, device="cuda")
    st_a = torch.zeros(S.STATS_LEN, dtype=torch.int64, device="cuda")
    st_b = torch.zeros(S.STATS_LEN, dtype=torch.int64, device="cuda")
    db.trace(seed=7, flags=S.RT_TRACE_GENERATE | S.RT_TRACE_STORE, rays=rays, flag=flag, stats=st_a)
    want = db.group_spots(rays, flag, G).cpu().numpy()
    got = db.trace_group_spots(G, seed=7, stats=st_b).cpu().numpy()
    torch.cuda.synchronize()
    assert np.array_equal(st_a.cpu().numpy(), st_b.cpu().numpy())
    for i in range(len(inst)):
        a, b = got[i], want[i]
        assert np.array_equal(a[:, 0], b[:, 0]), i
        assert np.all(np.abs(a[:, 1:3] - b[:, 1:3]) <= 1e-12 * np.maximum(np.abs(b[:, 1:3]), 1.0)), i
        assert np.all(np.abs(a[:, 3:6] - b[:, 3:6]) <= 1e-10 * np.abs(b[:, 3:6]) + 1e-12), i
    db.close()


def test_batch_adjoint_matches_single_scheme_adjoints(hip):
    print("[batch adjoint]", flush=True)
    base, inst, counts, ps, pf, db = setup_batch(hip)
    total = int(sum(counts))
    rays = torch.zeros((7, total), dtype=torch.float64, device="cuda")
    flag = torch.zeros(total, dtype=torch.int8, device="cuda")
    db.trace(seed=7, flags=S.RT_TRACE_GENERATE | S.RT_TRACE_STORE, rays=rays, flag=flag)
    sp = db.group_spots(rays, flag, G)
    loss, seeds = hip.spot_loss(sp.reshape(len(inst) * G, hip.SPOT_STRIDE).contiguous())
    seeds = seeds.reshape(len(inst), G, 3).contiguous()
    stats = torch.zeros(S.STATS_LEN, dtype=torch.int64, device="cuda")
    gB = db.adjoint(seeds, ps, pf, seed=7, stats=stats).cpu().numpy()
    assert int(stats.cpu().numpy()[S.ST_RAYS]) == total
    for i in range(len(inst)):
        ds = hip.DeviceScheme(inst[i])
        g1 = ds.adjoint(seeds[i].contiguous(), ps, pf, seed=7 + i, n=int(counts[i])).cpu().numpy()
        ds.close()
        scale = float(np.max(np.abs(g1)))
        assert scale > 0
        assert np.all(np.abs(gB[i] - g1) <= 1e-11 * scale), (i, float(np.max(np.abs(gB[i] - g1))), scale)
    # the generic kernel, to the same bar
    prev = hip.set_jit(False)
    try:
        gG = db.adjoint(seeds, ps, pf, seed=7).cpu().numpy()
    finally:
        hip.set_jit(prev)
    for i in range(len(inst)):
        scale = float(np.max(np.abs(gB[i])))
        assert np.all(np.abs(gG[i] - gB[i]) <= 1e-11 * scale), i
    db.close()


def _design_params(base):
    """the 48 design parameters plus the compensators (CCD focus and tilts)"""
    spec = diff.echelle_design_params(base)
    ps = [p.surface for p in spec.params]
    pf = [p.field for p in spec.params]
    ccd = base.n_surf - 1
    for f in (S.S_DIST + 2, S.S_ANG + 0, S.S_ANG + 1):
        if (ccd, f) not in set(zip(ps, pf)):
            ps.append(ccd)
            pf.append(f)
    return np.array(ps, dtype=np.int32), np.array(pf, dtype=np.int32)


def test_batch_descend_matches_host_update(hip):
    """rt_hip_batch_descend against the same step applied to the packed
    tables on the host (sin/cos of stepped tilts from the host libm): the
    traces of the two batches agree to the parity bar."""
    print("[descend]", flush=True)
    base, inst, counts, _, _, db = setup_batch(hip, n_inst=3, seed=9)
    ps, pf = _design_params(base)
    rng = np.random.default_rng(4)
    grad = rng.normal(0.0, 1.0, (len(inst), len(ps)))
    lr = np.zeros(len(ps))
    ccd = base.n_surf - 1
    for j, (s, f) in enumerate(zip(ps, pf)):
        if s == ccd:
            lr[j] = 1e-3 if f == S.S_DIST + 2 else 2e-5
    lr[0] = 1e-9                                   # one curvature too
    db.descend(ps, pf, torch.tensor(grad, device="cuda"), torch.tensor(lr, device="cuda"))
    host = []
    for i, sch in enumerate(inst):
        t = np.array(sch.table, dtype=np.float64).copy()
        for j, (s, f) in enumerate(zip(ps, pf)):
            t[s * S.SURF_STRIDE + f] -= lr[j] * grad[i, j]
            if S.S_ANG <= f < S.S_ANG + 3:
                a = t[s * S.SURF_STRIDE + f]
                t[s * S.SURF_STRIDE + S.S_SIN + f - S.S_ANG] = np.sin(a)
                t[s * S.SURF_STRIDE + S.S_COS + f - S.S_ANG] = np.cos(a)
        c = S.Scheme(list(sch.surfaces), sch.beam, name=sch.name)
        c.table = t
        host.append(c)
    dh = hip.DeviceBatch(host, counts)
    total = int(sum(counts))
    out = []
    for b in (db, dh):
        rays = torch.zeros((7, total), dtype=torch.float64, device="cuda")
        flag = torch.zeros(total, dtype=torch.int8, device="cuda")
        b.trace(seed=3, flags=S.RT_TRACE_GENERATE | S.RT_TRACE_STORE, rays=rays, flag=flag)
        out.append((rays.cpu().numpy(), flag.cpu().numpy()))
    (ra, fa), (rb, fb) = out
    assert np.array_equal(fa, fb)
    ok = fa == 1
    L = 1000.0
    assert np.max(np.abs(ra[:3, ok] - rb[:3, ok])) <= 1e-12 * L
    assert np.max(np.abs(ra[3:6, ok] - rb[3:6, ok])) <= 1e-12
    db.close()
    dh.close()


@pytest.mark.parametrize("fused", [False, True], ids=["stored-rays", "spots-in-trace"])
def test_design_step_graph_replays_match_eager(hip, fused):
    """One design-loop step (batched trace with stored rays and spots from
    them, or the trace forming the spots itself; loss and seeds, batched
    adjoint, descent on the compensators) captured in a graph: three replays
    give bit for bit the losses and tables of three eager steps on an
    identical batch."""
    print("[design graph]", "spots in the trace" if fused else "stored rays", flush=True)

    def make():
        base, inst, counts, _, _, db = setup_batch(hip, n_inst=3, seed=13)
        return base, inst, counts, db

    base, inst, counts, db_e = make()
    _, _, _, db_g = make()
    ps, pf = _design_params(base)
    ccd = base.n_surf - 1
    lr = torch.tensor([1e-4 if (s == ccd and f == S.S_DIST + 2) else (1e-6 if s == ccd and f in (S.S_ANG, S.S_ANG + 1)
                                                                      else 0.0) for s, f in zip(ps, pf)],
                      dtype=torch.float64, device="cuda")
    total = int(sum(counts))
    n = len(inst)

    def buffers():
        return dict(rays=torch.zeros((7, total), dtype=torch.float64, device="cuda"),
                    flag=torch.zeros(total, dtype=torch.int8, device="cuda"),
                    spots=torch.zeros((n, G, hip.SPOT_STRIDE), dtype=torch.float64, device="cuda"),
                    seeds=torch.zeros((n * G, 3), dtype=torch.float64, device="cuda"),
                    loss=torch.zeros(1, dtype=torch.float64, device="cuda"),
                    grad=torch.zeros((n, len(ps)), dtype=torch.float64, device="cuda"))

    def step(db, b):
        if fused:
            db.trace_group_spots(G, seed=11, out=b["spots"])
        else:
            db.trace(seed=11, flags=S.RT_TRACE_GENERATE | S.RT_TRACE_STORE, rays=b["rays"], flag=b["flag"])
            db.group_spots(b["rays"], b["flag"], G, out=b["spots"])
        hip.spot_loss(b["spots"].view(n * G, hip.SPOT_STRIDE), seeds=b["seeds"], loss=b["loss"])
        db.adjoint(b["seeds"].view(n, G, 3), ps, pf, seed=11, grad=b["grad"])
        db.descend(ps, pf, b["grad"], lr)

    be, bg = buffers(), buffers()
    losses_e = []
    for _ in range(3):
        step(db_e, be)
        losses_e.append(float(be["loss"].item()))
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        # the first step runs eagerly (kernels compiled, buffers allocated), then one is captured
        step(db_g, bg)
    torch.cuda.synchronize()
    losses_g = [float(bg["loss"].item())]
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s, capture_error_mode="thread_local"):
        step(db_g, bg)
    for _ in range(2):
        g.replay()
        torch.cuda.synchronize()
        losses_g.append(float(bg["loss"].item()))
    assert losses_g == losses_e, (losses_g, losses_e)
    assert torch.equal(bg["grad"], be["grad"])
    assert torch.equal(bg["rays"], be["rays"])
    db_e.close()
    db_g.close()


def test_adjoint_table_matches_per_ray_splines(hip):
    """The generated adjoints of a grid-wavelength beam read their per-
    wavelength values from a table in the forward and the backward steps (one
    scheme: mu and the grating term from the derived table; a batch: n1 and n2
    from the spline-value table).  The batch's gradients equal those of the
    per-ray spline evaluation bit for bit.  The single scheme's agree per
    component within 1e-8 relative (plus 1e-13 of the largest component):
    every per-ray value is the same, but the table kernel's register count
    sets its occupancy, hence its grid, and the gradient sums run in a fixed
    order per grid (bit-identity is reported)."""
    print("[adjoint table]", flush=True)
    base, inst, counts, ps, pf, db = setup_batch(hip, n_inst=3, seed=21)
    gen = torch.Generator().manual_seed(4)
    seeds = (torch.rand((len(inst), G, 3), generator=gen, dtype=torch.float64) * 1e-3).cuda()
    prev_jit = hip.set_jit(True)
    try:
        ds = hip.DeviceScheme(base)
        out = {}
        for tab in (False, True):
            prev = hip.set_ltab(tab)
            try:
                g1 = ds.adjoint(seeds[0].contiguous(), ps, pf, seed=5).cpu().numpy()
                gb = db.adjoint(seeds, ps, pf, seed=5).cpu().numpy()
            finally:
                hip.set_ltab(prev)
            out[tab] = (g1, gb)
        assert np.all(np.isfinite(out[True][0])) and np.abs(out[True][0]).max() > 0
        a, b = out[True][0], out[False][0]
        scale = np.abs(b).max()
        rel = float(np.max(np.abs(a - b) / (np.abs(b) + 1e-13 * scale)))
        print(f"[adjoint table] single scheme: worst relative difference {rel:.3e}, bit-identical "
              f"{np.array_equal(a, b)}", flush=True)
        assert np.all(np.abs(a - b) <= 1e-8 * np.abs(b) + 1e-13 * scale)
        assert np.array_equal(out[True][1], out[False][1])
        ds.close()
    finally:
        hip.set_jit(prev_jit)
    db.close()


def test_batch_adjoint_with_a_beam_per_instance(hip):
    """The batched adjoint of instances with their own wavelength sets
    (rt_hip_batch_set_beams) equals the single-scheme adjoint of each instance
    with its own beam, within 1e-11 of the instance's largest component."""
    import copy
    print("[batch adjoint beams]", flush=True)
    base, inst, counts, ps, pf, db = setup_batch(hip, n_inst=3, seed=31)
    beams = []
    for i, sch in enumerate(inst):
        b = copy.deepcopy(sch.beam)
        b.range = (0.45 + 0.05 * i, 0.65 + 0.05 * i)
        beams.append(b)
    db.set_beams(beams)
    gen = torch.Generator().manual_seed(6)
    seeds = (torch.rand((len(inst), G, 3), generator=gen, dtype=torch.float64) * 1e-3).cuda()
    gb = db.adjoint(seeds, ps, pf, seed=5).cpu().numpy()
    for i in range(len(inst)):
        own = copy.deepcopy(inst[i])
        own.beam = beams[i]
        ds = hip.DeviceScheme(own)
        g1 = ds.adjoint(seeds[i].contiguous(), ps, pf, seed=5 + i, n=int(counts[i])).cpu().numpy()
        ds.close()
        scale = max(float(np.abs(g1).max()), 1e-300)
        assert np.abs(gb[i] - g1).max() <= 1e-11 * scale, i
    db.close()


def test_design_graph_survives_later_buffer_growth(hip):
    """A captured design step keeps valid working memory when later eager calls
    need larger buffers: between replays, an eager adjoint of the same batch
    over more parameters (larger gradient partials), an eager adjoint of a
    larger batch and a single-scheme adjoint run; every replay still equals
    the eager step of an identical batch bit for bit."""
    print("[design graph growth]", flush=True)

    def make(n_inst, seed):
        base, inst, counts, _, _, db = setup_batch(hip, n_inst=n_inst, seed=seed)
        return base, inst, counts, db

    base, inst, counts, db_e = make(2, 17)
    _, _, _, db_g = make(2, 17)
    ps, pf = _design_params(base)
    ccd = base.n_surf - 1
    lr = torch.tensor([1e-4 if (s == ccd and f == S.S_DIST + 2) else 0.0 for s, f in zip(ps, pf)],
                      dtype=torch.float64, device="cuda")
    # a shorter list for the captured step, so the later eager call over the full one grows the partials
    keep = [j for j, (s, f) in enumerate(zip(ps, pf)) if s >= ccd - 3]
    ps_s, pf_s, lr_s = ps[keep], pf[keep], lr[keep].contiguous()
    total = int(sum(counts))
    n = len(inst)

    def buffers(npar):
        return dict(rays=torch.zeros((7, total), dtype=torch.float64, device="cuda"),
                    flag=torch.zeros(total, dtype=torch.int8, device="cuda"),
                    spots=torch.zeros((n, G, hip.SPOT_STRIDE), dtype=torch.float64, device="cuda"),
                    seeds=torch.zeros((n * G, 3), dtype=torch.float64, device="cuda"),
                    loss=torch.zeros(1, dtype=torch.float64, device="cuda"),
                    grad=torch.zeros((n, npar), dtype=torch.float64, device="cuda"))

    def step(db, b):
        db.trace(seed=11, flags=S.RT_TRACE_GENERATE | S.RT_TRACE_STORE, rays=b["rays"], flag=b["flag"])
        db.group_spots(b["rays"], b["flag"], G, out=b["spots"])
        hip.spot_loss(b["spots"].view(n * G, hip.SPOT_STRIDE), seeds=b["seeds"], loss=b["loss"])
        db.adjoint(b["seeds"].view(n, G, 3), ps_s, pf_s, seed=11, grad=b["grad"])
        db.descend(ps_s, pf_s, b["grad"], lr_s)

    be, bg = buffers(len(ps_s)), buffers(len(ps_s))
    eager = []
    for _ in range(4):
        step(db_e, be)
        eager.append((float(be["loss"].item()), be["grad"].clone()))
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        step(db_g, bg)
    torch.cuda.synchronize()
    got = [(float(bg["loss"].item()), bg["grad"].clone())]
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s, capture_error_mode="thread_local"):
        step(db_g, bg)
    big_base, big_inst, big_counts, db_big = make(5, 19)
    seeds_big = torch.full((len(big_inst), G, 3), 1e-4, dtype=torch.float64, device="cuda")
    seeds_g = torch.full((n, G, 3), 1e-4, dtype=torch.float64, device="cuda")
    ds = hip.DeviceScheme(base)
    for k in range(3):
        with torch.cuda.stream(s):
            g.replay()
        torch.cuda.synchronize()
        got.append((float(bg["loss"].item()), bg["grad"].clone()))
        # eager calls between replays that need larger working buffers
        db_g.adjoint(seeds_g, ps, pf, seed=3 + k)                 # more parameters on the same batch
        db_big.adjoint(seeds_big, ps, pf, seed=3 + k)             # another, larger batch
        ds.adjoint(seeds_g[0].contiguous(), ps, pf, seed=3 + k)   # a single scheme
        torch.cuda.synchronize()
    for k, ((le, ge), (lg, gg)) in enumerate(zip(eager, got)):
        assert le == lg, (k, le, lg)
        assert torch.equal(ge, gg), k
    ds.close()
    db_big.close()
    db_e.close()
    db_g.close()
