"""The adjoint kernel (rt_hip_adjoint) and the differentiable spot loss.

* GPU adjoint against the host build of the same derivative formulas reversed
  over the oracle's forward states (tests/native): 1e-9 relative.
* Gradient of the spot loss (sum over wavelengths of RMS radius plus the
  squared distance of the centroid from a dispersion line) over the echelle's
  48 design parameters against central finite differences with h = 1e-7
  relative (1e-7 absolute for parameters at 0), of the GPU forward and of the
  oracle: |adjoint - FD| <= 1e-6 |FD| + the FD's rounding noise (8 eps |loss| /
  2h), for parameters whose FD runs keep the same surviving rays and Newton
  iteration counts (checked; at least 40 of the 48 must qualify).
* The linear dispersion (SpotFunction over linear_dispersion) against the
  same finite differences.
* The torch.autograd.Function returns the same gradient.
"""
import math

import numpy as np
import pytest
import torch

from rt_amd import diff
from rt_amd import scheme as S
from rt_amd import schemes
from oracle import oracle as O

from test_adjoint_cpu import chain_adjoint, host  # noqa: F401  (fixture)

pytestmark = pytest.mark.gpu


def gpu_adjoint(hip, sch, seed, seeds_np, ps, pf):
    ds = hip.DeviceScheme(sch)
    seeds = torch.from_numpy(np.ascontiguousarray(seeds_np)).cuda()
    st = torch.zeros(S.STATS_LEN, dtype=torch.int64, device="cuda")
    print(f"[adjoint] {sch.name} n={sch.beam.n_rays} params={len(ps)}", flush=True)
    g = ds.adjoint(seeds, ps, pf, seed=seed, stats=st)
    torch.cuda.synchronize()
    ds.close()
    return g.cpu().numpy(), st.cpu().numpy().view(np.uint64)


def test_gpu_adjoint_matches_host_chain(hip, host):  # noqa: F811
    sch = schemes.echelle(n_per_lambda=3, nlambda=64)
    seed = 5
    G = 64
    rng = np.random.default_rng(1)
    seeds = np.stack([rng.normal(size=G), rng.normal(size=G), rng.normal(size=G) * 0.1], axis=1)
    ref = O.trace(sch, seed=seed)
    gidx = np.arange(sch.beam.n_rays) % G
    u = seeds[gidx, 0] + seeds[gidx, 2] * ref.rays[0]
    v = seeds[gidx, 1] + seeds[gidx, 2] * ref.rays[1]
    H, _ = chain_adjoint(host, sch, seed, u, v)
    spec = diff.echelle_design_params(sch)
    ps = [p.surface for p in spec.params]
    pf = [p.field for p in spec.params]
    g, st = gpu_adjoint(hip, sch, seed, seeds, ps, pf)
    assert int(st[S.ST_ADJ_SKIP]) == 0
    assert int(st[S.ST_GOOD]) == int(np.sum(ref.flag == 1))
    h = np.array([H[s, host.adj_host_slot(f)] for s, f in zip(ps, pf)])
    scale = np.maximum(np.abs(h), 1e-12 * np.max(np.abs(h)))
    err = np.abs(g - h) / scale
    assert np.max(err) <= 1e-9, (int(np.argmax(err)), float(np.max(err)))


def _oracle_loss(sch, seed, n_groups, target, weight):
    r = O.trace(sch, seed=seed)
    ok = r.flag == 1
    g = np.arange(sch.beam.n_rays) % n_groups
    loss = 0.0
    for k in range(n_groups):
        m = ok & (g == k)
        if not m.any():
            continue
        x, y = r.rays[0, m], r.rays[1, m]
        cx, cy = x.mean(), y.mean()
        rms = math.sqrt(np.mean((x - cx) ** 2 + (y - cy) ** 2))
        loss += rms + weight * ((cx - target[k, 0]) ** 2 + (cy - target[k, 1]) ** 2)
    return loss, r


def _counters(st, n_surf):
    return np.array([int(st[S.ST_SURF + 4 * s + k]) for s in range(n_surf) for k in (1, 2, 3)])


def test_spot_loss_gradient_matches_finite_differences(hip):
    base = schemes.echelle(n_per_lambda=48, nlambda=64)
    spec = diff.echelle_design_params(base)
    n_groups, seed, weight = 64, 9, 0.5
    v0 = spec.values(base)
    # target: the least-squares dispersion line through the nominal centroids, in wavelength
    loss0 = diff.SpotLoss(base, spec, n_groups=n_groups, seed=seed)
    cnt, cx, cy, rms = (t.cpu().numpy() for t in loss0.spots(v0))
    lam = np.array([O.gen_rays(base.beam.pack(), seed, k, 1)[6, 0] for k in range(n_groups)])
    tx = np.polyval(np.polyfit(lam, cx, 1), lam)
    ty = np.polyval(np.polyfit(lam, cy, 1), lam)
    target = np.stack([tx, ty], axis=1)
    L = diff.SpotLoss(base, spec, n_groups=n_groups, target=target, weight=weight, seed=seed)
    val, grad = L.value_and_grad(v0)
    oval, oref = _oracle_loss(spec.apply(base, v0), seed, n_groups, target, weight)
    assert abs(val - oval) <= 1e-11 * abs(oval)
    base_cnt = _counters(oref.stats, base.n_surf)
    checked = 0
    for j, p in enumerate(spec.params):
        x = v0[j]
        h = 1e-7 * abs(x) if x != 0.0 else 1e-7
        vp, vm = v0.copy(), v0.copy()
        vp[j] += h
        vm[j] -= h
        lp, rp = _oracle_loss(spec.apply(base, vp), seed, n_groups, target, weight)
        lm, rm = _oracle_loss(spec.apply(base, vm), seed, n_groups, target, weight)
        same = (np.array_equal(rp.flag, oref.flag) and np.array_equal(rm.flag, oref.flag) and
                np.array_equal(_counters(rp.stats, base.n_surf), base_cnt) and
                np.array_equal(_counters(rm.stats, base.n_surf), base_cnt))
        if not same:
            print(f"[fd] parameter {j} ({p.surface}, {p.field}): rays or iterations change, skipped")
            continue
        noise = 8 * np.finfo(float).eps * abs(val) / (2 * h)
        fd_cpu = (lp - lm) / (2 * h)
        gp_, _ = L.value_and_grad(vp, grad=False)
        gm_, _ = L.value_and_grad(vm, grad=False)
        fd_gpu = (gp_ - gm_) / (2 * h)
        print(f"[fd] {j:2d} s{p.surface} f{p.field} adj {grad[j]:+.9e} fd_cpu {fd_cpu:+.9e} fd_gpu {fd_gpu:+.9e}")
        assert abs(grad[j] - fd_cpu) <= 1e-6 * abs(fd_cpu) + noise, (j, grad[j], fd_cpu, noise)
        assert abs(grad[j] - fd_gpu) <= 1e-6 * abs(fd_gpu) + noise, (j, grad[j], fd_gpu, noise)
        checked += 1
    assert checked >= 40, checked


def _oracle_dispersion(sch, seed, n_groups, lam, axis):
    r = O.trace(sch, seed=seed)
    ok = r.flag == 1
    g = np.arange(sch.beam.n_rays) % n_groups
    cnt, cx, cy, rms = (np.zeros(n_groups) for _ in range(4))
    for k in range(n_groups):
        m = ok & (g == k)
        cnt[k] = m.sum()
        if cnt[k]:
            cx[k], cy[k] = r.rays[0, m].mean(), r.rays[1, m].mean()
    return diff.linear_dispersion(lam, axis)(cnt, cx, cy, rms)[0], r


def test_linear_dispersion_gradient_matches_finite_differences(hip):
    """Gradient of the linear dispersion (least-squares slope of the y
    centroids against wavelength, mm per um) over the 48 design parameters,
    through SpotFunction's seeds and the same adjoint: against central finite
    differences of the GPU forward and of the oracle, h = 1e-7 relative, bar
    1e-6 relative + the FD's rounding noise (8 eps times the slope's rounding
    scale, max |centroid| sum |l - l_mean| / sum (l - l_mean)^2, over 2h), for
    parameters whose FD runs keep the same surviving rays and Newton iteration
    counts (at least 40)."""
    base = schemes.echelle(n_per_lambda=48, nlambda=64)
    spec = diff.echelle_design_params(base)
    n_groups, seed, axis = 64, 9, 1
    lam = np.array([O.gen_rays(base.beam.pack(), seed, k, 1)[6, 0] for k in range(n_groups)])
    F = diff.SpotFunction(base, spec, n_groups=n_groups, f=diff.linear_dispersion(lam, axis), seed=seed)
    v0 = spec.values(base)
    val, grad = F.value_and_grad(v0)
    oval, oref = _oracle_dispersion(spec.apply(base, v0), seed, n_groups, lam, axis)
    print(f"[dispersion] {val:.12e} oracle {oval:.12e}", flush=True)
    assert abs(val - oval) <= 1e-11 * abs(oval)
    base_cnt = _counters(oref.stats, base.n_surf)
    # rounding scale of D: centroid rounding (eps |c|) times the slope's
    # sensitivity sum |l - l_mean| / sum (l - l_mean)^2
    cnt, cx, cy, _ = (t.cpu().numpy() for t in F.loss.spots(v0))
    c = (cx if axis == 0 else cy)[cnt > 0]
    dl = lam[cnt > 0] - lam[cnt > 0].mean()
    d_scale = max(abs(val), float(np.abs(c).max() * np.abs(dl).sum() / np.sum(dl * dl)))
    checked = 0
    for j, p in enumerate(spec.params):
        x = v0[j]
        h = 1e-7 * abs(x) if x != 0.0 else 1e-7
        vp, vm = v0.copy(), v0.copy()
        vp[j] += h
        vm[j] -= h
        dp, rp = _oracle_dispersion(spec.apply(base, vp), seed, n_groups, lam, axis)
        dm, rm = _oracle_dispersion(spec.apply(base, vm), seed, n_groups, lam, axis)
        same = (np.array_equal(rp.flag, oref.flag) and np.array_equal(rm.flag, oref.flag) and
                np.array_equal(_counters(rp.stats, base.n_surf), base_cnt) and
                np.array_equal(_counters(rm.stats, base.n_surf), base_cnt))
        if not same:
            continue
        noise = 8 * np.finfo(float).eps * d_scale / (2 * h)
        fd_cpu = (dp - dm) / (2 * h)
        gp_, _ = F.value_and_grad(vp, grad=False)
        gm_, _ = F.value_and_grad(vm, grad=False)
        fd_gpu = (gp_ - gm_) / (2 * h)
        assert abs(grad[j] - fd_cpu) <= 1e-6 * abs(fd_cpu) + noise, (j, grad[j], fd_cpu, noise)
        assert abs(grad[j] - fd_gpu) <= 1e-6 * abs(fd_gpu) + noise, (j, grad[j], fd_gpu, noise)
        checked += 1
    print(f"[dispersion] {checked} parameters checked", flush=True)
    assert checked >= 40, checked


def test_autograd_function(hip):
    base = schemes.echelle(n_per_lambda=16, nlambda=64)
    spec = diff.echelle_design_params(base)
    L = diff.SpotLoss(base, spec, n_groups=64, seed=2)
    v0 = spec.values(base)
    _, grad = L.value_and_grad(v0)
    SchemeLoss = diff.scheme_loss_function()
    x = torch.tensor(v0, dtype=torch.float64, requires_grad=True)
    loss = SchemeLoss.apply(x, L)
    loss.backward()
    assert np.array_equal(x.grad.numpy(), grad)


@pytest.mark.parametrize("n,n_groups", [(10007, 64), (4096, 1), (3001, 7), (20000, 300), (0, 5)])
def test_group_spots_match_two_pass_moments(hip, n, n_groups):
    """rt_hip_group_spots (fixed-order Welford per chunk, Chan across chunks)
    against two-pass moments in numpy: counts exact, centroids, variances and
    RMS radii within 1e-12 relative; groups without a live ray give zeros."""
    print("[group spots]", n, n_groups, flush=True)
    rng = np.random.default_rng(n + n_groups)
    rays = np.zeros((7, max(n, 1)))
    rays[0] = rng.normal(30.0, 2.0, rays.shape[1])
    rays[1] = rng.normal(-4.0, 0.5, rays.shape[1])
    flag = (rng.random(rays.shape[1]) < 0.8).astype(np.int8)
    if n_groups > 2 and n > 0:
        flag[2::n_groups] = 0                           # group 2 has no live ray
    R = torch.tensor(rays[:, :n] if n else rays[:, :0], dtype=torch.float64, device="cuda").contiguous()
    F = torch.tensor(flag[:n], dtype=torch.int8, device="cuda")
    sp = hip.group_spots(R, F, n_groups).cpu().numpy()
    torch.cuda.synchronize()
    for g in range(n_groups):
        idx = np.arange(g, n, n_groups)
        live = idx[flag[idx] == 1] if n else idx
        assert sp[g, 0] == len(live)
        if len(live) == 0:
            assert np.all(sp[g] == 0.0)
            continue
        x, y = rays[0, live], rays[1, live]
        cx, cy = x.mean(), y.mean()
        vx, vy = np.mean((x - cx) ** 2), np.mean((y - cy) ** 2)
        want = np.array([cx, cy, vx, vy, math.sqrt(vx + vy)])
        assert np.all(np.abs(sp[g, 1:] - want) <= 1e-12 * np.maximum(np.abs(want), 1.0)), (g, sp[g], want)


def spots_close(a, b, what):
    """group spots of the same rays summed in two orders: counts exact,
    centroids within 1e-12 of their size (floor 1 mm), variances and RMS
    radii within 1e-10 relative (floor 1e-12 mm^2 / mm)"""
    assert np.array_equal(a[:, 0], b[:, 0]), what
    c = np.abs(a[:, 1:3] - b[:, 1:3]) <= 1e-12 * np.maximum(np.abs(b[:, 1:3]), 1.0)
    v = np.abs(a[:, 3:6] - b[:, 3:6]) <= 1e-10 * np.abs(b[:, 3:6]) + 1e-12
    assert np.all(c) and np.all(v), (what, np.argwhere(~c), np.argwhere(~v))


@pytest.mark.parametrize("n_groups,launch", [(64, None), (7, None), (300, None), (64, (128, 0, 2)), (5, "generic")])
def test_trace_spots_equal_stored_rays(hip, n_groups, launch):
    """rt_hip_trace_spots (spots formed per workgroup tile inside the trace)
    equals rt_hip_trace with the rays stored followed by rt_hip_group_spots,
    on groups that divide the workgroup, that do not, that outnumber it, on a
    128-lane launch and on the generic kernel; the trace's counters are the
    same."""
    sch = schemes.echelle(n_per_lambda=157, nlambda=64)        # 10,048 rays: a partial last tile
    n = sch.beam.n_rays
    print(f"[trace spots] groups {n_groups} launch {launch}", flush=True)
    ds = hip.DeviceScheme(sch)
    prev = hip.set_jit(launch != "generic")
    try:
        if isinstance(launch, tuple):
            ds.set_launch(*launch)
        R = torch.zeros((7, n), dtype=torch.float64, device="cuda")
        F = torch.zeros(n, dtype=torch.int8, device="cuda")
        st_a = torch.zeros(S.STATS_LEN, dtype=torch.int64, device="cuda")
        st_b = torch.zeros(S.STATS_LEN, dtype=torch.int64, device="cuda")
        ds.trace(R, F, seed=11, flags=S.RT_TRACE_GENERATE | S.RT_TRACE_STORE, stats=st_a)
        want = hip.group_spots(R, F, n_groups).cpu().numpy()
        got = ds.trace_spots(n_groups, seed=11, stats=st_b).cpu().numpy()
        torch.cuda.synchronize()
        kernel = hip.last_kernel()
    finally:
        hip.set_jit(prev)
    assert kernel == ("generic" if launch == "generic" else "generated"), kernel
    assert np.array_equal(st_a.cpu().numpy(), st_b.cpu().numpy())
    assert want[:, 0].sum() == int(st_a.cpu().numpy()[S.ST_GOOD]) > 0
    spots_close(got, want, f"groups {n_groups}")
    ds.close()


def test_spot_loss_and_seeds(hip):
    """rt_hip_spot_loss against the loss and seeds written out in numpy, with
    and without a target, including a group without live rays."""
    print("[spot loss]", flush=True)
    rng = np.random.default_rng(11)
    G = 37
    sp = np.zeros((G, 6))
    sp[:, 0] = rng.integers(1, 500, G)
    sp[5, 0] = 0
    sp[:, 1] = rng.normal(10, 3, G)
    sp[:, 2] = rng.normal(-2, 1, G)
    sp[:, 3] = rng.random(G) * 1e-3
    sp[:, 4] = rng.random(G) * 1e-3
    sp[:, 5] = np.sqrt(sp[:, 3] + sp[:, 4])
    sp[5, 1:] = 0.0
    tgt = rng.normal(0, 1, (G, 2))
    for target, w in ((None, 0.0), (tgt, 0.7)):
        t = np.zeros((G, 2)) if target is None else target
        live = sp[:, 0] > 0
        n, cx, cy, rms = sp[:, 0], sp[:, 1], sp[:, 2], sp[:, 5]
        loss = np.sum((rms + w * ((cx - t[:, 0]) ** 2 + (cy - t[:, 1]) ** 2))[live])
        b = np.where(live, 1.0 / np.maximum(n * rms, 1e-300), 0.0)
        ax = np.where(live, -b * cx + 2 * w * (cx - t[:, 0]) / np.maximum(n, 1), 0.0)
        ay = np.where(live, -b * cy + 2 * w * (cy - t[:, 1]) / np.maximum(n, 1), 0.0)
        S_ = torch.tensor(sp, dtype=torch.float64, device="cuda")
        T_ = None if target is None else torch.tensor(target, dtype=torch.float64, device="cuda")
        gl, gs = hip.spot_loss(S_, T_, w)
        gs = gs.cpu().numpy()
        assert abs(float(gl.item()) - loss) <= 1e-12 * abs(loss)
        want = np.stack([ax, ay, b], axis=1)
        assert np.all(np.abs(gs - want) <= 1e-12 * np.maximum(np.abs(want), 1e-30))


def _oracle_loss_fixed(sch, seed, n_groups, target, weight, iters, mode):
    """the spot loss of the oracle's trace with the Newton iteration counts
    recorded (mode 1) or held at the given ones (mode 2); vectorised over
    groups (two-pass centroid and spread)"""
    r = O.trace(sch, seed=seed, nthreads=16, iters=iters, iters_mode=mode)
    ok = r.flag == 1
    g = (np.arange(sch.beam.n_rays) % n_groups)[ok]
    x, y = r.rays[0][ok], r.rays[1][ok]
    cnt = np.bincount(g, minlength=n_groups).astype(np.float64)
    cx = np.bincount(g, weights=x, minlength=n_groups) / np.maximum(cnt, 1)
    cy = np.bincount(g, weights=y, minlength=n_groups) / np.maximum(cnt, 1)
    m2 = np.bincount(g, weights=(x - cx[g]) ** 2 + (y - cy[g]) ** 2, minlength=n_groups)
    live = cnt > 0
    rms = np.sqrt(m2[live] / cnt[live])
    loss = float(np.sum(rms + weight * ((cx[live] - target[live, 0]) ** 2 + (cy[live] - target[live, 1]) ** 2)))
    return loss, r


def test_c4_spot_loss_gradient_full_size(hip):
    """The differentiable configuration at its stated size: the echelle at
    15625 rays x 64 wavelengths (1e6 rays), the spot loss of 64 wavelength
    groups (RMS radius + 0.5 x squared deviation of the centroid from the
    dispersion line), the adjoint over the 48 design parameters against
    central finite differences of the oracle with h = 1e-7 relative, bar
    1e-6 relative + the FD rounding noise.  At 1e6 rays a perturbation of
    1e-7 makes some ray's Newton loop stop one iteration earlier or later,
    which moves that ray by up to the 1e-7 mm stopping tolerance (a jump, not
    a derivative); the finite differences therefore hold every ray's
    iteration counts at the nominal trace's (oracle_trace_iters), which is
    the function the adjoint differentiates.  Parameters whose perturbed
    traces change the set of surviving rays are skipped; at least 44 of the 48
    must qualify.  The finite differences' rounding noise is measured, not
    modelled: a second central difference with h = 1e-6 relative has 10x less
    noise and a truncation error far below the bar (the two agree with
    h = 1e-5 to 1e-7 relative where the noise is small), so |fd(1e-7) -
    fd(1e-6)| bounds the noise of fd(1e-7).  It matters for components many
    orders of magnitude below the loss scale: d/dk of the paraboloid
    collimator is 0.33 against a loss of 462, and fd(1e-7) misses it by 2e-5
    relative from rounding of the 1e6-ray sums alone."""
    base = schemes.echelle(n_per_lambda=15625, nlambda=64)
    assert base.beam.n_rays == 1_000_000
    spec = diff.echelle_design_params(base)
    assert len(spec) == 48
    n_groups, seed, weight = 64, 9, 0.5
    v0 = spec.values(base)
    loss0 = diff.SpotLoss(base, spec, n_groups=n_groups, seed=seed)
    cnt, cx, cy, rms = (t.cpu().numpy() for t in loss0.spots(v0))
    lam = np.array([O.gen_rays(base.beam.pack(), seed, k, 1)[6, 0] for k in range(n_groups)])
    target = np.stack([np.polyval(np.polyfit(lam, cx, 1), lam), np.polyval(np.polyfit(lam, cy, 1), lam)], axis=1)
    L = diff.SpotLoss(base, spec, n_groups=n_groups, target=target, weight=weight, seed=seed)
    print("[C4] adjoint at 1e6 rays", flush=True)
    val, grad = L.value_and_grad(v0)
    iters = np.zeros((base.n_surf, base.beam.n_rays), dtype=np.uint32)
    oval, oref = _oracle_loss_fixed(spec.apply(base, v0), seed, n_groups, target, weight, iters, 1)
    assert abs(val - oval) <= 1e-11 * abs(oval), (val, oval)
    checked = 0
    worst = 0.0
    def central(j, rel):
        x = v0[j]
        h = rel * abs(x) if x != 0.0 else rel
        vp, vm = v0.copy(), v0.copy()
        vp[j] += h
        vm[j] -= h
        lp, rp = _oracle_loss_fixed(spec.apply(base, vp), seed, n_groups, target, weight, iters.copy(), 2)
        lm, rm = _oracle_loss_fixed(spec.apply(base, vm), seed, n_groups, target, weight, iters.copy(), 2)
        same = np.array_equal(rp.flag, oref.flag) and np.array_equal(rm.flag, oref.flag)
        return (lp - lm) / (2 * h), h, same

    for j, p in enumerate(spec.params):
        fd, h, same = central(j, 1e-7)
        fd6, _, same6 = central(j, 1e-6)
        if not (same and same6):
            print(f"[C4 fd] parameter {j} ({p.surface}, {p.field}): surviving rays change, skipped", flush=True)
            continue
        noise = max(8 * np.finfo(float).eps * abs(val) / (2 * h), abs(fd - fd6))
        rel = abs(grad[j] - fd) / max(abs(fd), 1e-300)
        worst = max(worst, rel)
        print(f"[C4 fd] {j:2d} s{p.surface} f{p.field} adjoint {grad[j]:+.9e} fd {fd:+.9e} (h 1e-6: {fd6:+.9e}) "
              f"rel {rel:.1e} noise {noise:.1e}", flush=True)
        assert abs(grad[j] - fd) <= 1e-6 * abs(fd) + 2 * noise, (j, grad[j], fd, fd6, noise)
        checked += 1
    print(f"[C4] {checked} of 48 parameters checked, worst relative difference {worst:.2e}", flush=True)
    assert checked >= 44, checked
