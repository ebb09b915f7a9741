"""Parity of the fused HIP trace (rt_hip_trace) with the CPU oracle.

Bar (tolerance written here, 1e-12 relative):
  positions   |a-b| <= 1e-12 * max(|a|, |b|, L)   L = length scale of the scheme
              (largest translation between surface frames, >= 1 mm): a
              coordinate near zero that came out of a frame transform over
              ~1 m carries rounding of ~ulp(1 m), so "relative" is taken
              against the magnitudes the transform handled;
  cosines     |a-b| <= 1e-12 * max(|a|, |b|, 1);
  wavelength  |a-b| <= 1e-12 * |a|;
  flags, integer CCD counts, statistics counters: bit-exact.
Every surface's state is compared (dump mode) so a wrong stage is named.
The oracle built with FMA contraction (what hipcc emits) differs from the
strict oracle by up to 1.1e-12 * 1 mm on the echelle, i.e. 1e-15 of L.
"""
import numpy as np
import pytest
import torch

from rt_amd import scheme as S
from rt_amd import schemes
from oracle import oracle as O

pytestmark = pytest.mark.gpu

RTOL = 1e-12


def length_scale(sch):
    t = np.asarray(sch.table).reshape(sch.n_surf, S.SURF_STRIDE)
    return max(1.0, float(np.max(np.abs(t[:, S.S_DIST:S.S_DIST + 3]))))


def rel_err(a, b, mask, L=1.0):
    """max relative error over rays in `mask` of (7, n) state arrays"""
    if not mask.any():
        return 0.0
    floor = np.array([L, L, L, 1.0, 1.0, 1.0, 0.0])[:, None]
    scale = np.maximum(np.maximum(np.abs(a), np.abs(b)), floor)
    return float(np.max((np.abs(a - b) / scale)[..., mask]))


@pytest.fixture(params=["generated", "generic"])
def kernel(request, hip):
    """run each parity test through the per-scheme generated kernel and the generic one"""
    prev = hip.set_jit(request.param == "generated")
    yield request.param
    hip.set_jit(prev)


def gpu_trace(hip, sch, seed=0, dump=True, flux=False, rays=None, flag=None, expect_kernel=None):
    n = sch.beam.n_rays if rays is None else rays.shape[1]
    ds = hip.DeviceScheme(sch)
    det = sch.detector()
    npix = det.ccd["ydim"] * det.ccd["xdim"] if det is not None else 1      # no CCD frame: nothing is added
    hist = torch.zeros(npix, dtype=torch.int32, device="cuda")
    fl = torch.zeros(npix, dtype=torch.float64, device="cuda") if flux else None
    stats = torch.zeros(S.STATS_LEN, dtype=torch.int64, device="cuda")
    D = DF = None
    if dump:
        D = torch.zeros((sch.n_surf, 7, n), dtype=torch.float64, device="cuda")
        DF = torch.zeros((sch.n_surf, n), dtype=torch.int8, device="cuda")
    flags = S.RT_TRACE_STORE | S.RT_TRACE_HIST | (S.RT_TRACE_FLUX if flux else 0)
    if rays is None:
        R = torch.zeros((7, n), dtype=torch.float64, device="cuda")
        F = torch.zeros(n, dtype=torch.int8, device="cuda")
        flags |= S.RT_TRACE_GENERATE
    else:
        R = torch.from_numpy(np.ascontiguousarray(rays)).cuda()
        F = torch.from_numpy(np.ascontiguousarray(flag)).cuda()
    print(f"[gpu trace] {sch.name} n={n} seed={seed}", flush=True)
    ds.trace(R, F, seed=seed, flags=flags, hist=hist, flux=fl, stats=stats, dump=D, dump_flag=DF)
    torch.cuda.synchronize()
    if expect_kernel:
        assert hip.last_kernel() == expect_kernel, f"ran the {hip.last_kernel()} kernel"
    out = dict(rays=R.cpu().numpy(), flag=F.cpu().numpy(), hist=hist.cpu().numpy().view(np.uint32),
               stats=stats.cpu().numpy().view(np.uint64))
    if dump:
        out["dump"] = D.cpu().numpy()
        out["dump_flag"] = DF.cpu().numpy()
    if flux:
        out["flux"] = fl.cpu().numpy()
    ds.close()
    return out


def compare(sch, g, ref):
    # per-surface states: a failing stage is reported by index
    L = length_scale(sch)
    worst = 0.0
    for s in range(sch.n_surf):
        fg, fr = g["dump_flag"][s], ref.dump_flag[s]
        assert np.array_equal(fg, fr), f"flags differ after surface {s} ({sch.surfaces[s].comment}): " \
                                       f"{int(np.sum(fg != fr))} rays"
        e = rel_err(g["dump"][s], ref.dump[s], fr == 1, L)
        worst = max(worst, e)
        assert e <= RTOL, f"state after surface {s} ({sch.surfaces[s].comment}) differs by {e:.3e}"
    assert np.array_equal(g["flag"], ref.flag)
    assert rel_err(g["rays"], ref.rays, ref.flag == 1, L) <= RTOL
    print(f"[parity] {sch.name}: worst relative error {worst:.2e} (L = {L:.1f} mm)")
    assert np.array_equal(g["hist"], ref.hist), "CCD hit counts differ"
    assert np.array_equal(g["stats"], ref.stats), "statistics differ"


@pytest.mark.parametrize("seed", [0, 7])
def test_czerny_turner_parity(hip, kernel, seed):
    sch = schemes.czerny_turner(n_per_lambda=4000, nlambda=5)
    g = gpu_trace(hip, sch, seed=seed, expect_kernel=kernel)
    ref = O.trace(sch, seed=seed, hist=True, dump=True)
    compare(sch, g, ref)
    assert ref.stats[S.ST_CCD_HITS] > 0.9 * sch.beam.n_rays


@pytest.mark.parametrize("seed", [0, 11])
def test_echelle_parity(hip, kernel, seed):
    sch = schemes.echelle(n_per_lambda=8, nlambda=2048)
    g = gpu_trace(hip, sch, seed=seed, flux=True, expect_kernel=kernel)
    ref = O.trace(sch, seed=seed, hist=True, flux=True, dump=True)
    compare(sch, g, ref)
    # FP64 flux accumulates in a different order: relative tolerance per pixel
    nz = ref.flux != 0
    assert np.array_equal(g["flux"] != 0, nz)
    assert np.max(np.abs(g["flux"][nz] - ref.flux[nz]) / ref.flux[nz]) < 1e-12


def test_loaded_rays_parity(hip, kernel):
    """rays handed in (not generated): the per-surface loop on given input"""
    sch = schemes.echelle(n_per_lambda=4, nlambda=512)
    rays = O.gen_rays(sch.beam.pack(), 3, 0, sch.beam.n_rays)
    flag = np.ones(sch.beam.n_rays, dtype=np.int8)
    flag[::17] = 0
    g = gpu_trace(hip, sch, rays=rays, flag=flag, expect_kernel=kernel)
    ref = O.trace(sch, rays=rays, flag=flag, hist=True, dump=True)
    compare(sch, g, ref)


def test_full_size_echelle_properties(hip):
    """The echelle at its full 10,000,384 rays: properties that need no oracle
    at this size (determinism, shards compose to the whole bit for bit, counts
    conserved), plus oracle parity on three ray blocks of the same launch."""
    sch = schemes.echelle()
    n = sch.beam.n_rays
    ds = hip.DeviceScheme(sch)
    det = sch.detector()
    npix = det.ccd["ydim"] * det.ccd["xdim"]
    flags = S.RT_TRACE_GENERATE | S.RT_TRACE_STORE | S.RT_TRACE_HIST | S.RT_TRACE_FLUX

    def run(begin, count, seed=21):
        R = torch.zeros((7, count), dtype=torch.float64, device="cuda")
        F = torch.zeros(count, dtype=torch.int8, device="cuda")
        H = torch.zeros(npix, dtype=torch.int32, device="cuda")
        X = torch.zeros(npix, dtype=torch.float64, device="cuda")
        st = torch.zeros(S.STATS_LEN, dtype=torch.int64, device="cuda")
        print(f"[full size] begin={begin} n={count}", flush=True)
        ds.trace(R, F, seed=seed, begin=begin, n=count, flags=flags, hist=H, flux=X, stats=st)
        torch.cuda.synchronize()
        return R, F, H, X, st.cpu().numpy().view(np.uint64)

    R1, F1, H1, X1, st1 = run(0, n)
    R2, F2, H2, X2, st2 = run(0, n)
    assert torch.equal(R1, R2) and torch.equal(F1, F2) and torch.equal(H1, H2), "trace is not deterministic"
    assert np.array_equal(st1, st2)
    del R2, F2, H2, X2
    h = n // 2 + 123                    # shard boundary inside a wave and inside a wavelength group
    Ra, Fa, Ha, Xa, sta = run(0, h)
    Rb, Fb, Hb, Xb, stb = run(h, n - h)
    assert torch.equal(torch.cat([Ra, Rb], dim=1), R1), "shards differ from the whole"
    assert torch.equal(torch.cat([Fa, Fb]), F1)
    assert torch.equal(Ha + Hb, H1), "shard hit counts do not add up"
    assert np.array_equal(sta + stb, st1)
    x1 = X1.cpu().numpy()
    xs = (Xa + Xb).cpu().numpy()
    assert np.all(np.abs(xs - x1) <= 1e-12 * np.abs(x1)), "shard flux does not add up"
    del Ra, Rb, Fa, Fb, Ha, Hb, Xa, Xb
    # conservation
    good = int(F1.sum().item())
    assert int(st1[S.ST_RAYS]) == n and int(st1[S.ST_GOOD]) == good
    assert int(H1.sum().item()) == int(st1[S.ST_CCD_HITS])
    assert int(st1[S.ST_CCD_HITS]) + int(st1[S.ST_CCD_OFFCHIP]) == good      # the CCD is the last surface
    for s in range(sch.n_surf - 1):
        entering, lost_i, lost_a = (int(st1[S.ST_SURF + 4 * s + k]) for k in (0, 2, 3))
        nxt = int(st1[S.ST_SURF + 4 * (s + 1)])
        assert nxt == entering - lost_i - lost_a, f"ray count not conserved after surface {s}"
    assert good > 0.95 * n
    # oracle parity on three blocks of the full launch
    Lsc = length_scale(sch)
    for b0 in (0, n // 2 - 1000, n - 2048):
        ref = O.trace(sch, seed=21, begin=b0, n=2048)
        g = R1[:, b0:b0 + 2048].cpu().numpy()
        f = F1[b0:b0 + 2048].cpu().numpy()
        assert np.array_equal(f, ref.flag)
        assert rel_err(g, ref.rays, ref.flag == 1, Lsc) <= RTOL
    ds.close()


def test_launch_configurations_agree_and_autotune(hip):
    """Every launch of the generated kernel computes the same rays: the launch
    (lanes per workgroup, waves per SIMD, grid size) changes only the schedule."""
    sch = schemes.echelle(n_per_lambda=8, nlambda=512)
    base = gpu_trace(hip, sch, seed=4, dump=False, expect_kernel="generated")
    for block, waves, mult in [(128, 5, 1), (256, 6, 2), (512, 0, 1), (64, 0, 4), (256, 7, 8), (256, 8, 16)]:
        ds = hip.DeviceScheme(sch)
        ds.set_launch(block, waves, mult)
        assert ds.get_launch() == (block, waves, mult)
        n = sch.beam.n_rays
        R = torch.zeros((7, n), dtype=torch.float64, device="cuda")
        F = torch.zeros(n, dtype=torch.int8, device="cuda")
        det = sch.detector()
        H = torch.zeros(det.ccd["ydim"] * det.ccd["xdim"], dtype=torch.int32, device="cuda")
        st = torch.zeros(S.STATS_LEN, dtype=torch.int64, device="cuda")
        print(f"[launch] block={block} waves={waves} grid_mult={mult}", flush=True)
        ds.trace(R, F, seed=4, flags=S.RT_TRACE_GENERATE | S.RT_TRACE_STORE | S.RT_TRACE_HIST, hist=H, stats=st)
        torch.cuda.synchronize()
        assert hip.last_kernel() == "generated"
        assert np.array_equal(R.cpu().numpy(), base["rays"]), (block, waves, mult)
        assert np.array_equal(F.cpu().numpy(), base["flag"])
        assert np.array_equal(H.cpu().numpy().view(np.uint32), base["hist"])
        assert np.array_equal(st.cpu().numpy().view(np.uint64), base["stats"])
        ds.close()
    # surface unroll depth 0: the runtime loop of the generic kernel; flags,
    # hits and counters as the unrolled kernel, rays within the parity bar
    ds = hip.DeviceScheme(sch)
    ds.set_unroll(False)
    assert ds.get_unroll() is False
    n = sch.beam.n_rays
    R = torch.zeros((7, n), dtype=torch.float64, device="cuda")
    F = torch.zeros(n, dtype=torch.int8, device="cuda")
    det = sch.detector()
    H = torch.zeros(det.ccd["ydim"] * det.ccd["xdim"], dtype=torch.int32, device="cuda")
    st = torch.zeros(S.STATS_LEN, dtype=torch.int64, device="cuda")
    print("[launch] surface loop", flush=True)
    ds.trace(R, F, seed=4, flags=S.RT_TRACE_GENERATE | S.RT_TRACE_STORE | S.RT_TRACE_HIST, hist=H, stats=st)
    torch.cuda.synchronize()
    assert hip.last_kernel() == "generic"
    assert np.array_equal(F.cpu().numpy(), base["flag"])
    assert rel_err(R.cpu().numpy(), base["rays"], base["flag"] == 1, length_scale(sch)) <= 2 * RTOL
    assert np.array_equal(H.cpu().numpy().view(np.uint32), base["hist"])
    assert np.array_equal(st.cpu().numpy().view(np.uint64), base["stats"])
    ds.close()
    ds = hip.DeviceScheme(sch)
    print("[launch] autotune", flush=True)
    ms = ds.autotune(n_sample=200000, flags=S.RT_TRACE_STORE | S.RT_TRACE_HIST | S.RT_TRACE_FLUX)
    b, w, m = ds.get_launch()
    assert ms > 0 and b in (128, 256) and w in (0, 5, 6, 7, 8) and m in (1, 2, 4, 8, 16)
    assert ds.get_unroll() in (True, False)
    ds.close()


def test_clustered_ccd_hits_are_counted_exactly(hip, kernel):
    """A 0.05 mm beam puts each wavelength's rays on one or two pixels, so every
    wave's hits merge before the atomics (the pre-aggregation path); the
    frame must still equal the oracle's count for count."""
    sch = schemes.czerny_turner(n_per_lambda=4000, nlambda=5)
    sch.beam.params = (4000, 0.0, 0.05)
    g = gpu_trace(hip, sch, seed=2, dump=False, expect_kernel=kernel)
    ref = O.trace(sch, seed=2, hist=True)
    assert np.array_equal(g["flag"], ref.flag)
    assert np.array_equal(g["hist"], ref.hist), "CCD hit counts differ"
    assert np.array_equal(g["stats"], ref.stats)
    nz = np.count_nonzero(ref.hist)
    assert 0 < nz <= 40 and ref.hist.sum() > 0.9 * sch.beam.n_rays, nz


def _echelle_mono():
    sch = schemes.echelle(2000, 8)
    sch.beam.range_distr = S.RT_RANGE_MONO
    return sch


def _echelle_oob():
    # 0.30-0.90 um: the glass tables start at 0.35 um, so the bluest grid
    # wavelengths are out of range at the first lens
    sch = schemes.echelle(600, 64)
    sch.beam.range = (0.30, 0.90)
    return sch


@pytest.mark.parametrize("make", [lambda: schemes.echelle(600, 64), _echelle_mono, _echelle_oob],
                         ids=["grid", "mono", "out-of-range"])
def test_wavelength_table_matches_per_ray_splines(hip, make):
    """Generated traces of grid-wavelength beams read their refractive indices
    and detector QE from the per-wavelength table (filled by the same spline
    code at the same wavelengths): rays, flags, hit counts and counters equal
    those of the per-ray spline evaluation bit for bit, the flux frame to
    1e-12 (FP64 atomics add in any order)."""
    sch = make()
    prev_jit = hip.set_jit(True)
    prev = hip.set_ltab(False)
    try:
        off = gpu_trace(hip, sch, seed=3, dump=False, flux=True, expect_kernel="generated")
    finally:
        hip.set_ltab(prev)
    try:
        on = gpu_trace(hip, sch, seed=3, dump=False, flux=True, expect_kernel="generated")
    finally:
        hip.set_jit(prev_jit)
    assert np.array_equal(on["rays"].view(np.uint64), off["rays"].view(np.uint64))
    assert np.array_equal(on["flag"], off["flag"])
    assert np.array_equal(on["hist"], off["hist"])
    assert np.array_equal(on["stats"], off["stats"])
    scale = max(float(np.max(np.abs(off["flux"]))), 1e-300)
    assert float(np.max(np.abs(on["flux"] - off["flux"]))) <= 1e-12 * scale
    assert int(on["stats"][S.ST_RAYS]) == sch.beam.n_rays
    if make is _echelle_oob:
        assert int(on["stats"][S.ST_SPLINE_OOB]) > 0


def test_derived_table_follows_beam_stream_and_graph(hip):
    """The derived per-wavelength table of a scheme is refilled when the
    wavelength grid or the stream changes, and on every launch once a launch
    has been captured in a graph: traces of two beams alternated on one
    scheme, on two streams, and a captured trace replayed after an eager trace
    of the other beam each equal the trace of a fresh scheme that has seen
    only that beam (bit for bit: same kernel, same table values)."""
    import dataclasses
    sch = schemes.echelle(300, 64)
    beam_a = sch.beam
    beam_b = dataclasses.replace(beam_a, range=(beam_a.range[0] + 0.05, beam_a.range[1] + 0.02))
    n = beam_a.n_rays
    flags = S.RT_TRACE_GENERATE | S.RT_TRACE_STORE

    def run(ds, beam, stream=None):
        R = torch.zeros((7, n), dtype=torch.float64, device="cuda")
        F = torch.zeros(n, dtype=torch.int8, device="cuda")
        ds.trace(R, F, seed=5, flags=flags, beam=beam, stream=stream)
        return R, F

    def fresh(beam):
        ds = hip.DeviceScheme(sch)
        R, F = run(ds, beam)
        torch.cuda.synchronize()
        assert hip.last_kernel() == "generated"
        out = R.cpu().numpy().view(np.uint64).copy(), F.cpu().numpy().copy()
        ds.close()
        return out

    ref = {"a": fresh(beam_a), "b": fresh(beam_b)}
    assert not np.array_equal(ref["a"][0][6], ref["b"][0][6])

    def same(R, F, key):
        torch.cuda.synchronize()
        assert np.array_equal(R.cpu().numpy().view(np.uint64), ref[key][0]), key
        assert np.array_equal(F.cpu().numpy(), ref[key][1]), key

    ds = hip.DeviceScheme(sch)
    print("[derived table] alternating beams", flush=True)
    for key in ("a", "b", "a", "a", "b"):
        same(*run(ds, beam_a if key == "a" else beam_b), key)
    print("[derived table] second stream", flush=True)
    s2 = torch.cuda.Stream()
    with torch.cuda.stream(s2):
        R, F = run(ds, beam_b, stream=s2.cuda_stream)
    s2.synchronize()
    same(R, F, "b")
    print("[derived table] graph replays between eager traces of the other beam", flush=True)
    Rg = torch.zeros((7, n), dtype=torch.float64, device="cuda")
    Fg = torch.zeros(n, dtype=torch.int8, device="cuda")
    g = torch.cuda.CUDAGraph()
    cs = torch.cuda.Stream()
    torch.cuda.synchronize()
    with torch.cuda.graph(g, stream=cs):
        ds.trace(Rg, Fg, seed=5, flags=flags, beam=beam_a, stream=cs.cuda_stream)
    for _ in range(3):
        same(*run(ds, beam_b), "b")
        Rg.zero_()
        g.replay()
        same(Rg, Fg, "a")
    same(*run(ds, beam_b), "b")
    ds.close()
