"""Batched traces (rt_hip_batch_*): many tolerance instances of the echelle
with uneven ray counts in one persistent launch, checked against the oracle
instance by instance.

Bars (written here): every ray of instance i equals the oracle's trace of that
instance alone with seed + i, positions within 1e-12 relative with the length
scale as floor, cosines 1e-12, flags bit-exact; statistics counters summed over
instances bit-exact; per-instance spot (count exact, centroid 1e-12 relative to
the length scale, variances and RMS radius 1e-9 relative: the oracle side sums
in another order); two runs bit-identical; every graph replay bit-identical to
the eager trace of the same tables.
"""
import numpy as np
import pytest
import torch

from rt_amd import scheme as S
from rt_amd import schemes
from oracle import oracle as O

pytestmark = pytest.mark.gpu

RTOL = 1e-12


@pytest.fixture(params=["generated", "generic"])
def kernel(request, hip):
    prev = hip.set_jit(request.param == "generated")
    yield request.param
    hip.set_jit(prev)


def make_batch(n_inst=7, seed=5):
    base = schemes.echelle(n_per_lambda=4, nlambda=64)
    inst, counts = schemes.tolerance_instances(base, n_inst, seed=seed, min_rays=40, max_rays=3000)
    counts[1] = 1              # single-ray instance
    counts[2] = 256            # exactly one workgroup tile
    return inst, counts


def oracle_spot(rays, flag):
    ok = flag == 1
    x, y = rays[0, ok], rays[1, ok]
    n = int(ok.sum())
    if n == 0:
        return np.zeros(6)
    mx, my = x.mean(), y.mean()
    vx, vy = np.mean((x - mx) ** 2), np.mean((y - my) ** 2)
    return np.array([n, mx, my, vx, vy, np.sqrt(vx + vy)])


def run_batch(hip, inst, counts, seed, store=True):
    db = hip.DeviceBatch(inst, counts)
    total = int(sum(counts))
    R = torch.zeros((7, total), dtype=torch.float64, device="cuda") if store else None
    F = torch.zeros(total, dtype=torch.int8, device="cuda") if store else None
    spots = torch.zeros((len(inst), hip.SPOT_STRIDE), dtype=torch.float64, device="cuda")
    st = torch.zeros(S.STATS_LEN, dtype=torch.int64, device="cuda")
    print(f"[batch] n_inst={len(inst)} rays={total} seed={seed}", flush=True)
    db.trace(seed=seed, flags=S.RT_TRACE_STORE if store else 0, rays=R, flag=F, spots=spots, stats=st)
    torch.cuda.synchronize()
    out = dict(spots=spots.cpu().numpy(), stats=st.cpu().numpy().view(np.uint64))
    if store:
        out["rays"] = R.cpu().numpy()
        out["flag"] = F.cpu().numpy()
    return db, out


def test_batch_matches_oracle_per_instance(hip, kernel):
    inst, counts = make_batch()
    seed = 40
    db, g = run_batch(hip, inst, counts, seed)
    assert hip.last_kernel() == kernel
    t = np.asarray(inst[0].table).reshape(inst[0].n_surf, S.SURF_STRIDE)
    L = max(1.0, float(np.max(np.abs(t[:, S.S_DIST:S.S_DIST + 3]))))
    floor = np.array([L, L, L, 1.0, 1.0, 1.0, 0.0])[:, None]
    off = 0
    stats = np.zeros(S.STATS_LEN, dtype=np.uint64)
    for i, (sch, n) in enumerate(zip(inst, counts)):
        ref = O.trace(sch, seed=seed + i, n=n)
        stats += ref.stats
        gr, gf = g["rays"][:, off:off + n], g["flag"][off:off + n]
        assert np.array_equal(gf, ref.flag), f"instance {i}: flags differ"
        ok = ref.flag == 1
        if ok.any():
            scale = np.maximum(np.maximum(np.abs(gr), np.abs(ref.rays)), floor)
            e = float(np.max((np.abs(gr - ref.rays) / scale)[:, ok]))
            assert e <= RTOL, f"instance {i}: rays differ by {e:.3e}"
        sp, rs = g["spots"][i], oracle_spot(ref.rays, ref.flag)
        assert sp[0] == rs[0], f"instance {i}: spot count"
        if rs[0] > 0:
            assert abs(sp[1] - rs[1]) <= RTOL * L and abs(sp[2] - rs[2]) <= RTOL * L, f"instance {i}: centroid"
            assert np.all(np.abs(sp[3:] - rs[3:]) <= 1e-9 * np.maximum(np.abs(rs[3:]), 1e-300)), i
        off += n
    assert np.array_equal(g["stats"], stats), "statistics differ"
    # bit-reproducible
    db2, g2 = run_batch(hip, inst, counts, seed, store=False)
    assert np.array_equal(g2["spots"], g["spots"]) and np.array_equal(g2["stats"], g["stats"])
    db.close()
    db2.close()


def test_batch_launch_configurations_agree(hip):
    inst, counts = make_batch(5, seed=9)
    db, base = run_batch(hip, inst, counts, 3)
    for block, waves, mult in [(128, 5, 1), (512, 0, 2), (64, 6, 1)]:
        db.set_launch(block, waves, mult)
        total = int(sum(counts))
        R = torch.zeros((7, total), dtype=torch.float64, device="cuda")
        F = torch.zeros(total, dtype=torch.int8, device="cuda")
        spots = torch.zeros((len(inst), hip.SPOT_STRIDE), dtype=torch.float64, device="cuda")
        print(f"[batch launch] {block} {waves} {mult}", flush=True)
        db.trace(seed=3, flags=S.RT_TRACE_STORE, rays=R, flag=F, spots=spots)
        torch.cuda.synchronize()
        assert np.array_equal(R.cpu().numpy(), base["rays"]) and np.array_equal(F.cpu().numpy(), base["flag"])
        s = spots.cpu().numpy()
        assert np.array_equal(s[:, 0], base["spots"][:, 0])
        # tiles differ with the workgroup size: the moments combine in another grouping
        assert np.all(np.abs(s[:, 1:] - base["spots"][:, 1:]) <= 1e-9 * np.maximum(np.abs(base["spots"][:, 1:]), 1e-12))
    db.close()


def test_batch_graph_replays_match_eager(hip):
    inst, counts = make_batch(6, seed=11)
    inst2, _ = schemes.tolerance_instances(schemes.echelle(n_per_lambda=4, nlambda=64), 6, seed=12,
                                           min_rays=40, max_rays=3000)
    db = hip.DeviceBatch(inst, counts)
    spots = torch.zeros((len(inst), hip.SPOT_STRIDE), dtype=torch.float64, device="cuda")
    stream = torch.cuda.Stream()
    g = db.graph(stream=stream, seed=77, spots=spots)
    ref_spots = torch.zeros_like(spots)
    for k, tables in enumerate([inst, inst2, inst, inst2]):
        db.update(tables)
        print(f"[graph] replay {k}", flush=True)
        g.launch(stream)
        stream.synchronize()
        got = spots.cpu().numpy().copy()
        db.trace(seed=77, spots=ref_spots)          # eager, same tables, on the current stream
        torch.cuda.synchronize()
        assert np.array_equal(got, ref_spots.cpu().numpy()), f"replay {k} differs from the eager trace"
        spots.fill_(-1.0)                            # eager work between replays
        torch.cuda.synchronize()
    g.close()
    db.close()


def test_batch_with_a_beam_per_instance(hip):
    """rt_hip_batch_set_beams: every instance its own wavelength set (range and
    number of wavelengths) in one launch; instance i equals the oracle's trace
    of that instance with its own beam and seed + i (same bars)."""
    import copy
    inst, counts = make_batch(n_inst=5, seed=8)
    beams = []
    for i, sch in enumerate(inst):
        b = copy.deepcopy(sch.beam)
        b.range = (0.40 + 0.03 * i, 0.62 + 0.05 * i)
        b.nlambda = (8, 16, 32, 64, 5)[i]
        beams.append(b)
    seed = 12
    db = hip.DeviceBatch(inst, counts)
    db.set_beams(beams)
    total = int(sum(counts))
    R = torch.zeros((7, total), dtype=torch.float64, device="cuda")
    F = torch.zeros(total, dtype=torch.int8, device="cuda")
    print(f"[batch beams] n_inst={len(inst)} rays={total}", flush=True)
    db.trace(seed=seed, flags=S.RT_TRACE_STORE, rays=R, flag=F)
    torch.cuda.synchronize()
    rays, flag = R.cpu().numpy(), F.cpu().numpy()
    t = np.asarray(inst[0].table).reshape(inst[0].n_surf, S.SURF_STRIDE)
    L = max(1.0, float(np.max(np.abs(t[:, S.S_DIST:S.S_DIST + 3]))))
    floor = np.array([L, L, L, 1.0, 1.0, 1.0, 0.0])[:, None]
    off = 0
    for i, (sch, n) in enumerate(zip(inst, counts)):
        own = copy.deepcopy(sch)
        own.beam = beams[i]
        ref = O.trace(own, seed=seed + i, n=n)
        gr, gf = rays[:, off:off + n], flag[off:off + n]
        assert np.array_equal(gr[6], ref.rays[6]), f"instance {i}: wavelengths differ"
        assert np.array_equal(gf, ref.flag), f"instance {i}: flags differ"
        ok = ref.flag == 1
        if ok.any():
            scale = np.maximum(np.maximum(np.abs(gr), np.abs(ref.rays)), floor)
            assert float(np.max((np.abs(gr - ref.rays) / scale)[:, ok])) <= RTOL, i
        off += n
    # back to one beam: the first instance's for all
    db.set_beams(None)
    db.trace(seed=seed, flags=S.RT_TRACE_STORE, rays=R, flag=F)
    torch.cuda.synchronize()
    ref0 = O.trace(inst[1], seed=seed + 1, n=counts[1])
    assert np.array_equal(R.cpu().numpy()[6, counts[0]:counts[0] + counts[1]], ref0.rays[6])
    db.close()
