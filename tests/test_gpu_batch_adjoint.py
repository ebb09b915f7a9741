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
