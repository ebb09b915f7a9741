"""Every optical branch of the engine on the GPU against the oracle.

* branches() (tests/branch_schemes.py): conic and toric gratings, reflective
  and transmissive (the four curved-grating kernels of
  rt_cuda_engine_surface.cu:1320-1520), circular and rectangular Shades
  (surface.cpp:652-699), plane-lens refraction (surface.cu:956-989), a plane
  mirror, total internal reflection at a tilted exit face and evanescent
  grating orders, with random wavelengths: the fused trace (generated and
  generic kernels) surface by surface, and the per-stage kernels of the
  plugin API along the whole scheme.  Bars as in test_gpu_trace.py: flags,
  lost-ray counts, hit counts and counters bit-exact, states within 1e-12.
* newton_giveups(): rays that make the Newton intersection give up, on
  |dF| < eps (a ray tangent to the surface) and at the 1e5 iteration cap (a
  line with no root on the surface), beside ordinary rays: flags and lost-ray
  counts bit-exact through the fused trace (both kernels) and the
  per-stage intersection.  The capped ray's iterates wander chaotically, so
  the GPU's and the oracle's rounding lead them apart: its iteration count is
  not compared beyond "gave up" (between 1 and the cap).
"""
import numpy as np
import pytest

from rt_amd import scheme as S
from oracle import oracle as O

import branch_schemes as B
from test_gpu_stages import run_chain
from test_gpu_trace import compare, gpu_trace, kernel  # noqa: F401  (fixture)

pytestmark = pytest.mark.gpu

CAP = 100001


def _lost_counts(st, n_surf):
    return [(int(st[S.ST_SURF + 4 * s]), int(st[S.ST_SURF + 4 * s + 2]), int(st[S.ST_SURF + 4 * s + 3]))
            for s in range(n_surf)]


@pytest.mark.parametrize("seed", [1, 8])
def test_branches_fused_trace(hip, kernel, seed):
    sch = B.branches(20000)
    g = gpu_trace(hip, sch, seed=seed, flux=True, expect_kernel=kernel)
    ref = O.trace(sch, seed=seed, hist=True, flux=True, dump=True)
    compare(sch, g, ref)
    lost = _lost_counts(ref.stats, sch.n_surf)
    # every loss path is taken: TIR at the 40 deg exit face, evanescent orders
    # at the order-2 grating, both shades
    assert lost[0][2] > 0 and lost[3][2] > 0 and lost[4][2] > 0 and lost[9][2] > 0, lost
    assert int(ref.stats[S.ST_GOOD]) > 0.2 * sch.beam.n_rays
    nz = ref.flux != 0
    assert np.array_equal(g["flux"] != 0, nz)
    assert np.max(np.abs(g["flux"][nz] - ref.flux[nz]) / ref.flux[nz]) < 1e-12


def test_branches_stage_chain(hip):
    sch = B.branches(4000)
    state, alive = run_chain(hip, sch, 3)
    fused = O.trace(sch, seed=3)
    assert np.array_equal(alive, fused.flag == 1), "stage chain and fused trace lose different rays"
    assert np.array_equal(state[:, alive], fused.rays[:, alive])
    assert 0 < alive.sum() < len(alive)


def test_newton_giveups_fused(hip, kernel):
    sch, R, F = B.newton_giveups()
    g = gpu_trace(hip, sch, rays=R, flag=F, expect_kernel=kernel)
    it = np.zeros((1, R.shape[1]), dtype=np.uint32)
    ref = O.trace(sch, rays=R, flag=F, hist=True, dump=True, iters=it, iters_mode=1)
    assert np.array_equal(g["flag"], ref.flag), (g["flag"], ref.flag)
    assert list(ref.flag[-2:]) == [0, 0] and np.all(ref.flag[:-2] == 1)
    assert it[0, -2] == 0 and it[0, -1] == CAP, it          # |dF| < eps at once; the iteration cap
    ok = ref.flag == 1
    L = 50.0
    scale = np.maximum(np.maximum(np.abs(g["rays"][:, ok]), np.abs(ref.rays[:, ok])), L)
    assert float(np.max(np.abs(g["rays"][:, ok] - ref.rays[:, ok]) / scale)) <= 1e-12
    gs, rs = g["stats"], ref.stats
    for k in range(S.STATS_LEN):
        if k != S.ST_SURF + 1:
            assert gs[k] == rs[k], (k, gs[k], rs[k])
    ordinary = int(it[0, :-1].sum())
    giter = int(gs[S.ST_SURF + 1]) - ordinary
    print(f"[newton give-ups] capped ray: oracle {CAP} iterations, GPU {giter}", flush=True)
    assert 1 <= giter <= CAP


def test_newton_giveups_stage(hip):
    sch, R, F = B.newton_giveups()
    state, alive = run_chain(hip, sch, 0, rays=R, flag=F, iters_exact=False)
    assert list(alive) == [True] * 4 + [False, False]
