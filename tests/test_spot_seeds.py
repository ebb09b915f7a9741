"""Adjoint seeds of functions of the group spots (rt_amd.diff.spot_seeds) and
the linear dispersion's derivatives, on the host.

spot_seeds turns dF/dcentroid and dF/drms per group into per-ray derivatives
dF/dX_r = ax + b X_r, dF/dY_r = ay + b Y_r; checked against central finite
differences of F in every ray coordinate of random groups (F = sum of
radii + centroid terms + the dispersion), bar 1e-7 relative to the largest
derivative.
"""
import numpy as np

from rt_amd import diff


def spots(x, y, g, G):
    cnt = np.zeros(G); cx = np.zeros(G); cy = np.zeros(G); rms = np.zeros(G)
    for k in range(G):
        m = g == k
        cnt[k] = m.sum()
        if cnt[k]:
            cx[k], cy[k] = x[m].mean(), y[m].mean()
            rms[k] = np.sqrt(np.mean((x[m] - cx[k]) ** 2 + (y[m] - cy[k]) ** 2))
    return cnt, cx, cy, rms


def test_spot_seeds_match_finite_differences():
    rng = np.random.default_rng(5)
    G, n = 6, 90
    x, y = rng.normal(size=n), rng.normal(size=n) * 0.5
    g = rng.integers(0, G, n)
    g[g == 4] = 5                                  # group 4 empty
    lam = np.linspace(0.4, 0.9, G)
    disp = diff.linear_dispersion(lam, axis=0)
    wx, wy = rng.normal(size=G), rng.normal(size=G)

    def F(x, y):
        cnt, cx, cy, rms = spots(x, y, g, G)
        D = disp(cnt, cx, cy, rms)[0]
        return float(np.sum(rms) + np.sum(wx * cx ** 2) + np.sum(wy * cy) + 3.0 * D)

    cnt, cx, cy, rms = spots(x, y, g, G)
    _, dD, _, _ = disp(cnt, cx, cy, rms)
    seeds = diff.spot_seeds(cnt, cx, cy, rms, 2 * wx * cx + 3.0 * dD, wy, np.ones(G))
    assert np.all(seeds[4] == 0.0)
    ax, ay, b = seeds[g, 0], seeds[g, 1], seeds[g, 2]
    dx, dy = ax + b * x, ay + b * y
    h = 1e-6
    fdx = np.array([(F(x + h * e, y) - F(x - h * e, y)) / (2 * h) for e in np.eye(n)])
    fdy = np.array([(F(x, y + h * e) - F(x, y - h * e)) / (2 * h) for e in np.eye(n)])
    scale = max(np.abs(fdx).max(), np.abs(fdy).max())
    assert np.abs(dx - fdx).max() <= 1e-7 * scale
    assert np.abs(dy - fdy).max() <= 1e-7 * scale


def test_linear_dispersion_is_the_least_squares_slope():
    rng = np.random.default_rng(2)
    lam = np.linspace(0.38, 0.9, 9)
    cx = 12.0 * lam + rng.normal(size=9) * 1e-3
    cnt = np.full(9, 10.0)
    cnt[2] = 0.0
    D = diff.linear_dispersion(lam, axis=0)(cnt, cx, np.zeros(9), np.ones(9))[0]
    live = cnt > 0
    assert abs(D - np.polyfit(lam[live], cx[live], 1)[0]) <= 1e-12 * abs(D)
