"""The occlusion bounds of checkpoint frames, checked on the CPU against brute force.

A checkpoint frame's span pass gives every (triangle, tile) entry a lower bound of
its fragments' depth order and lets a triangle that covers a whole tile raise the
tile's occluder bound; the raster skips entries whose lower bound lies above the
least occluder bound (kernels.hip, tile_bounds and big_tasks).  That is exact only
if the bounds hold for the values the raster really computes: weights reached from
the row start x0 by one f32 addition per pixel (triangles.rs:257-301), then
u2 = (w2 - w0) - w1 and z = (hz0 w0 + hz1 w1) + hz2 u2.

This test restates tile_bounds' f32 arithmetic (the same operations in the same
order) and walks thousands of random weight planes in f32 from random row starts
left of random tiles, then checks every pixel of every tile:
  * each walked z lies in [zlo, zhi];
  * where the bounds claim the triangle covers the tile, each walked w0, w1, u2 is
    >= 0 (every pixel is a fragment).
It also checks that it can fail: with the margins taken out, the same walks break
both claims.  The planes come in three kinds: well inside the tile, with a weight
that reaches zero within a few ulps of a corner, and crossing the tile.
"""
import numpy as np

F = np.float32
W, H, TW, TH = 640, 480, 32, 8


def bounds(o, dx, dy, hz, xmin, xmax, ymax, X0, Y0, margins=True):
    """tile_bounds' arithmetic in f32 (kernels.hip), vectorised over cases:
    returns zlo, zhi and the full-cover claim (no z clip)."""
    Xm = xmax.astype(F) + F(32.0)
    Ym = ymax.astype(F)
    K = (xmax - xmin + 8).astype(F) * F(2.0 ** -22)
    M = [(np.abs(o[i]) + np.abs(dy[i]) * Ym + np.abs(dx[i]) * Xm) * F(1.001) for i in range(3)]
    e = [K * M[i] + F(2.0 ** -125) for i in range(3)]
    S = M[0] + M[1] + M[2]
    eu = (e[0] + e[1] + e[2]) + F(2.0 ** -21) * S
    h = [np.abs(hz[i]) for i in range(3)]
    P = h[0] * M[0] + h[1] * M[1] + h[2] * S
    ez = F(2.0) * ((h[0] * e[0] + h[1] * e[1] + h[2] * eu) + F(2.0 ** -18) * P) + F(2.0 ** -124)
    xs = [X0.astype(F), (X0 + TW - 1).astype(F)]
    ys = [Y0.astype(F), (Y0 + TH - 1).astype(F)]
    zmin = np.full(o[0].shape, np.inf, F)
    zmax = np.full(o[0].shape, -np.inf, F)
    wmin = [np.full(o[0].shape, np.inf, F) for _ in range(3)]
    for a in range(2):
        for c in range(2):
            w = [(o[i] + dy[i] * ys[c]) + dx[i] * xs[a] for i in range(3)]
            u2 = (w[2] - w[0]) - w[1]
            z = (hz[0] * w[0] + hz[1] * w[1]) + hz[2] * u2
            zmin, zmax = np.fmin(zmin, z), np.fmax(zmax, z)
            wmin = [np.fmin(wmin[0], w[0]), np.fmin(wmin[1], w[1]), np.fmin(wmin[2], u2)]
    if not margins:
        e, eu, ez = [F(0.0)] * 3, F(0.0), F(0.0)
        M, S = [F(0.0)] * 3, F(0.0)
    finite = np.isfinite(zmin) & np.isfinite(zmax) & np.isfinite(ez)
    zlo, zhi = zmin - ez, zmax + ez
    full = (finite & (wmin[0] > e[0] + F(2.0 ** -21) * M[0]) & (wmin[1] > e[1] + F(2.0 ** -21) * M[1])
            & (wmin[2] > eu + F(2.0 ** -20) * S))
    return zlo, zhi, finite, full


def cases(seed, n):
    """n random (triangle box, tile, weight planes, depth coefficients)."""
    rng = np.random.default_rng(seed)
    # tiles anywhere right of the box's left edge: walks of up to W pixels
    xmin = rng.integers(0, W - TW, n)
    first, last = (xmin + TW - 1) // TW, W // TW - 1
    X0 = (first + (rng.random(n) * (last - first + 1)).astype(np.int64)) * TW
    xmin = np.minimum(xmin, X0)
    xmax = np.minimum(W, X0 + TW + rng.integers(0, W, n))
    Y0 = rng.integers(0, H // TH, n) * TH
    ymax = np.minimum(H, Y0 + TH + rng.integers(0, H, n))
    kind = rng.integers(0, 3, n)  # 0 inside, 1 a weight near zero at a corner, 2 crossing
    mag = 10.0 ** rng.uniform(-3, 2, (3, n))
    dx = (rng.choice([-1.0, 1.0], (3, n)) * mag * 10.0 ** rng.uniform(-4, -1, (3, n)))
    dy = (rng.choice([-1.0, 1.0], (3, n)) * mag * 10.0 ** rng.uniform(-4, -1, (3, n)))
    # the planes' values at the tile's corners: w = a + dx (x - X0) + dy (y - Y0)
    span = np.abs(dx) * (TW - 1) + np.abs(dy) * (TH - 1)
    lo_corner = -(np.minimum(dx, 0) * (TW - 1) + np.minimum(dy, 0) * (TH - 1))  # a that puts the minimum at 0
    a = np.where(kind == 0, lo_corner + mag * rng.uniform(0.2, 2.0, (3, n)),
                 np.where(kind == 1, lo_corner + span * rng.uniform(-1e-5, 1e-5, (3, n))
                          + mag * rng.uniform(-1e-6, 1e-6, (3, n)),
                          lo_corner - span * rng.uniform(0.1, 0.9, (3, n))))
    # u2 = w2 - w0 - w1: w2's plane is w0 + w1 + a third edge's plane
    a[2] = a[0] + a[1] + a[2]
    dx[2] = dx[0] + dx[1] + dx[2]
    dy[2] = dy[0] + dy[1] + dy[2]
    o = a - dx * X0 - dy * Y0
    hz = rng.choice([-1.0, 1.0], (3, n)) * 10.0 ** rng.uniform(-3, 1, (3, n))
    f32 = lambda v: [np.asarray(v[i], dtype=F) for i in range(3)]
    return f32(o), f32(dx), f32(dy), f32(hz), xmin, xmax, ymax, X0, Y0, rng


def walk(o, dx, dy, hz, xmin, X0, Y0, rng):
    """Every tile pixel's w0, w1, u2 and z as the raster computes them: each row
    walked from a random start x0 in [xmin, X0] (triangles.rs:257-301)."""
    n = X0.size
    out_w = np.empty((3, n, TH, TW), F)
    out_u = np.empty((n, TH, TW), F)
    out_z = np.empty((n, TH, TW), F)
    for r in range(TH):
        x0 = xmin + (rng.random(n) * (X0 - xmin + 1)).astype(np.int64)
        Y = (Y0 + r).astype(F)
        X = x0.astype(F)
        w = [(o[i] + dy[i] * Y) + dx[i] * X for i in range(3)]
        steps = X0 + TW - x0
        for k in range(int(steps.max())):
            x = x0 + k
            at = (x >= X0) & (x < X0 + TW)
            if at.any():
                idx = np.nonzero(at)[0]
                col = (x - X0)[idx]
                u2 = (w[2][idx] - w[0][idx]) - w[1][idx]
                z = (hz[0][idx] * w[0][idx] + hz[1][idx] * w[1][idx]) + hz[2][idx] * u2
                for i in range(3):
                    out_w[i, idx, r, col] = w[i][idx]
                out_u[idx, r, col] = u2
                out_z[idx, r, col] = z
            w = [w[i] + dx[i] for i in range(3)]
    return out_w, out_u, out_z


def check(seed, n, margins=True):
    o, dx, dy, hz, xmin, xmax, ymax, X0, Y0, rng = cases(seed, n)
    zlo, zhi, finite, full = bounds(o, dx, dy, hz, xmin, xmax, ymax, X0, Y0, margins)
    w, u2, z = walk(o, dx, dy, hz, xmin, X0, Y0, rng)
    zf = z.reshape(n, -1)
    ok_z = np.all((zf >= zlo[:, None]) & (zf <= zhi[:, None]) | ~np.isfinite(zf), axis=1) | ~finite
    covered = np.all((w[0] >= 0) & (w[1] >= 0) & (u2 >= 0), axis=(1, 2))
    ok_full = ~full | covered
    return ok_z, ok_full, full, finite


def test_bounds_hold_for_walked_values():
    bad_z = bad_full = 0
    n_full = n_cases = 0
    for seed in range(4):
        ok_z, ok_full, full, finite = check(seed, 3000)
        bad_z += int((~ok_z).sum())
        bad_full += int((~ok_full).sum())
        n_full += int(full.sum())
        n_cases += full.size
        assert finite.all()
    assert bad_z == 0, f"{bad_z} tiles with a walked z outside [zlo, zhi]"
    assert bad_full == 0, f"{bad_full} tiles claimed covered with an uncovered pixel"
    assert n_full > n_cases // 10, "too few full-cover cases to test the coverage claim"


def test_check_fails_without_margins():
    # the same walks against bounds with no margins: the walks' roundings must
    # carry some z outside the corners' range and some claimed-full tile below
    # zero, or the test above could not tell a wrong margin from a right one
    bad_z = bad_full = 0
    for seed in range(4):
        ok_z, ok_full, _, _ = check(seed, 3000, margins=False)
        bad_z += int((~ok_z).sum())
        bad_full += int((~ok_full).sum())
    assert bad_z > 0
    assert bad_full > 0
