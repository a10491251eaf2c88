"""Adversarial parity for the FP32 + edge-recheck method: isolated pairs
placed a few float ulps either side of every bin edge of the s axis (s-mu,
monopole, sigma-pi range) and of the angular axis, far from the origin (where
FP32 absolute positions would be useless).  Each pair is alone within the
maximum separation, so every histogram entry is one constructed pair; the
counts of both GPU methods must equal the oracle's."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _pairs_at_edges(L, s_edges, n_dir, spacing, seed):
    """Point pairs (A, B) with |A - B| = edge * (1 + m 2^-24), m in -6..6,
    n_dir random orientations each, pair centres on a lattice of the given
    spacing around (1500, -900, 700) Mpc/h."""
    rng = np.random.Generator(np.random.PCG64(seed))
    seps = []
    for e in s_edges:
        if not (e > 0 and np.isfinite(e)):
            continue
        for m in range(-6, 7):
            seps += [e * (1.0 + m * 2.0 ** -24)] * n_dir
    seps = np.array(seps)
    n = seps.size
    side = int(np.ceil(n ** (1 / 3)))
    idx = np.arange(n)
    centre = np.stack([idx % side, (idx // side) % side, idx // (side * side)], 1) * spacing
    centre = centre + np.array([1500.0, -900.0, 700.0]) - 0.5 * side * spacing
    u = rng.normal(size=(n, 3))
    u /= np.linalg.norm(u, axis=1)[:, None]
    a = centre - 0.5 * seps[:, None] * u
    b = centre + 0.5 * seps[:, None] * u
    pts = np.concatenate([a, b])
    x, y, z = (np.ascontiguousarray(pts[:, k]) for k in range(3))
    _, nx, ny, nz = L.cart_norm(x, y, z)
    return x, y, z, nx, ny, nz


@pytest.mark.parametrize("kind,axes", [
    (2, [(0.0, 100.0, 40, 1.3), (0.0, 1.0, 50, 0.0)]),
    (0, [(0.0, 100.0, 30, 1.3)]),
    (0, [(5.0, 80.0, 25, 1.0)]),
    (1, [(0.0, 50.0, 50, 1.0), (0.0, 50.0, 50, 1.0)]),
    (1, [(1.0, 40.0, 20, 1.25), (0.5, 40.0, 20, 1.3)]),
])
def test_pairs_at_bin_edges(oracle_lib, kind, axes):
    from two_pcf_amd import lib as L
    o = oracle_lib
    ax1 = o.axis(*axes[0])
    ax2 = o.axis(*axes[1]) if len(axes) > 1 else None
    # edges of the s axis (or of sigma, pi and the sigma-pi range)
    if kind == 1:
        s_edges = list(np.sqrt(np.maximum(L.axis_thresholds(axes[0]), 0)))
        s_edges += list(np.abs(L.axis_thresholds(axes[1])))
        mx = np.sqrt(axes[0][1] ** 2 + axes[1][1] ** 2)
        s_edges += [mx, min(axes[0][0], axes[1][0])]   # the range limits (correlators.cpp:192-195)
        rmax = mx
    else:
        t = L.s2_thresholds(kind, axes[0], axes[1] if len(axes) > 1 else None)
        s_edges = list(np.sqrt(np.maximum(t.astype(np.float64), 0)))
        rmax = axes[0][1]
    x, y, z, nx, ny, nz = _pairs_at_edges(L, s_edges, 6, 2.2 * rmax, seed=kind + 7)
    w = np.ones(x.size, np.float32)
    jk = (np.arange(x.size) % 2).astype(np.int32)
    oc = o.Catalog(x, y, z, nx, ny, nz, w, jk)
    full, sub, acc = o.count(kind, oc, None, ax1, ax2, 2)
    assert acc > 0
    with L.Context(0) as ctx:
        c = ctx.upload(x, y, z, nx, ny, nz, w, jk)
        for method in (0, 1):
            print(f"[gpu] edges kind={kind} axes={axes} method={method} n={x.size}", flush=True)
            res, t = ctx.count([dict(kind=kind, a=c, b=None, axes=axes, n_jk=2)], method=method)
            np.testing.assert_array_equal(res[0]["full"], full)
            np.testing.assert_array_equal(res[0]["jk"], sub)
            assert res[0]["n_pairs"] == acc
            if method == 1:
                assert t["n_rechecked"] > 0   # the construction lands inside the margins


def test_angular_pairs_at_bin_edges(oracle_lib):
    """Isolated pairs on the sphere whose 1 - cos straddles every threshold of
    a log angular axis by a few float ulps."""
    from two_pcf_amd import lib as L
    o = oracle_lib
    ax = (0.0, float(1.0 - np.cos(0.02)), 16, 1.6)
    oax = o.axis(*ax)
    t = L.axis_thresholds(ax).astype(np.float64)
    vals = []
    for e in t:
        if e > 0 and np.isfinite(e):
            for m in range(-6, 7):
                vals.append(e * (1.0 + m * 2.0 ** -24))
    vals = np.array(vals)
    theta = np.arccos(1.0 - vals)
    # pair centres on a (ra, dec) lattice 0.05 rad apart
    n = theta.size
    side = int(np.ceil(np.sqrt(n)))
    k = np.arange(n)
    ra0 = 0.3 + (k % side) * 0.05
    dec0 = -0.6 + (k // side) * 0.05
    rng = np.random.Generator(np.random.PCG64(3))
    phi = rng.uniform(0, 2 * np.pi, n)
    def unit(ra, dec):
        return np.stack([np.cos(dec) * np.cos(ra), np.cos(dec) * np.sin(ra), np.sin(dec)], 1)
    c = unit(ra0, dec0)
    e1 = np.cross(c, [0.0, 0.0, 1.0])
    e1 /= np.linalg.norm(e1, axis=1)[:, None]
    e2 = np.cross(c, e1)
    d = np.cos(phi)[:, None] * e1 + np.sin(phi)[:, None] * e2
    a = np.cos(theta / 2)[:, None] * c - np.sin(theta / 2)[:, None] * d
    b = np.cos(theta / 2)[:, None] * c + np.sin(theta / 2)[:, None] * d
    pts = np.concatenate([a, b]) * 800.0
    x, y, z = (np.ascontiguousarray(pts[:, i]) for i in range(3))
    _, nx, ny, nz = L.cart_norm(x, y, z)
    w = np.ones(x.size, np.float32)
    jk = np.zeros(x.size, np.int32)
    oc = o.Catalog(x, y, z, nx, ny, nz, w, jk)
    full, sub, acc = o.count(3, oc, None, oax, None, 1)
    assert acc > 0
    with L.Context(0) as ctx:
        cat = ctx.upload(x, y, z, nx, ny, nz, w, jk)
        for method in (0, 1):
            print(f"[gpu] angular edges method={method} n={x.size}", flush=True)
            res, tm = ctx.count([dict(kind=L.ANGULAR, a=cat, b=None, axes=[ax], n_jk=1)], method=method)
            np.testing.assert_array_equal(res[0]["full"], full)
            assert res[0]["n_pairs"] == acc
            if method == 1:
                assert tm["n_rechecked"] > 0
