"""Pin the CPU oracle against the golden vectors written by the reference
binary (tests/golden/make_golden.py), against brute force, and against
closed forms.  CPU only."""
import math

import numpy as np
import pytest

from cases import load_case

CASES = ["cart_uniform", "eq_clustered", "edge", "uniform_monopole_50k", "cart_midpoint", "eq_midpoint"]
KIND = {"monopole": 0, "sigma_pi": 1, "s_mu": 2}


def oracle_cats(o, case):
    cats = {}
    if case.coord == "equatorial":
        tab = o.cosmo_table(case.fp("omega_m"), case.fp("h"), case.fp("z_min"), case.fp("z_max"))
    for tag in ("data", "randoms"):
        a = case.inputs[tag]
        if case.coord == "cartesian":
            d, nx, ny, nz = o.cart_norm(a["x"], a["y"], a["z"])
            x, y, z = a["x"], a["y"], a["z"]
        else:
            d = o.comov_dist(tab, case.fp("z_min"), a["z"])
            x, y, z, nx, ny, nz = o.eq_to_cart(a["ra"], a["dec"], d)
        cats[tag] = o.Catalog(x, y, z, nx, ny, nz, a["w"], a["jk"])
        cats[tag].d = d
    return cats


def oracle_axes(o, case, corr):
    ax = [o.axis(*t) for t in case.axes(corr)]
    return ax[0], (ax[1] if len(ax) > 1 else None)


@pytest.mark.parametrize("name", CASES)
def test_axis_constants_match_reference(oracle_lib, name):
    o = oracle_lib
    case = load_case(name)
    for corr in case.corrs():
        axes = oracle_axes(o, case, corr)
        for k, ax in enumerate(a for a in axes if a is not None):
            g = case.g[f"{corr}_axis{k + 1}"]
            mine = np.array([ax.min, ax.max, ax.step, ax.inv_step, ax.ilb, ax.ils], dtype=np.float32)
            np.testing.assert_array_equal(mine.view(np.uint32), g.view(np.uint32))
            assert [ax.n_bins, ax.log_binning] == list(case.g[f"{corr}_axis{k + 1}_meta"])


def test_catalogue_geometry_matches_reference(oracle_lib):
    """Cosmology table + Galaxy::set reproduce the reference's positions and
    unit vectors bit for bit (cosmology.cpp, dataset.cpp, galaxy.cpp)."""
    o = oracle_lib
    case = load_case("eq_clustered")
    cats = oracle_cats(o, case)
    for tag in ("data", "randoms"):
        c = cats[tag]
        for fld, mine in (("x", c.x), ("y", c.y), ("z", c.z), ("nx", c.nx), ("ny", c.ny), ("nz", c.nz), ("d", c.d)):
            np.testing.assert_array_equal(mine, case.g[f"cat_{tag}_{fld}"], err_msg=f"{tag}.{fld}")


@pytest.mark.parametrize("name", CASES)
def test_counts_match_reference(oracle_lib, name):
    o = oracle_lib
    case = load_case(name)
    cats = oracle_cats(o, case)
    for corr in case.corrs():
        ax1, ax2 = oracle_axes(o, case, corr)
        kind = KIND[corr]
        for h, a, b in (("DD", "data", None), ("DR", "data", "randoms"), ("RR", "randoms", None)):
            wo = o.WOpt(0, 0, None, o.Axis(), case.pi_def)
            full, sub, acc = o.count(kind, cats[a], cats[b] if b else None, ax1, ax2, case.n_jk, method=1, wopt=wo)
            ff, fs = o.finish(b is None, full, sub)
            gf, gs = case.golden(corr, h)
            np.testing.assert_array_equal(ff, gf, err_msg=f"{name} {corr} {h} full")
            np.testing.assert_array_equal(fs, gs, err_msg=f"{name} {corr} {h} sub")
            if np.all(cats[a].w == 1) and (b is None or np.all(cats[b].w == 1)):
                assert full.sum() == acc
            # normalise (hist.cpp:512-589), incl. the n_jk<=1 all-NaN block
            if b is None:
                nf, ns = o.normalise_auto(ff, fs, cats[a])
            else:
                nf, ns = o.normalise_cross(ff, fs, cats[a], cats[b])
            gnf, gns = case.golden(corr, h, norm=True)
            np.testing.assert_array_equal(nf, gnf, err_msg=f"{name} {corr} {h} norm full")
            if gns is not None:
                np.testing.assert_array_equal(ns, gns, err_msg=f"{name} {corr} {h} norm sub")


def test_midpoint_fixture_differs_from_angular(oracle_lib):
    """The pair-midpoint golden vectors are not the angular ones: the two
    line-of-sight definitions bin some pairs differently."""
    o = oracle_lib
    case = load_case("eq_midpoint")
    cats = oracle_cats(o, case)
    ax1, ax2 = oracle_axes(o, case, "s_mu")
    full, sub, _ = o.count(2, cats["data"], None, ax1, ax2, case.n_jk, method=1)
    ff, _ = o.finish(True, full, sub)
    gf, _ = case.golden("s_mu", "DD")
    assert ff.sum() == gf.sum() and not np.array_equal(ff, gf)


@pytest.mark.parametrize("name", ["cart_uniform", "edge"])
def test_cells_equal_brute_force(oracle_lib, name):
    o = oracle_lib
    case = load_case(name)
    cats = oracle_cats(o, case)
    for corr in case.corrs():
        ax1, ax2 = oracle_axes(o, case, corr)
        for a, b in (("data", None), ("data", "randoms")):
            r1 = o.count(KIND[corr], cats[a], cats[b] if b else None, ax1, ax2, case.n_jk, method=1)
            r0 = o.count(KIND[corr], cats[a], cats[b] if b else None, ax1, ax2, case.n_jk, method=0)
            np.testing.assert_array_equal(r1[0], r0[0])
            np.testing.assert_array_equal(r1[1], r0[1])
            assert r1[2] == r0[2]


def test_jackknife_closed_form(oracle_lib):
    """sub[r] = sum over pairs with neither member in r (after invert),
    checked against a pure-Python enumeration of a tiny catalogue."""
    o = oracle_lib
    rng = np.random.Generator(np.random.PCG64(5))
    n = 120
    p = rng.uniform(0, 60, (n, 3))
    jk = rng.integers(0, 3, n).astype(np.int32)
    w = rng.uniform(0.5, 1.5, n).astype(np.float32)
    d, nx, ny, nz = o.cart_norm(p[:, 0], p[:, 1], p[:, 2])
    cat = o.Catalog(p[:, 0], p[:, 1], p[:, 2], nx, ny, nz, w, jk)
    ax = o.axis(0.0, 50.0, 10, 1.0)
    full, sub, _ = o.count(0, cat, None, ax, None, 3, method=0)
    ff, fs = o.finish(True, full, sub)
    exp_full = np.zeros(10)
    exp_sub = np.zeros((3, 10))
    for i in range(n):
        for j in range(i + 1, n):
            dd = [np.float32(p[i, k] - p[j, k]) for k in range(3)]
            s2 = np.float32(np.float32(dd[0] * dd[0] + dd[1] * dd[1]) + dd[2] * dd[2])
            if not (0 < s2 < np.float32(2500.0)):
                continue
            b = int(np.float32(np.sqrt(s2, dtype=np.float32) * np.float32(1.0 / np.float32(5.0))))
            if b >= 10:
                continue
            ww = float(np.float32(w[i] * w[j]))
            exp_full[b] += ww
            for r in range(3):
                if jk[i] != r and jk[j] != r:
                    exp_sub[r, b] += ww
    np.testing.assert_allclose(ff, 2 * exp_full, rtol=1e-12)
    np.testing.assert_allclose(fs, 2 * exp_sub, rtol=1e-12)


def test_log_axis_closed_form(oracle_lib):
    """The log index lands in bin k exactly between the double edges
    e_k = (exp(k/ilb)-1)/ils (hist.cpp:55-60), away from float rounding."""
    o = oracle_lib
    ax = o.axis(0.0, 100.0, 40, 1.3)
    L = o.lib()
    for k in range(40):
        lo, hi = L.or_axis_value_from_bin(ax, k), L.or_axis_value_from_bin(ax, k + 1)
        for t in (0.25, 0.5, 0.75):
            v = lo + t * (hi - lo)
            assert int(L.or_axis_bin_num(ax, v)) == k
    # bin edges grow geometrically by the base
    w = [L.or_axis_bin_width(ax, k) for k in range(40)]
    assert all(abs(w[k + 1] / w[k] - 1.3) < 1e-5 for k in range(39))


def test_cosmology_closed_form(oracle_lib):
    """The trapezoid table agrees with a direct quadrature of
    c/H0 * int dz / E(z) to 1e-8 relative."""
    from scipy.integrate import quad
    o = oracle_lib
    tab = o.cosmo_table(0.25, 1.0, 0.0, 1.0)
    for z in (0.05, 0.2, 0.4037, 0.6, 0.99):
        exact = 2997.92458 * quad(lambda t: 1 / math.sqrt(0.25 * (1 + t) ** 3 + 0.75), 0, z, epsabs=0, epsrel=1e-13)[0]
        got = o.comov_dist(tab, 0.0, z)[0]
        assert abs(got / exact - 1) < 1e-8
