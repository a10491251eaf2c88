"""Pin the oracle's inverse-p pipeline against the reference built here with
_USE_INV_WEIGHTS (2 and 32 mask words) and _USE_DITHER_WEIGHTS (3 words)
from its own sources (oracle/Makefile; fixtures by
tests/golden/make_invp_golden.py):

- IIP weights at read (io.cpp:521-588), with the weights computed by the
  test itself from the masks;
- the four angular histograms of calc_angular_upweight
  (correlators_invpweight.cpp:713-1033: parent and inverse-p weights, DD and
  DR), including the angular value 1 - n1.n2 and its float binning;
- the up-weight tables (divide_by_hist, hist.cpp:224-231) and
  JKHists::interp (hist.cpp:369-391) at every bin centre and edge and at
  float neighbours of every segment switch;
- the 3-D DD with inverse-p weights, DR with the up-weight and RR, raw and
  normalised, for monopole, sigma-pi and s-mu; on the 32-word case on the
  angular axis theta < 0.175 rad, 40 log bins base 1.6, where the first
  angular bins are empty and the reference's up-weights there are 0/0 = NaN.

Everything is bit-exact except weighted sums whose additions the reference
makes in another order: those agree to a few units in the last place (the
bound below, 1e-14 relative), far below what a changed rounding of the
per-pair float weight would move them (about 1e-10 on these sizes)."""
import ctypes as C

import numpy as np
import pytest

from invp_cases import ANGULAR, CASES, InvpCase

SUM_RTOL = 1e-14
KIND = {"monopole": 0, "sigma_pi": 1, "s_mu": 2}


def same(got, want, rtol=0.0):
    got, want = np.asarray(got, np.float64), np.asarray(want, np.float64)
    assert got.shape == want.shape
    assert np.array_equal(np.isnan(got), np.isnan(want)), "NaN bins differ"
    ok = ~np.isnan(want)
    if rtol == 0.0:
        assert np.array_equal(got[ok], want[ok])
    else:
        np.testing.assert_allclose(got[ok], want[ok], rtol=rtol, atol=0)


@pytest.fixture(scope="module", params=CASES)
def case(request, oracle_lib):
    c = InvpCase(request.param)
    D, R = c.oracle_cats(oracle_lib)
    return c, D, R


def test_iip_weights_and_geometry(case):
    """The catalogue as the reference's hdf5 reader leaves it: positions and
    unit vectors (cosmology table + Galaxy::set), input weight, and the IIP
    weight n_runs/total (dither_total/total) where total != n_runs."""
    c, D, R = case
    for tag, cat in (("data", D), ("rand", R)):
        for f in ("x", "y", "z", "nx", "ny", "nz", "w", "iw"):
            same(getattr(cat, f), c.g[f"cat_{tag}_{f}"])
        # the rule is exercised: data weights change (the first three objects
        # have all-zero masks: total counts as 1); randoms with all-ones masks
        # keep theirs
        changed = cat.w != cat.iw
        if tag == "data":
            assert changed[:3].all()
        elif c.p["n_bitwise_runs"] == "2048":
            assert not changed.any()
        else:
            assert changed.any() and not changed[:5].any()


def _fin(o, c, kind, a, b, ax1, ax2, wo):
    full, sub, _ = o.count(kind, a, b, ax1, ax2, c.n_jk, wopt=wo)
    return o.finish(b is None, full, sub)


def _angular(o, c, D, R):
    ang = o.axis(*c.angular_axis())
    none = o.Axis()
    dith = int(c.dither)
    return ang, dict(
        DD=_fin(o, c, 3, D, None, ang, None, o.WOpt(o.W_ANG_IW, c.n_runs, None, none)),
        DDinvp=_fin(o, c, 3, D, None, ang, None, o.WOpt(o.W_ANG_INVP, c.n_runs, None, none, 0, dith)),
        DR=_fin(o, c, 3, D, R, ang, None, o.WOpt(o.W_ANG_IW, c.n_runs, None, none)),
        DRinvp=_fin(o, c, 3, D, R, ang, None, o.WOpt(o.W_ANG_W, c.n_runs, None, none)))


def test_angular_histograms(case, oracle_lib):
    c, D, R = case
    _, res = _angular(oracle_lib, c, D, R)
    for k in ANGULAR:
        same(res[k][0], c.g[f"ang_{k}_full"], SUM_RTOL)
        same(res[k][1], c.g[f"ang_{k}_sub"], SUM_RTOL)
        assert np.nansum(res[k][0]) > 0


def test_upweight_tables_and_interp(case, oracle_lib):
    """divide_by_hist of the oracle's angular histograms gives the
    reference's tables; or_interp at the probes gives JKHists::interp's
    values bit for bit (the probes straddle every segment switch)."""
    o = oracle_lib
    c, D, R = case
    ang, res = _angular(o, c, D, R)
    with np.errstate(invalid="ignore", divide="ignore"):
        same(res["DD"][0] / res["DDinvp"][0], c.g["up_dd_bins"], SUM_RTOL)
        same(res["DR"][0] / res["DRinvp"][0], c.g["up_dr_bins"], SUM_RTOL)
    L = o.lib()
    for k in ("dd", "dr"):
        bins = np.ascontiguousarray(c.g[f"up_{k}_bins"])
        got = [L.or_interp(C.byref(ang), bins.ctypes.data, float(v)) for v in c.g["probes"]]
        same(got, c.g[f"up_{k}_interp"])


def test_3d_inverse_p_counts(case, oracle_lib):
    """auto_pair_count_*_invpweights (DD), cross_pair_count_*_invpweights (DR,
    the up-weight rule) and auto_pair_count_* (RR), each with the
    reference's own up-weight table, raw and after normalise()."""
    o = oracle_lib
    c, D, R = case
    ang = o.axis(*c.angular_axis())
    up_dd = np.ascontiguousarray(c.g["up_dd_bins"])
    up_dr = np.ascontiguousarray(c.g["up_dr_bins"])
    dith = int(c.dither)
    for corr in c.corrs():
        axs = c.axes(corr)
        ax1 = o.axis(*axs[0])
        ax2 = o.axis(*axs[1]) if len(axs) > 1 else None
        kind = KIND[corr]
        got = dict(
            DD=_fin(o, c, kind, D, None, ax1, ax2, o.WOpt(o.W_INVP_3D, c.n_runs, up_dd.ctypes.data, ang, 0, dith)),
            DR=_fin(o, c, kind, D, R, ax1, ax2, o.WOpt(o.W_ANGUP_3D, c.n_runs, up_dr.ctypes.data, ang)),
            RR=_fin(o, c, kind, R, None, ax1, ax2, None))
        norm = dict(DD=o.normalise_auto(*got["DD"], D), DR=o.normalise_cross(*got["DR"], D, R),
                    RR=o.normalise_auto(*got["RR"], R))
        for t in ("DD", "DR", "RR"):
            same(got[t][0], c.g[f"{corr}_{t}_full"], SUM_RTOL)
            same(got[t][1], c.g[f"{corr}_{t}_sub"], SUM_RTOL)
            same(norm[t][0], c.g[f"{corr}_norm_{t}_full"], SUM_RTOL)
            same(norm[t][1], c.g[f"{corr}_norm_{t}_sub"], SUM_RTOL)


def test_fixtures_exercise_the_rules(case, oracle_lib):
    """The fixtures are not degenerate: the inverse-p DD differs from the
    parent DD; the 32-word case has NaN up-weights in its empty first bins;
    the dither case differs from the same masks' n_runs numerator; the
    probes land on both sides of each segment switch."""
    o = oracle_lib
    c, D, R = case
    g = c.g
    assert not np.allclose(np.nan_to_num(g["ang_DD_full"]), np.nan_to_num(g["ang_DDinvp_full"]))
    if c.name == "invp_w32_axis":
        assert int(c.p["theta_n_bins"]) == 40 and float(c.p["theta_max"]) == 0.175
        assert np.isnan(g["up_dd_bins"]).any() and np.isnan(g["s_mu_DD_full"]).any()
    if c.dither:
        ang = o.axis(*c.angular_axis())
        plain = _fin(o, c, 3, D, None, ang, None, o.WOpt(o.W_ANG_INVP, c.n_runs, None, o.Axis(), 0, 0))[0]
        assert np.abs(plain - g["ang_DDinvp_full"]).max() > 1e-3 * np.abs(plain).max()
    # interp segments: for a log axis the probes include values whose
    # bin_num - 0.5 falls on both sides of each integer
    L = o.lib()
    ang = o.axis(*c.angular_axis())
    idx = np.array([L.or_axis_bin_num(C.byref(ang), float(np.float32(v))) - 0.5 for v in g["probes"]])
    fl = np.floor(idx[(idx >= 0) & (idx < ang.n_bins - 1)]).astype(int)
    assert set(range(ang.n_bins - 1)) <= set(fl.tolist())
