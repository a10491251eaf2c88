"""The benchmark workloads at full size on the GPU, through properties that
hold at any size: the batched FP32 + edge-recheck method gives the plain
FP64 method's counts at every cell subdivision (the plain method evaluates
every candidate pair with the reference's own arithmetic,
correlators.cpp:309-343); weighted sums agree to 1e-12; the jackknife
sums obey the identity of the "touching" convention (hist.cpp:484-502:
each pair touches one region, or two when its regions differ, so
full <= sum over regions of touching <= 2 full); and the accepted-pair
counter equals the sum of the unweighted histogram.  The seeded generators
make every workload's pair count a fixed number, which is asserted too
(the plain method's count)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

S_MU = [(0.0, 100.0, 40, 1.3), (0.0, 1.0, 50, 0.0)]
SIGMA_PI = [(0.0, 50.0, 50, 1.0), (0.0, 50.0, 50, 1.0)]


def upload(L, ctx, a, masks=False):
    d = L.comoving_distance(0.25, 1.0, 0.0, 1.0, a["z"])
    x, y, z, nx, ny, nz = L.eq_to_cart(a["ra"], a["dec"], d)
    return ctx.upload(x, y, z, nx, ny, nz, a["w"], a["jk"], iw=a.get("iw"), masks=a.get("masks") if masks else None)


def agree(got, want, what, rtol=0.0):
    for k, (g, w) in enumerate(zip(got, want)):
        assert g["n_pairs"] == w["n_pairs"], f"{what} job {k}"
        for f in ("full", "jk"):
            assert np.array_equal(np.isnan(g[f]), np.isnan(w[f])), f"{what} job {k} {f}: NaN bins"
            ok = ~np.isnan(w[f])
            if rtol == 0.0:
                np.testing.assert_array_equal(g[f][ok], w[f][ok], err_msg=f"{what} job {k} {f}")
            else:
                np.testing.assert_allclose(g[f][ok], w[f][ok], rtol=rtol, atol=0, err_msg=f"{what} job {k} {f}")


def test_jackknife_batch_full_size():
    """2M + 10M, 100 regions, s-mu 40 x 50: plain and batched at cell
    subdivisions 1, 2 and 3 identical; the touching identity per bin."""
    from two_pcf_amd import lib as L
    from two_pcf_amd import synth
    c = synth.jackknife_batch()
    with L.Context(0) as ctx:
        D, R = upload(L, ctx, c["data"]), upload(L, ctx, c["randoms"])
        jobs = [dict(kind=L.S_MU, a=D, b=None, axes=S_MU, n_jk=100), dict(kind=L.S_MU, a=D, b=R, axes=S_MU, n_jk=100),
                dict(kind=L.S_MU, a=R, b=None, axes=S_MU, n_jk=100)]
        print("[gpu] full-size jackknife batch, plain", flush=True)
        plain, t = ctx.count(jobs, method=L.PLAIN_FP64)
        assert t["n_pairs"] == 146175129905
        for r in plain:
            assert r["full"].sum() == r["n_pairs"]
            touching = r["jk"].sum(axis=0)
            assert np.all(touching >= r["full"]) and np.all(touching <= 2 * r["full"])
        for subdiv in (1, 2, 3):
            print(f"[gpu] full-size jackknife batch, batched, cell subdivision {subdiv}", flush=True)
            res, _ = ctx.count(jobs, method=L.FP32_RECHECK, tuning={"cell_subdiv": subdiv})
            agree(res, plain, f"subdivision {subdiv}")


def test_equatorial_full_size():
    """1M + 5M with weights U[0.5, 1.5], s-mu and sigma-pi DD, DR, RR in one
    call: counts identical, weighted sums within 1e-12."""
    from two_pcf_amd import lib as L
    from two_pcf_amd import synth
    c = synth.equatorial_smu_sigma_pi()
    with L.Context(0) as ctx:
        D, R = upload(L, ctx, c["data"]), upload(L, ctx, c["randoms"])
        jobs = []
        for kind, axes in ((L.S_MU, S_MU), (L.SIGMA_PI, SIGMA_PI)):
            jobs += [dict(kind=kind, a=D, b=None, axes=axes, n_jk=1), dict(kind=kind, a=D, b=R, axes=axes, n_jk=1),
                     dict(kind=kind, a=R, b=None, axes=axes, n_jk=1)]
        print("[gpu] full-size equatorial, plain and batched", flush=True)
        plain, t = ctx.count(jobs, method=L.PLAIN_FP64)
        assert t["n_pairs"] == 43800261027
        res, _ = ctx.count(jobs, method=L.FP32_RECHECK)
        agree(res, plain, "equatorial", rtol=1e-12)


def test_inverse_p_full_size():
    """500k with 2048-bit masks + 2.5M: the angular DD/DR with parent and
    inverse-p weights, the up-weights formed from them (empty bins give NaN,
    as in the reference), then s-mu DD (inverse-p), DR (up-weight) and RR:
    both methods agree, NaN bins included."""
    from two_pcf_amd import lib as L
    from two_pcf_amd import synth
    c = synth.inverse_p_catalogs()
    n_runs = c["n_runs"]
    ang = (0.0, float(1.0 - np.cos(0.175)), 40, 1.6)
    with L.Context(0) as ctx:
        D, R = upload(L, ctx, c["data"], masks=True), upload(L, ctx, c["randoms"], masks=True)
        ajobs = [dict(kind=L.ANGULAR, a=D, b=None, axes=[ang], weight_mode=L.W_ANG_IW, n_runs=n_runs),
                 dict(kind=L.ANGULAR, a=D, b=None, axes=[ang], weight_mode=L.W_ANG_INVP, n_runs=n_runs),
                 dict(kind=L.ANGULAR, a=D, b=R, axes=[ang], weight_mode=L.W_ANG_IW, n_runs=n_runs),
                 dict(kind=L.ANGULAR, a=D, b=R, axes=[ang], weight_mode=L.W_ANG_W, n_runs=n_runs)]
        print("[gpu] full-size inverse-p, angular", flush=True)
        plain_a, ta = ctx.count(ajobs, method=L.PLAIN_FP64)
        assert ta["n_pairs"] == 153955235474
        res_a, _ = ctx.count(ajobs, method=L.FP32_RECHECK)
        agree(res_a, plain_a, "angular", rtol=1e-12)
        fin = [L.finish(j["b"] is None, r["full"], r["jk"])[0] for r, j in zip(plain_a, ajobs)]
        with np.errstate(invalid="ignore", divide="ignore"):
            up_dd, up_dr = fin[0] / fin[1], fin[2] / fin[3]
        tjobs = [dict(kind=L.S_MU, a=D, b=None, axes=S_MU, weight_mode=L.W_INVP_3D, n_runs=n_runs, up_bins=up_dd,
                      up_axis=ang),
                 dict(kind=L.S_MU, a=D, b=R, axes=S_MU, weight_mode=L.W_ANGUP_3D, n_runs=n_runs, up_bins=up_dr,
                      up_axis=ang),
                 dict(kind=L.S_MU, a=R, b=None, axes=S_MU)]
        print("[gpu] full-size inverse-p, 3-D", flush=True)
        plain_t, tt = ctx.count(tjobs, method=L.PLAIN_FP64)
        assert tt["n_pairs"] == 9145198924
        res_t, _ = ctx.count(tjobs, method=L.FP32_RECHECK)
        agree(res_t, plain_t, "3-D", rtol=1e-12)
