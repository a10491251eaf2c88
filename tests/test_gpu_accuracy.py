"""Weighted accuracy against the FP64 oracle on the equatorial s-mu +
sigma-pi workload's own axes (s-mu 40 log bins base 1.3 to 100 x 50 mu;
sigma-pi 50 x 50 linear to 50 Mpc/h; weights U[0.5, 1.5]) at 5 % of its
size (1k parents x Poisson(50) and 250k randoms on the full footprint):
DD, DR and RR from both GPU methods against the oracle, then normalise()
(hist.cpp:290-367) and the LS estimator (io.cpp:643-651) on each.  The
targets are a relative difference of at most 1e-10 on every nonzero sum and
at most 1e-8 on xi; the largest differences seen are printed."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SUM_RTOL = 1e-10
XI_ATOL = 1e-8


def test_weighted_xi_equatorial(oracle_lib):
    from two_pcf_amd import lib as L
    from two_pcf_amd import synth
    o = oracle_lib
    c = synth.equatorial_smu_sigma_pi(scale=0.05, seed=20260101)
    host, dev, oc = {}, {}, {}
    with L.Context(0) as ctx:
        for tag in ("data", "randoms"):
            a = c[tag]
            d = L.comoving_distance(0.25, 1.0, 0.0, 1.0, a["z"])
            x, y, z, nx, ny, nz = L.eq_to_cart(a["ra"], a["dec"], d)
            dev[tag] = ctx.upload(x, y, z, nx, ny, nz, a["w"], a["jk"])
            oc[tag] = o.Catalog(x, y, z, nx, ny, nz, a["w"], a["jk"])
        D, R = oc["data"], oc["randoms"]
        spec = (("data", None), ("data", "randoms"), ("randoms", None))
        for name, kind, axes in (("s_mu", 2, [(0.0, 100.0, 40, 1.3), (0.0, 1.0, 50, 0.0)]),
                                 ("sigma_pi", 1, [(0.0, 50.0, 50, 1.0), (0.0, 50.0, 50, 1.0)])):
            ax1, ax2 = o.axis(*axes[0]), o.axis(*axes[1])
            want = [o.finish(b is None, *o.count(kind, oc[a], oc[b] if b else None, ax1, ax2, 1)[:2])
                    for a, b in spec]

            def xi(raw):
                dd = o.normalise_auto(*raw[0], D)[0]
                dr = o.normalise_cross(*raw[1], D, R)[0]
                rr = o.normalise_auto(*raw[2], R)[0]
                return np.array([o.lib().or_ls(u, v, w) for u, v, w in zip(dd, dr, rr)])

            xi_want = xi(want)
            for method in (0, 1):
                res, _ = ctx.count([dict(kind=kind, a=dev[a], b=dev[b] if b else None, axes=axes, n_jk=1)
                                    for a, b in spec], method=method)
                got = [L.finish(b is None, r["full"], r["jk"]) for r, (a, b) in zip(res, spec)]
                rel = 0.0
                for g, w in zip(got, want):
                    nz_ = w[0] != 0
                    assert np.array_equal(g[0] == 0, w[0] == 0)
                    rel = max(rel, float(np.max(np.abs(g[0][nz_] - w[0][nz_]) / np.abs(w[0][nz_]))))
                dxi = float(np.max(np.abs(xi(got) - xi_want)))
                print(f"[gpu] accuracy {name} method={method}: largest relative difference of the sums "
                      f"{rel:.3g}, largest |xi difference| {dxi:.3g}, pairs DD+DR+RR "
                      f"{sum(r['n_pairs'] for r in res)}", flush=True)
                assert rel <= SUM_RTOL
                assert dxi <= XI_ATOL
