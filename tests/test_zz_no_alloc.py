"""twopcf_count_noalloc (the entry point without device allocation) against
twopcf_count, through the C ABI, on the GPU: two calls in a row, the second
with a smaller input, both methods; the reference's golden counts; and the
error paths (a catalogue not yet gridded for the job set, a job set larger
than the scratch array)."""
import numpy as np
import pytest

from cases import load_case

pytestmark = pytest.mark.gpu

KIND = {"monopole": 0, "sigma_pi": 1, "s_mu": 2}


def upload(L, ctx, inputs, n=None):
    """Equatorial catalogues as the reader leaves them (first n objects)."""
    cats = {}
    for tag, a in inputs.items():
        k = slice(None) if n is None else slice(0, n)
        d = L.comoving_distance(0.25, 1.0, 0.0, 1.0, a["z"][k])
        x, y, z, nx, ny, nz = L.eq_to_cart(a["ra"][k], a["dec"][k], d)
        cats[tag] = ctx.upload(x, y, z, nx, ny, nz, a["w"][k], a["jk"][k])
    return cats


def jobs_of(L, case, cats, corr):
    axes = case.axes(corr)
    return [dict(kind=KIND[corr], a=cats["data"], b=None, axes=axes, n_jk=case.n_jk),
            dict(kind=KIND[corr], a=cats["data"], b=cats["randoms"], axes=axes, n_jk=case.n_jk),
            dict(kind=KIND[corr], a=cats["randoms"], b=None, axes=axes, n_jk=case.n_jk)]


def same(got, want, what):
    """Counts bit for bit; weighted sums (added by atomics in any order)
    to 1e-12 relative."""
    for g, w in zip(got, want):
        assert g["n_pairs"] == w["n_pairs"], what
        assert g["n_rechecked"] == w["n_rechecked"], what
        np.testing.assert_allclose(g["full"], w["full"], rtol=1e-12, atol=0, err_msg=what)
        np.testing.assert_allclose(g["jk"], w["jk"], rtol=1e-12, atol=0, err_msg=what)


@pytest.mark.parametrize("method", [0, 1], ids=["fp64", "fp32"])
def test_noalloc_matches_count_twice_smaller_second(method):
    from two_pcf_amd import lib as L
    case = load_case("eq_clustered")
    with L.Context(0) as ctx:
        big = upload(L, ctx, case.inputs)
        small = upload(L, ctx, case.inputs, n=400)
        for tag, cats in (("full", big), ("small", small)):
            for corr in case.corrs():
                print(f"[gpu] noalloc {tag} {corr} method={method}", flush=True)
                jobs = jobs_of(L, case, cats, corr)
                want, _ = ctx.count(jobs, method=method)        # grids the catalogues
                got, t = ctx.count(jobs, method=method, noalloc=True)
                same(got, want, f"{tag} {corr} method={method}")
                assert t["n_pairs"] == sum(w["n_pairs"] for w in want)
                if tag == "full":
                    # the reference's own counts
                    for h, r, j in zip(("DD", "DR", "RR"), got, jobs):
                        full, sub = L.finish(j["b"] is None, r["full"], r["jk"])
                        gf, gs = case.golden(corr, h)
                        np.testing.assert_array_equal(full, gf, err_msg=f"{corr} {h}")
                        np.testing.assert_array_equal(sub, gs, err_msg=f"{corr} {h}")


def test_noalloc_refuses_ungridded_and_oversized():
    from two_pcf_amd import lib as L
    case = load_case("eq_clustered")
    with L.Context(0) as ctx:
        cats = upload(L, ctx, case.inputs, n=300)
        jobs = jobs_of(L, case, cats, "s_mu")
        print("[gpu] noalloc refuses a catalogue not gridded for the job set", flush=True)
        with pytest.raises(L.TwopcfError, match="not yet gridded"):
            ctx.count(jobs, method=1, noalloc=True)
        # 2000 s-mu bins x 5001 histograms x 8 B = 80 MB of output pools
        print("[gpu] noalloc refuses a job set larger than its scratch", flush=True)
        big = [dict(kind=2, a=cats["data"], b=None, axes=case.axes("s_mu"), n_jk=5000)]
        want, _ = ctx.count(big, method=0)
        with pytest.raises(L.TwopcfError, match="MiB scratch"):
            ctx.count(big, method=0, noalloc=True)
        # the refusals leave the entry point usable
        ctx.count(jobs, method=1)
        got, _ = ctx.count(jobs, method=1, noalloc=True)
        want, _ = ctx.count(jobs, method=1)
        same(got, want, "after refusals")
