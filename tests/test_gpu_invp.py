"""GPU parity of the inverse-p pipeline (correlators_invpweight.cpp): angular
DD/DR with parent and inverse-p weights, the angular up-weight tables built
from them as calc_angular_upweight and the 3-D correlators do
(correlators_invpweight.cpp:985-1033, 584-588), then the 3-D s-mu DD with
bitwise weights, DR with the up-weight, and plain RR.  2048-bit masks (32
words).  Pair counts must be identical; weighted sums agree to 1e-12
relative (their summation order differs from the oracle's)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N_RUNS = 2048
WORDS = 32


def make(L, n, seed, ones=False, ra_l=(180.0, 190.0), dec_l=(10.0, 20.0), z_l=(0.1, 0.15)):
    from two_pcf_amd import synth
    rng = np.random.Generator(np.random.PCG64(seed))
    if ones:
        ra, dec = rng.uniform(*ra_l, n), np.degrees(np.arcsin(rng.uniform(np.sin(np.radians(dec_l[0])),
                                                                            np.sin(np.radians(dec_l[1])), n)))
        z = rng.uniform(*z_l, n)
    else:
        ra, dec, z = synth.neyman_scott_equatorial(max(1, n // 25), 25.0, 4.0, ra_l, dec_l, z_l, seed=seed)
    m = ra.size
    d = L.comoving_distance(0.25, 1.0, 0.0, 1.0, z)
    x, y, zz, nx, ny, nz = L.eq_to_cart(ra, dec, d)
    if ones:
        masks = np.full((m, WORDS), np.uint64(0xFFFFFFFFFFFFFFFF), dtype=np.uint64)
    else:
        p = rng.uniform(0.3, 0.95, m)
        bits = rng.random((m, WORDS * 64)) < p[:, None]
        masks = np.packbits(bits.astype(np.uint8), axis=1, bitorder="little").view(np.uint64).reshape(m, WORDS)
    iw = rng.uniform(0.5, 1.5, m).astype(np.float32)
    # inverse-probability weight at read time (io.cpp:578-582)
    tot = np.array([sum(int(int(v)).bit_count() for v in row) for row in masks])
    tot[tot == 0] = 1
    w = np.where(tot != N_RUNS, (iw.astype(np.float64) * (N_RUNS / tot)).astype(np.float32), iw)
    jk = (ra > 0.5 * (ra_l[0] + ra_l[1])).astype(np.int32)
    return dict(x=x, y=y, z=zz, nx=nx, ny=ny, nz=nz, w=w, iw=iw, masks=masks, jk=jk)


def test_invp_pipeline(oracle_lib):
    from two_pcf_amd import lib as L
    o = oracle_lib
    D = make(L, 2500, 31)
    R = make(L, 5000, 32, ones=True)
    oc = {k: o.Catalog(v["x"], v["y"], v["z"], v["nx"], v["ny"], v["nz"], v["w"], v["jk"], iw=v["iw"], masks=v["masks"])
          for k, v in (("d", D), ("r", R))}
    th = 0.03
    ang = (0.0, float(1.0 - np.cos(th)), 16, 1.6)
    oang = o.axis(*ang)
    with L.Context(0) as ctx:
        dc = {k: ctx.upload(v["x"], v["y"], v["z"], v["nx"], v["ny"], v["nz"], v["w"], v["jk"], iw=v["iw"],
                            masks=v["masks"]) for k, v in (("d", D), ("r", R))}
        # angular: parent DD (iw), inverse-p DD, parent DR (iw), inverse-p DR (w)
        spec = [("d", None, L.W_ANG_IW), ("d", None, L.W_ANG_INVP), ("d", "r", L.W_ANG_IW), ("d", "r", L.W_ANG_W)]
        for method in (0, 1):
            print(f"[gpu] angular method={method}", flush=True)
            jobs = [dict(kind=L.ANGULAR, a=dc[a], b=dc[b] if b else None, axes=[ang], n_jk=2, weight_mode=m,
                         n_runs=N_RUNS) for a, b, m in spec]
            res, _ = ctx.count(jobs, method=method)
            for r, (a, b, m) in zip(res, spec):
                wo = o.WOpt(m, N_RUNS, None, o.Axis())
                full, sub, acc = o.count(3, oc[a], oc[b] if b else None, oang, None, 2, wopt=wo)
                assert r["n_pairs"] == acc
                np.testing.assert_allclose(r["full"], full, rtol=1e-12, atol=0)
                np.testing.assert_allclose(r["jk"], sub, rtol=1e-12, atol=0)
        # up-weight tables: DD/DDinvp and DR/DRinvp per bin (ratios of finished hists)
        fin = [L.finish(b is None, r["full"], r["jk"])[0] for r, (a, b, m) in zip(res, spec)]
        with np.errstate(invalid="ignore", divide="ignore"):
            up_dd = fin[0] / fin[1]
            up_dr = fin[2] / fin[3]
        up_dd = np.where(np.isfinite(up_dd), up_dd, 1.0)
        up_dr = np.where(np.isfinite(up_dr), up_dr, 1.0)
        axes = [(0.0, 40.0, 20, 1.3), (0.0, 1.0, 10, 0.0)]
        ax1, ax2 = o.axis(*axes[0]), o.axis(*axes[1])
        spec3 = [("d", None, L.W_INVP_3D, up_dd), ("d", "r", L.W_ANGUP_3D, up_dr), ("r", None, L.W_PLAIN, None)]
        for method in (0, 1):
            print(f"[gpu] 3-D inverse-p method={method}", flush=True)
            jobs = [dict(kind=L.S_MU, a=dc[a], b=dc[b] if b else None, axes=axes, n_jk=2, weight_mode=m,
                         n_runs=N_RUNS, up_bins=u, up_axis=ang if u is not None else None)
                    for a, b, m, u in spec3]
            res3, _ = ctx.count(jobs, method=method)
            for r, (a, b, m, u) in zip(res3, spec3):
                wo = o.WOpt(m, N_RUNS, u.ctypes.data if u is not None else None, oang)
                full, sub, acc = o.count(2, oc[a], oc[b] if b else None, ax1, ax2, 2, wopt=wo)
                assert r["n_pairs"] == acc
                np.testing.assert_allclose(r["full"], full, rtol=1e-12, atol=0)
                np.testing.assert_allclose(r["jk"], sub, rtol=1e-12, atol=0)


@pytest.mark.parametrize("axis", [(0.0, None, 12, 1.6), (0.0, None, 10, 1.0)])
def test_angular_counts_all_subdivisions(oracle_lib, axis):
    """Unweighted angular DD/DR with 3 regions, duplicate objects included
    (1 - cos = 0), on a log and a linear axis: counts equal the oracle's at
    every cell subdivision of the batched FP32 method and in the plain method."""
    from two_pcf_amd import lib as L
    o = oracle_lib
    D = make(L, 1500, 51, ra_l=(150.0, 156.0), dec_l=(0.0, 6.0))
    R = make(L, 3000, 52, ones=True, ra_l=(150.0, 156.0), dec_l=(0.0, 6.0))
    for cat in (D, R):
        cat["jk"] = (np.arange(cat["x"].size) % 3).astype(np.int32)
        cat["w"] = np.ones(cat["x"].size, np.float32)
        for k in ("x", "y", "z", "nx", "ny", "nz"):
            cat[k][:20] = cat[k][20]   # 21 identical objects
    ax = (axis[0], float(1.0 - np.cos(0.02)), axis[2], axis[3])
    oax = o.axis(*ax)
    oc = {k: o.Catalog(v["x"], v["y"], v["z"], v["nx"], v["ny"], v["nz"], v["w"], v["jk"]) for k, v in (("d", D), ("r", R))}
    want = [o.count(3, oc["d"], None, oax, None, 3), o.count(3, oc["d"], oc["r"], oax, None, 3)]
    with L.Context(0) as ctx:
        dc = {k: ctx.upload(v["x"], v["y"], v["z"], v["nx"], v["ny"], v["nz"], v["w"], v["jk"]) for k, v in (("d", D), ("r", R))}
        jobs = [dict(kind=L.ANGULAR, a=dc["d"], b=None, axes=[ax], n_jk=3),
                dict(kind=L.ANGULAR, a=dc["d"], b=dc["r"], axes=[ax], n_jk=3)]
        for method, tu in ((0, None), (1, {"cell_subdiv": 1}), (1, {"cell_subdiv": 2}), (1, {"cell_subdiv": 3})):
            print(f"[gpu] angular counts method={method} tuning={tu}", flush=True)
            res, _ = ctx.count(jobs, method=method, tuning=tu)
            for r, (full, sub, acc) in zip(res, want):
                np.testing.assert_array_equal(r["full"], full)
                np.testing.assert_array_equal(r["jk"], sub)
                assert r["n_pairs"] == acc


def test_twin_angular_jobs(oracle_lib):
    """Weighted angular jobs of the same catalogues and axis without regions,
    one with the products of the input weights and one with the products of
    the weights (correlators_invpweight.cpp:901, 968 and 853), are counted in
    one pass of the batched method, each pair binned into both histograms
    (twins); a job with the same rule twice is a twin too.  Every histogram
    and accepted-pair count equals the oracle's and the same job counted
    alone, in both methods (sums to 1e-12: the order of addition differs)."""
    from two_pcf_amd import lib as L
    o = oracle_lib
    D = make(L, 2500, 41)
    R = make(L, 5000, 42)   # random masks: w differs from iw
    for cat in (D, R):
        cat["jk"][:] = 0    # one region: none
    oc = {k: o.Catalog(v["x"], v["y"], v["z"], v["nx"], v["ny"], v["nz"], v["w"], v["jk"], iw=v["iw"], masks=v["masks"])
          for k, v in (("d", D), ("r", R))}
    ang = (0.0, float(1.0 - np.cos(0.03)), 16, 1.6)
    oang = o.axis(*ang)
    # the DR pair: twins of two rules; the first two DD jobs: twins of one
    # rule; the last DD job: no partner left
    spec = [("d", "r", L.W_ANG_IW), ("d", "r", L.W_ANG_W), ("d", None, L.W_ANG_IW), ("d", None, L.W_ANG_IW),
            ("d", None, L.W_ANG_W)]
    want = []
    for a, b, m in spec:
        wo = o.WOpt(m, N_RUNS, None, o.Axis())
        want.append(o.count(3, oc[a], oc[b] if b else None, oang, None, 1, wopt=wo))
    with L.Context(0) as ctx:
        dc = {k: ctx.upload(v["x"], v["y"], v["z"], v["nx"], v["ny"], v["nz"], v["w"], v["jk"], iw=v["iw"],
                            masks=v["masks"]) for k, v in (("d", D), ("r", R))}
        jobs = [dict(kind=L.ANGULAR, a=dc[a], b=dc[b] if b else None, axes=[ang], n_jk=1, weight_mode=m,
                     n_runs=N_RUNS) for a, b, m in spec]
        for method in (0, 1):
            print(f"[gpu] twin angular jobs method={method}", flush=True)
            together, _ = ctx.count(jobs, method=method)
            for j, r, (full, _sub, acc) in zip(jobs, together, want):
                alone, _ = ctx.count([j], method=method)
                assert r["n_pairs"] == acc == alone[0]["n_pairs"]
                np.testing.assert_allclose(r["full"], full, rtol=1e-12, atol=0)
                np.testing.assert_allclose(r["full"], alone[0]["full"], rtol=1e-12, atol=0)
        # the twin's weights differ from its partner's: the sums must too
        assert not np.allclose(together[0]["full"], together[1]["full"], rtol=1e-6)


def test_dither_weights(oracle_lib):
    """Dither weights (_USE_DITHER_WEIGHTS, correlators_invpweight.cpp:62-70,
    749-757): the inverse-p DD numerator is popc(dither1 & dither2), 0
    counting as 1.  Angular and 3-D s-mu DD with sparse 32-word dither masks
    equal the oracle in both methods; the numerator really changes the sums;
    a dither job on catalogues without dither masks is refused."""
    from two_pcf_amd import lib as L
    o = oracle_lib
    D = make(L, 2000, 61)
    rng = np.random.Generator(np.random.PCG64(62))
    p = rng.uniform(0.0, 0.05, D["x"].size)
    bits = rng.random((D["x"].size, WORDS * 64)) < p[:, None]
    dm = np.packbits(bits.astype(np.uint8), axis=1, bitorder="little").view(np.uint64).reshape(-1, WORDS)
    oc = o.Catalog(D["x"], D["y"], D["z"], D["nx"], D["ny"], D["nz"], D["w"], D["jk"], iw=D["iw"],
                   masks=D["masks"], dither=dm)
    ang = (0.0, float(1.0 - np.cos(0.03)), 16, 1.6)
    oang = o.axis(*ang)
    up = np.linspace(1.4, 1.0, 16)
    axes = [(0.0, 40.0, 20, 1.3), (0.0, 1.0, 10, 0.0)]
    ax1, ax2 = o.axis(*axes[0]), o.axis(*axes[1])
    with L.Context(0) as ctx:
        dc = ctx.upload(D["x"], D["y"], D["z"], D["nx"], D["ny"], D["nz"], D["w"], D["jk"], iw=D["iw"],
                        masks=D["masks"], dither=dm)
        plain = ctx.upload(D["x"], D["y"], D["z"], D["nx"], D["ny"], D["nz"], D["w"], D["jk"], iw=D["iw"],
                           masks=D["masks"])
        jobs = [dict(kind=L.ANGULAR, a=dc, axes=[ang], n_jk=2, weight_mode=L.W_ANG_INVP, n_runs=N_RUNS, dither=True),
                dict(kind=L.S_MU, a=dc, axes=axes, n_jk=2, weight_mode=L.W_INVP_3D, n_runs=N_RUNS, up_bins=up,
                     up_axis=ang, dither=True)]
        want = [o.count(3, oc, None, oang, None, 2, wopt=o.WOpt(L.W_ANG_INVP, N_RUNS, None, o.Axis(), 0, 1)),
                o.count(2, oc, None, ax1, ax2, 2, wopt=o.WOpt(L.W_INVP_3D, N_RUNS, up.ctypes.data, oang, 0, 1))]
        for method in (0, 1):
            print(f"[gpu] dither weights method={method}", flush=True)
            res, _ = ctx.count(jobs, method=method)
            for r, (full, sub, acc) in zip(res, want):
                assert r["n_pairs"] == acc
                np.testing.assert_allclose(r["full"], full, rtol=1e-12, atol=0)
                np.testing.assert_allclose(r["jk"], sub, rtol=1e-12, atol=0)
        # without the dither flag the same catalogue gives the n_runs numerator
        res, _ = ctx.count([dict(j, dither=False) for j in jobs], method=1)
        for r, (full, _, _) in zip(res, want):
            assert np.abs(r["full"] - full).max() > 1e-3 * np.abs(full).max()
        with pytest.raises(L.TwopcfError, match="dither"):
            ctx.count([dict(jobs[0], a=plain)], method=1)


@pytest.mark.parametrize("words,dither", [(1, False), (2, False), (3, False), (4, False), (5, False), (3, True),
                                          (2, True)])
def test_popcount_rules_word_counts(oracle_lib, words, dither):
    """Angular DD with inverse-p weights and monopole DD with the 3-D
    inverse-p rule, masks of 1 to 5 words.  The batched kernels copy the
    first side's masks into LDS when a row (bit mask, then dither mask) has an
    even word count (rows padded to an odd number of 16-byte units: 2 -> 2,
    4 -> 6 words) and read both sides from global memory otherwise (the
    one-word tail loop); with dither masks of 3 words the dither popcount
    starts at an odd word of the LDS row.  Both methods equal the oracle."""
    from two_pcf_amd import lib as L
    o = oracle_lib
    n_runs = 64 * words
    rng = np.random.Generator(np.random.PCG64(100 + words))
    D = make(L, 1500, 40 + words)
    m = D["x"].size
    p = rng.uniform(0.3, 0.95, m)
    bits = rng.random((m, n_runs)) < p[:, None]
    masks = np.packbits(bits.astype(np.uint8), axis=1, bitorder="little").view(np.uint64).reshape(m, words)
    masks[:5] = 0   # popcount 0 counts as 1
    tot = np.array([sum(int(int(v)).bit_count() for v in row) for row in masks])
    tot[tot == 0] = 1
    w = np.where(tot != n_runs, (D["iw"].astype(np.float64) * (n_runs / tot)).astype(np.float32), D["iw"])
    dm = None
    if dither:
        dbits = rng.random((m, n_runs)) < rng.uniform(0.0, 0.3, m)[:, None]
        dm = np.packbits(dbits.astype(np.uint8), axis=1, bitorder="little").view(np.uint64).reshape(m, words)
    oc = o.Catalog(D["x"], D["y"], D["z"], D["nx"], D["ny"], D["nz"], w, D["jk"], iw=D["iw"], masks=masks, dither=dm)
    th = 0.03
    ang = (0.0, float(1.0 - np.cos(th)), 16, 1.6)
    oang = o.axis(*ang)
    up = np.ascontiguousarray(1.0 + 0.1 * rng.standard_normal(16))
    mono = [(0.0, 40.0, 15, 1.3)]
    with L.Context(0) as ctx:
        dc = ctx.upload(D["x"], D["y"], D["z"], D["nx"], D["ny"], D["nz"], w, D["jk"], iw=D["iw"], masks=masks,
                        dither=dm)
        jobs = [dict(kind=L.ANGULAR, a=dc, axes=[ang], n_jk=2, weight_mode=L.W_ANG_INVP, n_runs=n_runs,
                     dither=dither),
                dict(kind=L.MONOPOLE, a=dc, axes=mono, n_jk=2, weight_mode=L.W_INVP_3D, n_runs=n_runs, up_bins=up,
                     up_axis=ang, dither=dither)]
        d = int(dither)
        want = [o.count(3, oc, None, oang, None, 2, wopt=o.WOpt(L.W_ANG_INVP, n_runs, None, o.Axis(), 0, d)),
                o.count(0, oc, None, o.axis(*mono[0]), None, 2,
                        wopt=o.WOpt(L.W_INVP_3D, n_runs, up.ctypes.data, oang, 0, d))]
        for method in (0, 1):
            print(f"[gpu] popcount rules words={words} dither={dither} method={method}", flush=True)
            res, _ = ctx.count(jobs, method=method)
            for r, (full, sub, acc) in zip(res, want):
                assert r["n_pairs"] == acc
                np.testing.assert_allclose(r["full"], full, rtol=1e-12, atol=0)
                np.testing.assert_allclose(r["jk"], sub, rtol=1e-12, atol=0)
