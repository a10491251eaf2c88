"""GPU parity: libtwopcf_hip against the reference's golden vectors and the
oracle, through the C ABI.  Counts must be bit-identical; weighted sums whose
partial sums are exactly representable (weights in [0.5, 1.5]) must be too."""
import sys

import numpy as np
import pytest

from cases import load_case

pytestmark = pytest.mark.gpu

KIND = {"monopole": 0, "sigma_pi": 1, "s_mu": 2}
GOLDEN = ["cart_uniform", "eq_clustered", "edge", "uniform_monopole_50k", "cart_midpoint", "eq_midpoint"]


@pytest.fixture(scope="module")
def ctx():
    from two_pcf_amd import lib as L
    c = L.Context(0)
    yield c
    c.close()


def lib_cats(L, ctx, inputs, coord, om=0.25, h=1.0, z_min=0.0, z_max=1.0):
    cats = {}
    for tag, a in inputs.items():
        if coord == "cartesian":
            x, y, z = (np.asarray(a[k], np.float64) for k in "xyz")
            _, nx, ny, nz = L.cart_norm(x, y, z)
        else:
            d = L.comoving_distance(om, h, z_min, z_max, a["z"])
            x, y, z, nx, ny, nz = L.eq_to_cart(a["ra"], a["dec"], d)
        cats[tag] = ctx.upload(x, y, z, nx, ny, nz, a["w"], a["jk"])
    return cats


@pytest.mark.parametrize("method", [0, 1])
@pytest.mark.parametrize("name", GOLDEN)
def test_golden(ctx, name, method):
    from two_pcf_amd import lib as L
    case = load_case(name)
    print(f"[gpu] golden {name} method={method}", flush=True)
    cats = lib_cats(L, ctx, case.inputs, case.coord)
    for corr in case.corrs():
        axes = case.axes(corr)
        jobs = [dict(kind=KIND[corr], a=cats["data"], b=None, axes=axes, n_jk=case.n_jk, pi_def=case.pi_def),
                dict(kind=KIND[corr], a=cats["data"], b=cats["randoms"], axes=axes, n_jk=case.n_jk, pi_def=case.pi_def),
                dict(kind=KIND[corr], a=cats["randoms"], b=None, axes=axes, n_jk=case.n_jk, pi_def=case.pi_def)]
        res, _ = ctx.count(jobs, method=method)
        for h, r, j in zip(("DD", "DR", "RR"), res, jobs):
            full, sub = L.finish(j["b"] is None, r["full"], r["jk"])
            gf, gs = case.golden(corr, h)
            np.testing.assert_array_equal(full, gf, err_msg=f"{name} {corr} {h} full")
            np.testing.assert_array_equal(sub, gs, err_msg=f"{name} {corr} {h} sub")
    for c in cats.values():
        c.free()


def _oracle_cat(o, L, a):
    return o.Catalog(a["x"], a["y"], a["z"], a["nx"], a["ny"], a["nz"], a["w"], a["jk"])


def _jk_batch(scale, seed):
    from two_pcf_amd import lib as L
    from two_pcf_amd import synth
    c = synth.jackknife_batch(scale=scale, seed=seed)
    out = {}
    for tag in ("data", "randoms"):
        a = c[tag]
        d = L.comoving_distance(0.25, 1.0, 0.0, 1.0, a["z"])
        x, y, z, nx, ny, nz = L.eq_to_cart(a["ra"], a["dec"], d)
        out[tag] = dict(x=x, y=y, z=z, nx=nx, ny=ny, nz=nz, w=a["w"], jk=a["jk"])
    return out, c["n_jk"]


@pytest.mark.parametrize("method,slots", [(0, 0), (1, 0), (1, 1), (0, 2)])
def test_jackknife_batch_vs_oracle(ctx, oracle_lib, method, slots):
    """100 jackknife regions, s-mu 40x50, 1% of the 2M+10M batch.  slots=1/2
    forces most regions off the LDS slots onto the global-atomic path."""
    from two_pcf_amd import lib as L
    o = oracle_lib
    cats, n_jk = _jk_batch(0.01, 7)
    print(f"[gpu] jk batch method={method} slots={slots} nd={cats['data']['x'].size} nr={cats['randoms']['x'].size}", flush=True)
    axes = [(0.0, 100.0, 40, 1.3), (0.0, 1.0, 50, 0.0)]
    dc = {t: ctx.upload(a["x"], a["y"], a["z"], a["nx"], a["ny"], a["nz"], a["w"], a["jk"]) for t, a in cats.items()}
    oc = {t: _oracle_cat(o, L, a) for t, a in cats.items()}
    ax1, ax2 = o.axis(*axes[0]), o.axis(*axes[1])
    jobs = [dict(kind=2, a=dc["data"], b=None, axes=axes, n_jk=n_jk),
            dict(kind=2, a=dc["data"], b=dc["randoms"], axes=axes, n_jk=n_jk),
            dict(kind=2, a=dc["randoms"], b=None, axes=axes, n_jk=n_jk)]
    res, tm = ctx.count(jobs, method=method, tuning={"max_slots": slots} if slots else None)
    for r, (a, b) in zip(res, (("data", None), ("data", "randoms"), ("randoms", None))):
        full, sub, acc = o.count(2, oc[a], oc[b] if b else None, ax1, ax2, n_jk)
        np.testing.assert_array_equal(r["full"], full)
        np.testing.assert_array_equal(r["jk"], sub)
        assert r["n_pairs"] == acc == int(full.sum())
    for c in dc.values():
        c.free()


@pytest.mark.parametrize("weighted", [False, True])
def test_regions_in_slabs(ctx, oracle_lib, weighted):
    """Regions as slabs of x, 100 Mpc/h wide: most cells lie in one region
    (the batched method then counts every pair of their items in the full
    histogram, which is also that region's touching count) and the cells at
    a boundary straddle two.  s-mu and monopole, DD and DR, unweighted and
    weighted, both methods, with the default region slots and with one slot
    (cross pairs of one-region items then go to regions without a slot),
    against the oracle bit for bit."""
    from two_pcf_amd import lib as L
    o = oracle_lib
    rng = np.random.default_rng(11)

    def cat(n):
        x, y, z = (rng.uniform(0.0, 400.0, size=(3, n)) + np.array([[1000.0], [-300.0], [200.0]]))
        r = np.sqrt(x * x + y * y + z * z)
        w = rng.uniform(0.5, 1.5, n).astype(np.float32) if weighted else np.ones(n, np.float32)
        jk = np.floor((x - 1000.0) / 100.0).clip(0, 3).astype(np.int32)
        return dict(x=x, y=y, z=z, nx=x / r, ny=y / r, nz=z / r, w=w, jk=jk)

    cats = {"data": cat(15000), "randoms": cat(30000)}
    dc = {t: ctx.upload(a["x"], a["y"], a["z"], a["nx"], a["ny"], a["nz"], a["w"] if weighted else None, a["jk"])
          for t, a in cats.items()}
    oc = {t: _oracle_cat(o, L, a) for t, a in cats.items()}
    spec = (("data", None), ("data", "randoms"))
    for kind, axes in ((2, [(0.0, 100.0, 40, 1.3), (0.0, 1.0, 50, 0.0)]), (0, [(0.0, 100.0, 30, 1.3)])):
        ax1 = o.axis(*axes[0])
        ax2 = o.axis(*axes[1]) if len(axes) > 1 else None
        want = [o.count(kind, oc[a], oc[b] if b else None, ax1, ax2, 4) for a, b in spec]
        for method, slots in ((0, None), (1, None), (1, 1)):
            print(f"[gpu] slab regions weighted={weighted} kind={kind} method={method} slots={slots}", flush=True)
            res, _ = ctx.count([dict(kind=kind, a=dc[a], b=dc[b] if b else None, axes=axes, n_jk=4) for a, b in spec],
                               method=method, tuning={"max_slots": slots} if slots else None)
            for r, w, (a, b) in zip(res, want, spec):
                np.testing.assert_array_equal(r["full"], w[0], err_msg=f"kind {kind} {a}-{b or a} full")
                np.testing.assert_array_equal(r["jk"], w[1], err_msg=f"kind {kind} {a}-{b or a} regions")
                assert r["n_pairs"] == w[2]
    for c_ in dc.values():
        c_.free()


def test_methods_agree_weighted_sigma_pi(ctx, oracle_lib):
    """Weighted sigma-pi, s-mu and monopole (weights U[0.5,1.5]), DD, DR and
    RR: both GPU methods equal the oracle bit for bit (partial sums are
    exactly representable)."""
    from two_pcf_amd import lib as L
    from two_pcf_amd import synth
    o = oracle_lib
    c = synth.equatorial_smu_sigma_pi(scale=0.01, seed=3)
    cats = {}
    for tag in ("data", "randoms"):
        a = c[tag]
        d = L.comoving_distance(0.25, 1.0, 0.0, 1.0, a["z"])
        x, y, z, nx, ny, nz = L.eq_to_cart(a["ra"], a["dec"], d)
        cats[tag] = dict(x=x, y=y, z=z, nx=nx, ny=ny, nz=nz, w=a["w"], jk=a["jk"])
    dc = {t: ctx.upload(a["x"], a["y"], a["z"], a["nx"], a["ny"], a["nz"], a["w"], a["jk"]) for t, a in cats.items()}
    oc = {t: _oracle_cat(o, L, a) for t, a in cats.items()}
    for kind, axes in ((1, [(0.0, 50.0, 50, 1.0), (0.0, 50.0, 50, 1.0)]), (2, [(0.0, 100.0, 40, 1.3), (0.0, 1.0, 50, 0.0)]),
                       (0, [(0.0, 100.0, 30, 1.3)])):
        ax1 = o.axis(*axes[0])
        ax2 = o.axis(*axes[1]) if len(axes) > 1 else None
        spec = (("data", None), ("data", "randoms"), ("randoms", None))
        want = [o.count(kind, oc[a], oc[b] if b else None, ax1, ax2, 1) for a, b in spec]
        for method in (0, 1):
            print(f"[gpu] weighted kind={kind} method={method}", flush=True)
            res, _ = ctx.count([dict(kind=kind, a=dc[a], b=dc[b] if b else None, axes=axes, n_jk=1) for a, b in spec],
                               method=method)
            for r, w, (a, b) in zip(res, want, spec):
                np.testing.assert_array_equal(r["full"], w[0], err_msg=f"kind {kind} {a}-{b or a}")
                assert r["n_pairs"] == w[2]
    for c_ in dc.values():
        c_.free()


def test_edge_cases(ctx):
    """Empty catalogues, a single object, duplicates only, and region labels
    out of range."""
    from two_pcf_amd import lib as L
    axes = [(0.0, 100.0, 40, 1.3), (0.0, 1.0, 50, 0.0)]
    z = np.zeros(0)
    empty = ctx.upload(z, z, z, z, z, z)
    one = ctx.upload([1.0], [2.0], [3.0], [0.26], [0.53], [0.8])
    dup = ctx.upload([5.0] * 10, [5.0] * 10, [5.0] * 10, [0.57] * 10, [0.57] * 10, [0.57] * 10)
    for method in (0, 1):
        res, _ = ctx.count([dict(kind=2, a=empty, b=None, axes=axes), dict(kind=2, a=one, b=None, axes=axes),
                            dict(kind=2, a=one, b=empty, axes=axes), dict(kind=2, a=dup, b=None, axes=axes),
                            dict(kind=2, a=dup, b=one, axes=axes)], method=method)
        for r in res[:4]:
            assert r["full"].sum() == 0 and r["n_pairs"] == 0
        assert res[4]["n_pairs"] == 10
    bad = ctx.upload([1.0, 2.0], [1.0, 2.0], [1.0, 2.0], [0.6] * 2, [0.6] * 2, [0.5] * 2, jk=[0, 5])
    with pytest.raises(L.TwopcfError):
        ctx.count([dict(kind=2, a=bad, b=None, axes=axes, n_jk=3)], method=1)


def test_graph_replays_match_eager(ctx, oracle_lib):
    """The batched method replays a captured hipGraph on repeated calls; every
    replay, with other GPU work and synchronisations in between, must equal
    the plain method's eager result and the oracle."""
    from two_pcf_amd import lib as L
    o = oracle_lib
    cats, n_jk = _jk_batch(0.005, 11)
    axes = [(0.0, 100.0, 40, 1.3), (0.0, 1.0, 50, 0.0)]
    dc = {t: ctx.upload(a["x"], a["y"], a["z"], a["nx"], a["ny"], a["nz"], a["w"], a["jk"]) for t, a in cats.items()}
    jobs = [dict(kind=2, a=dc["data"], b=None, axes=axes, n_jk=n_jk),
            dict(kind=2, a=dc["data"], b=dc["randoms"], axes=axes, n_jk=n_jk),
            dict(kind=2, a=dc["randoms"], b=None, axes=axes, n_jk=n_jk)]
    eager, _ = ctx.count(jobs, method=0)
    oc = {t: _oracle_cat(o, L, a) for t, a in cats.items()}
    want = o.count(2, oc["data"], oc["randoms"], o.axis(*axes[0]), o.axis(*axes[1]), n_jk)
    np.testing.assert_array_equal(eager[1]["full"], want[0])
    for rep in range(4):
        print(f"[gpu] graph replay {rep}", flush=True)
        res, t = ctx.count(jobs, method=1)
        for a, b in zip(res, eager):
            np.testing.assert_array_equal(a["full"], b["full"])
            np.testing.assert_array_equal(a["jk"], b["jk"])
            assert a["n_pairs"] == b["n_pairs"]
        ctx.count([dict(kind=0, a=dc["data"], b=None, axes=[(0.0, 50.0, 10, 1.0)])], method=0)  # eager work between
    for c in dc.values():
        c.free()


@pytest.mark.parametrize("subdiv", [1, 3, 5])
def test_grid_independence(ctx, subdiv):
    """The batched method's cell size only changes which candidates are
    examined: every subdivision gives the reference's counts."""
    from two_pcf_amd import lib as L
    case = load_case("eq_clustered")
    print(f"[gpu] grid independence subdiv={subdiv}", flush=True)
    cats = lib_cats(L, ctx, case.inputs, case.coord)
    for corr in ("monopole", "sigma_pi", "s_mu"):
        if corr not in case.corrs():
            continue
        axes = case.axes(corr)
        jobs = [dict(kind=KIND[corr], a=cats["data"], b=None, axes=axes, n_jk=case.n_jk),
                dict(kind=KIND[corr], a=cats["data"], b=cats["randoms"], axes=axes, n_jk=case.n_jk)]
        res, _ = ctx.count(jobs, method=1, tuning={"cell_subdiv": subdiv})
        for h, r, j in zip(("DD", "DR"), res, jobs):
            full, sub = L.finish(j["b"] is None, r["full"], r["jk"])
            gf, gs = case.golden(corr, h)
            np.testing.assert_array_equal(full, gf, err_msg=f"subdiv {subdiv} {corr} {h} full")
            np.testing.assert_array_equal(sub, gs, err_msg=f"subdiv {subdiv} {corr} {h} sub")
    for c in cats.values():
        c.free()


def test_autotune_keeps_counts(ctx, oracle_lib):
    """The autotuner's choice is used by later calls without a tuning and
    gives the same counts as the defaults and the plain method."""
    from two_pcf_amd import lib as L
    cats, n_jk = _jk_batch(0.005, 13)
    axes = [(0.0, 100.0, 40, 1.3), (0.0, 1.0, 50, 0.0)]
    dc = {t: ctx.upload(a["x"], a["y"], a["z"], a["nx"], a["ny"], a["nz"], a["w"], a["jk"]) for t, a in cats.items()}
    jobs = [dict(kind=2, a=dc["data"], b=None, axes=axes, n_jk=n_jk),
            dict(kind=2, a=dc["data"], b=dc["randoms"], axes=axes, n_jk=n_jk),
            dict(kind=2, a=dc["randoms"], b=None, axes=axes, n_jk=n_jk)]
    print("[gpu] autotune", flush=True)
    tu, est = ctx.autotune(jobs)
    assert 1 <= tu["cell_subdiv"] <= 3 and 1 <= tu["max_slots"] <= 16 and tu["i_chunk"] in (128, 256)
    assert est > 0
    tuned, _ = ctx.count(jobs, method=1)
    dflt, _ = ctx.count(jobs, method=1, tuning={})
    plain, _ = ctx.count(jobs, method=0)
    for a, b, c in zip(tuned, dflt, plain):
        np.testing.assert_array_equal(a["full"], c["full"])
        np.testing.assert_array_equal(a["jk"], c["jk"])
        np.testing.assert_array_equal(b["full"], c["full"])
        np.testing.assert_array_equal(b["jk"], c["jk"])
    for c in dc.values():
        c.free()


@pytest.mark.parametrize("subdiv", [1, 2, 3])
def test_first_side_chunk_sizes(ctx, subdiv):
    """Work items of at most 64, 128, 192 and 256 first-side objects: the
    batched method deals an item's row blocks of 16 objects to 1, 2 or 4
    wave groups, partial items included (cell occupancies vary with the
    cell subdivision); its counts, with 100 regions, equal the plain
    method's for DD, DR and RR."""
    from two_pcf_amd import lib as L
    cats, n_jk = _jk_batch(0.01, 17)
    axes = [(0.0, 100.0, 40, 1.3), (0.0, 1.0, 50, 0.0)]
    dc = {t: ctx.upload(a["x"], a["y"], a["z"], a["nx"], a["ny"], a["nz"], a["w"], a["jk"]) for t, a in cats.items()}
    jobs = [dict(kind=L.S_MU, a=dc["data"], b=None, axes=axes, n_jk=n_jk),
            dict(kind=L.S_MU, a=dc["data"], b=dc["randoms"], axes=axes, n_jk=n_jk),
            dict(kind=L.S_MU, a=dc["randoms"], b=None, axes=axes, n_jk=n_jk)]
    plain, tp = ctx.count(jobs, method=L.PLAIN_FP64)
    for chunk in (64, 128, 192, 256):
        print(f"[gpu] first-side chunk {chunk}, cell subdivision {subdiv}", flush=True)
        res, t = ctx.count(jobs, method=L.FP32_RECHECK, tuning={"cell_subdiv": subdiv, "i_chunk": chunk})
        assert t["n_pairs"] == tp["n_pairs"]
        for k, (a, b) in enumerate(zip(res, plain)):
            np.testing.assert_array_equal(a["full"], b["full"], err_msg=f"chunk {chunk} job {k}")
            np.testing.assert_array_equal(a["jk"], b["jk"], err_msg=f"chunk {chunk} job {k}")
    for c in dc.values():
        c.free()


@pytest.mark.parametrize("method", [0, 1])
def test_uneven_multi_catalogue_batch(ctx, oracle_lib, method):
    """Many independent cartesian mocks of uneven size, monopole and s-mu
    DD/DR/RR of every mock in one call (the batched method bins each
    monopole in the pass of the s-mu job on the same catalogues: one
    persistent launch): every job equals the oracle's count of that mock
    alone."""
    from two_pcf_amd import lib as L
    from two_pcf_amd import synth
    o = oracle_lib
    mocks = synth.uneven_mocks(n_mocks=6, n_min=300, n_max=3000, density=3e-4, seed=5)
    mono = [(0.0, 100.0, 30, 1.3)]
    smu = [(0.0, 100.0, 40, 1.3), (0.0, 1.0, 50, 0.0)]
    jobs, want = [], []
    dcs = []
    for m in mocks:
        dcat = {}
        for tag in ("data", "randoms"):
            p = m[tag]
            _, nx, ny, nz = L.cart_norm(p[:, 0], p[:, 1], p[:, 2])
            dcat[tag] = (p[:, 0].copy(), p[:, 1].copy(), p[:, 2].copy(), nx, ny, nz)
        up = {t: ctx.upload(*v) for t, v in dcat.items()}
        dcs.extend(up.values())
        oc = {t: o.Catalog(*v, np.ones(v[0].size, np.float32), np.zeros(v[0].size, np.int32)) for t, v in dcat.items()}
        for kind, axes in ((0, mono), (2, smu)):
            ax1 = o.axis(*axes[0])
            ax2 = o.axis(*axes[1]) if len(axes) > 1 else None
            for a, b in (("data", None), ("data", "randoms"), ("randoms", None)):
                jobs.append(dict(kind=kind, a=up[a], b=up[b] if b else None, axes=axes, n_jk=1))
                want.append(o.count(kind, oc[a], oc[b] if b else None, ax1, ax2, 1))
    print(f"[gpu] multi-catalogue batch method={method}: {len(jobs)} jobs", flush=True)
    res, t = ctx.count(jobs, method=method)
    if method == 1:
        assert t["n_launches"] == 1   # each monopole binned in its s-mu job's pass: one persistent launch
    for r, (full, _, acc) in zip(res, want):
        np.testing.assert_array_equal(r["full"], full)
        assert r["n_pairs"] == acc
    for c in dcs:
        c.free()


@pytest.mark.parametrize("kind", [0, 1, 2])
def test_dense_cells_multi_tile(ctx, oracle_lib, kind):
    """Dense cells (about 1500 objects per cell of side rmax/2): every
    stencil row spans several 1024-entry tiles, every cell several
    256-object work items, and an auto count's own cell more than one tile,
    so the tile loop and the auto-count order test run across tile
    boundaries.  Points sit 1000 Mpc/h from the origin; 4 regions; both
    methods equal the oracle."""
    from two_pcf_amd import lib as L
    o = oracle_lib
    rng = np.random.Generator(np.random.PCG64(11 + kind))
    axes = {0: [(0.0, 40.0, 20, 1.3)], 1: [(0.0, 30.0, 15, 0.0), (0.0, 25.0, 10, 0.0)],
            2: [(0.0, 40.0, 20, 1.3), (0.0, 1.0, 10, 0.0)]}[kind]
    cats, dc, oc = {}, {}, {}
    for tag, n in (("a", 12000), ("b", 8000)):
        p = 1000.0 + rng.uniform(0.0, 40.0, (n, 3))
        x, y, z = (np.ascontiguousarray(p[:, k]) for k in range(3))
        _, nx, ny, nz = L.cart_norm(x, y, z)
        jk = ((x > 1020.0).astype(np.int32) + 2 * (y > 1020.0).astype(np.int32))
        w = np.ones(n, np.float32)
        dc[tag] = ctx.upload(x, y, z, nx, ny, nz, w, jk)
        oc[tag] = o.Catalog(x, y, z, nx, ny, nz, w, jk)
    oax = [o.axis(*a) for a in axes]
    want = [o.count(kind, oc["a"], None, oax[0], oax[-1] if kind else None, 4),
            o.count(kind, oc["a"], oc["b"], oax[0], oax[-1] if kind else None, 4)]
    jobs = [dict(kind=kind, a=dc["a"], b=None, axes=axes, n_jk=4),
            dict(kind=kind, a=dc["a"], b=dc["b"], axes=axes, n_jk=4)]
    for method in (0, 1):
        print(f"[gpu] dense cells kind={kind} method={method}", flush=True)
        res, _ = ctx.count(jobs, method=method)
        for r, (full, sub, acc) in zip(res, want):
            assert r["n_pairs"] == acc
            np.testing.assert_array_equal(r["full"], full)
            np.testing.assert_array_equal(r["jk"], sub)
    for c in dc.values():
        c.free()


@pytest.mark.parametrize("kind", [0, 1, 2, 3])
def test_long_column_z_window(ctx, oracle_lib, kind):
    """A column 30 Mpc/h wide and 600 long along z, about 300 objects per
    cell: stencil rows along z hold more than one 1024-entry tile, so the
    screening's z window skips blocks at both ends of every row.  A third of
    the objects sit on a 0.5 Mpc/h lattice in z (many equal z values, sort
    ties), and pairs at every separation up to the maximum occur along z.
    Monopole, sigma-pi, s-mu and angular counts, auto and cross, 4 regions,
    both methods equal the oracle."""
    from two_pcf_amd import lib as L
    o = oracle_lib
    rng = np.random.Generator(np.random.PCG64(41 + kind))
    axes = {0: [(0.0, 40.0, 20, 1.3)], 1: [(0.0, 30.0, 15, 0.0), (0.0, 40.0, 10, 0.0)],
            2: [(0.0, 40.0, 20, 1.3), (0.0, 1.0, 10, 0.0)],
            3: [(0.0, float(1.0 - np.cos(0.015)), 12, 1.6)]}[kind]
    cats, dc, oc = {}, {}, {}
    for tag, n in (("a", 30000), ("b", 10000)):
        p = np.stack([rng.uniform(2000.0, 2030.0, n), rng.uniform(-15.0, 15.0, n),
                      rng.uniform(-300.0, 300.0, n)], 1)
        snap = rng.random(n) < 1.0 / 3.0
        p[snap, 2] = np.round(p[snap, 2] * 2.0) / 2.0
        x, y, z = (np.ascontiguousarray(p[:, k]) for k in range(3))
        _, nx, ny, nz = L.cart_norm(x, y, z)
        jk = ((z > 0.0).astype(np.int32) + 2 * (x > 2015.0).astype(np.int32))
        w = np.ones(n, np.float32)
        dc[tag] = ctx.upload(x, y, z, nx, ny, nz, w, jk)
        oc[tag] = o.Catalog(x, y, z, nx, ny, nz, w, jk)
    oax = [o.axis(*a) for a in axes]
    two = kind in (1, 2)
    want = [o.count(kind, oc["a"], None, oax[0], oax[-1] if two else None, 4),
            o.count(kind, oc["a"], oc["b"], oax[0], oax[-1] if two else None, 4)]
    jobs = [dict(kind=kind, a=dc["a"], b=None, axes=axes, n_jk=4),
            dict(kind=kind, a=dc["a"], b=dc["b"], axes=axes, n_jk=4)]
    for method in (0, 1):
        print(f"[gpu] long column kind={kind} method={method}", flush=True)
        res, _ = ctx.count(jobs, method=method)
        for r, (full, sub, acc) in zip(res, want):
            assert acc > 0
            assert r["n_pairs"] == acc
            np.testing.assert_array_equal(r["full"], full)
            np.testing.assert_array_equal(r["jk"], sub)
    for c in dc.values():
        c.free()


@pytest.mark.parametrize("mono", [(40, 1.3), (30, 1.3), (25, 1.6), (12, 1.0)], ids=["log40", "log30", "log25", "linear12"])
@pytest.mark.parametrize("name", ["cart_uniform", "eq_clustered", "edge", "cart_midpoint", "eq_midpoint"])
def test_fused_monopole_s_mu(ctx, oracle_lib, name, mono):
    """A monopole over the same catalogues, weights and s range as an s-mu job
    is binned in the s-mu job's pass (one launch for DD, DR and RR of both).
    Its bins are the s-mu's s bins (40 bins, base 1.3), the upper 30 of them
    (30 bins to the same s maximum), or differ from them (other base, or
    linear); each pair is proved in both histograms or rechecked in both; the golden
    cases bring clustered and equatorial catalogues, edge separations,
    weights and the pair-midpoint line of sight.  Every histogram equals the
    oracle's count of that job alone; the plain method takes no fusion."""
    from two_pcf_amd import lib as L
    o = oracle_lib
    case = load_case(name)
    cats = lib_cats(L, ctx, case.inputs, case.coord)
    oc = {}
    for tag, a in case.inputs.items():
        if case.coord == "cartesian":
            x, y, z = (np.asarray(a[k], np.float64) for k in "xyz")
            _, nx, ny, nz = L.cart_norm(x, y, z)
        else:
            d = L.comoving_distance(0.25, 1.0, 0.0, 1.0, a["z"])
            x, y, z, nx, ny, nz = L.eq_to_cart(a["ra"], a["dec"], d)
        oc[tag] = o.Catalog(x, y, z, nx, ny, nz, a["w"], np.zeros(x.size, np.int32))
    smu = case.axes("s_mu")
    mono_ax = [(smu[0][0], smu[0][1], mono[0], mono[1])]
    jobs, want = [], []
    for a, b in (("data", None), ("data", "randoms"), ("randoms", None)):
        for kind, axes in ((2, smu), (0, mono_ax)):
            jobs.append(dict(kind=kind, a=cats[a], b=cats[b] if b else None, axes=axes, n_jk=1, pi_def=case.pi_def))
            ax1 = o.axis(*axes[0])
            ax2 = o.axis(*axes[1]) if len(axes) > 1 else None
            want.append(o.count(kind, oc[a], oc[b] if b else None, ax1, ax2, 1,
                                wopt=o.WOpt(0, 0, None, o.Axis(), case.pi_def, 0)))
    for method in (0, 1):
        print(f"[gpu] fused monopole {name} {mono} method={method}", flush=True)
        res, t = ctx.count(jobs, method=method)
        if method == 1:
            assert t["n_launches"] == 1
        for k, (r, (full, _, acc)) in enumerate(zip(res, want)):
            assert r["n_pairs"] == acc, f"job {k}"
            np.testing.assert_allclose(r["full"], full, rtol=1e-12, atol=0, err_msg=f"job {k}")
    for c in cats.values():
        c.free()


def _eq_cats(L, ctx, o, case, unit_w=False, regions=True):
    """Library and oracle catalogues of an equatorial golden case: unit
    weights when unit_w; the case's regions, or all objects in region 0."""
    cats, oc = {}, {}
    for tag, a in case.inputs.items():
        d = L.comoving_distance(0.25, 1.0, 0.0, 1.0, a["z"])
        x, y, z, nx, ny, nz = L.eq_to_cart(a["ra"], a["dec"], d)
        w = np.ones(x.size, np.float32) if unit_w else a["w"]
        jk = a["jk"] if regions else np.zeros(x.size, np.int32)
        cats[tag] = ctx.upload(x, y, z, nx, ny, nz, w, jk)
        oc[tag] = o.Catalog(x, y, z, nx, ny, nz, w, jk)
    return cats, oc


def test_fused_monopole_near_lds_limit(ctx, oracle_lib):
    """Unweighted s-mu 120 x 50 with a 120-bin monopole on the same
    catalogues: the pair fuses (its two histograms with their lane copies
    fit in LDS) and, fused or not, the call succeeds in both methods; every
    histogram equals the oracle's count of that job alone."""
    from two_pcf_amd import lib as L
    o = oracle_lib
    case = load_case("eq_clustered")
    cats, oc = _eq_cats(L, ctx, o, case, unit_w=True, regions=False)
    smu = [(0.0, 60.0, 120, 1.05), (0.0, 1.0, 50, 0.0)]
    mono = [(0.0, 60.0, 120, 1.05)]
    jobs, want = [], []
    for a, b in (("data", None), ("data", "randoms"), ("randoms", None)):
        for kind, axes in ((2, smu), (0, mono)):
            jobs.append(dict(kind=kind, a=cats[a], b=cats[b] if b else None, axes=axes, n_jk=1))
            ax1 = o.axis(*axes[0])
            ax2 = o.axis(*axes[1]) if len(axes) > 1 else None
            want.append(o.count(kind, oc[a], oc[b] if b else None, ax1, ax2, 1))
    for method in (0, 1):
        print(f"[gpu] fused monopole near the LDS limit method={method}", flush=True)
        res, t = ctx.count(jobs, method=method)
        for k, (r, (full, _, acc)) in enumerate(zip(res, want)):
            assert r["n_pairs"] == acc, f"job {k}"
            np.testing.assert_array_equal(r["full"], full, err_msg=f"job {k}")
    for c in cats.values():
        c.free()


@pytest.mark.parametrize("kind,axes", [
    (2, [(0.0, 60.0, 100, 1.05), (0.0, 1.0, 50, 0.0)]),        # weighted s-mu, 5000 bins
    (1, [(0.0, 40.0, 200, 0.0), (0.0, 40.0, 100, 0.0)]),       # weighted sigma-pi, 20000 bins
], ids=["s_mu_5000", "sigma_pi_20000"])
def test_large_weighted_histograms(ctx, oracle_lib, kind, axes):
    """Weighted histograms with 9 regions too large to keep beside region
    slots in LDS (5000 bins x 8 B, and 20000 bins x 8 B = 160 KB, which no
    LDS holds): the library drops lane copies, then keeps the full histogram
    in the global pool and as many region slots as fit (none for the second),
    the other regions' sums going to the global pool.  DD, DR and RR equal
    the oracle's in both methods."""
    from two_pcf_amd import lib as L
    o = oracle_lib
    case = load_case("eq_clustered")
    cats, oc = _eq_cats(L, ctx, o, case)
    ax1 = o.axis(*axes[0])
    ax2 = o.axis(*axes[1])
    spec = (("data", None), ("data", "randoms"), ("randoms", None))
    jobs = [dict(kind=kind, a=cats[a], b=cats[b] if b else None, axes=axes, n_jk=case.n_jk) for a, b in spec]
    want = [o.count(kind, oc[a], oc[b] if b else None, ax1, ax2, case.n_jk) for a, b in spec]
    for method in (0, 1):
        print(f"[gpu] large weighted histogram kind={kind} nb={axes[0][2] * axes[1][2]} method={method}", flush=True)
        res, _ = ctx.count(jobs, method=method)
        for k, (r, (full, sub, acc)) in enumerate(zip(res, want)):
            assert r["n_pairs"] == acc, f"job {k}"
            np.testing.assert_allclose(r["full"], full, rtol=1e-12, atol=0, err_msg=f"job {k}")
            np.testing.assert_allclose(r["jk"], sub, rtol=1e-12, atol=0, err_msg=f"job {k}")
    for c in cats.values():
        c.free()
