"""GPU parity of the inverse-p pipeline against the reference built here with
_USE_INV_WEIGHTS (2 and 32 mask words) and _USE_DITHER_WEIGHTS (3 words),
from the fixtures of tests/golden/make_invp_golden.py:

- through the C ABI, both methods: the four angular histograms
  (correlators_invpweight.cpp:713-983), the up-weight tables formed from
  them (hist.cpp:224-231), and the 3-D DD with inverse-p weights, DR with
  the up-weight and RR (correlators_invpweight.cpp:14-700) given the
  reference's own up-weight tables, NaN bins included (on the 32-word case
  the first angular bins are empty, 0/0);
- the inverse-p driver (twopcf_invp) on the same hdf5 catalogues and
  parameter file as the reference's own main.cpp: every output file the
  reference wrote (the angular hdf5 files, ascii and hdf5 correlation
  functions) equals ours: ascii byte for byte, hdf5 in groups, datasets,
  shapes, element sizes and attribute bytes, with values to 1e-12.

Positions and weights at the ABI are the catalogue as the reference's
reader left it (fixture dumps), so the ABI tests check the kernels alone;
the driver tests add the hdf5 reader and its IIP weights.

Pair counts are exact; weighted sums agree to 1e-12 relative (the GPU adds
them in another order; the largest difference seen is printed)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import ROOT
from invp_cases import ANGULAR, CASES, InvpCase

pytestmark = pytest.mark.gpu

RTOL = 1e-12
KIND = {"monopole": 0, "sigma_pi": 1, "s_mu": 2}
DRIVER_INVP = os.path.join(ROOT, "two_pcf_amd", "twopcf_invp")
DRIVER = os.path.join(ROOT, "two_pcf_amd", "twopcf")


def close(got, want, what, rtol=RTOL):
    """NaN in the same bins, the rest within rtol; returns the largest
    relative difference."""
    got, want = np.asarray(got, np.float64), np.asarray(want, np.float64)
    assert got.shape == want.shape, what
    assert np.array_equal(np.isnan(got), np.isnan(want)), f"{what}: NaN bins differ"
    ok = ~np.isnan(want)
    np.testing.assert_allclose(got[ok], want[ok], rtol=rtol, atol=0, err_msg=what)
    d = np.abs(got[ok] - want[ok]) / np.maximum(np.abs(want[ok]), 1e-300)
    return float(d.max()) if d.size else 0.0


def upload(ctx, c, tag):
    g = c.g
    cols = c.columns(tag)
    f = lambda k: np.ascontiguousarray(g[f"cat_{tag}_{k}"])
    return ctx.upload(f("x"), f("y"), f("z"), f("nx"), f("ny"), f("nz"), f("w").astype(np.float32), cols[4],
                      iw=f("iw").astype(np.float32), masks=cols[5], dither=cols[6])


@pytest.mark.parametrize("method", [0, 1], ids=["fp64", "fp32"])
@pytest.mark.parametrize("name", CASES)
def test_abi_against_reference(name, method):
    from two_pcf_amd import lib as L
    c = InvpCase(name)
    g = c.g
    print(f"[gpu] inverse-p golden {name} method={method}", flush=True)
    ang = c.angular_axis()
    worst = 0.0
    with L.Context(0) as ctx:
        D, R = upload(ctx, c, "data"), upload(ctx, c, "rand")
        spec = dict(DD=(D, None, L.W_ANG_IW), DDinvp=(D, None, L.W_ANG_INVP), DR=(D, R, L.W_ANG_IW),
                    DRinvp=(D, R, L.W_ANG_W))
        jobs = [dict(kind=L.ANGULAR, a=a, b=b, axes=[ang], n_jk=c.n_jk, weight_mode=m, n_runs=c.n_runs,
                     dither=c.dither and m == L.W_ANG_INVP) for a, b, m in spec.values()]
        res, _ = ctx.count(jobs, method=method)
        fin = {}
        for k, r in zip(ANGULAR, res):
            fin[k] = L.finish(spec[k][1] is None, r["full"], r["jk"])
            worst = max(worst, close(fin[k][0], g[f"ang_{k}_full"], f"{name} angular {k} full"),
                        close(fin[k][1], g[f"ang_{k}_sub"], f"{name} angular {k} sub"))
        with np.errstate(invalid="ignore", divide="ignore"):
            close(fin["DD"][0] / fin["DDinvp"][0], g["up_dd_bins"], f"{name} up_dd")
            close(fin["DR"][0] / fin["DRinvp"][0], g["up_dr_bins"], f"{name} up_dr")
        up_dd, up_dr = g["up_dd_bins"], g["up_dr_bins"]
        for corr in c.corrs():
            axes = c.axes(corr)
            kind = KIND[corr]
            spec3 = dict(DD=(D, None, L.W_INVP_3D, up_dd), DR=(D, R, L.W_ANGUP_3D, up_dr), RR=(R, None, L.W_PLAIN, None))
            jobs = [dict(kind=kind, a=a, b=b, axes=axes, n_jk=c.n_jk, weight_mode=m, n_runs=c.n_runs, up_bins=u,
                         up_axis=ang if u is not None else None, dither=c.dither and m == L.W_INVP_3D)
                    for a, b, m, u in spec3.values()]
            res3, _ = ctx.count(jobs, method=method)
            for t, r in zip(("DD", "DR", "RR"), res3):
                full, sub = L.finish(spec3[t][1] is None, r["full"], r["jk"])
                worst = max(worst, close(full, g[f"{corr}_{t}_full"], f"{name} {corr} {t} full"),
                            close(sub, g[f"{corr}_{t}_sub"], f"{name} {corr} {t} sub"))
    print(f"[gpu] {name} method={method}: largest relative difference of the sums {worst:.3g}", flush=True)


def h5_same(got_path, want_path, what):
    """Two hdf5 files with the same groups, datasets (shapes, element sizes),
    attributes (stored bytes) and values (to RTOL, NaN in the same places)."""
    import h5mini
    tg, tw = h5mini.tree(got_path), h5mini.tree(want_path)
    assert list(tg) == list(tw), f"{what}: groups {list(tg)} != {list(tw)}"
    for grp in tw:
        dg, dw = tg[grp]["datasets"], tw[grp]["datasets"]
        assert list(dg) == list(dw), f"{what}/{grp}: datasets differ"
        for n in dw:
            assert dg[n][1:] == dw[n][1:], f"{what}/{grp}/{n}: shape or element size differs"
            close(dg[n][0], dw[n][0], f"{what}/{grp}/{n}")
        assert tg[grp]["attrs"] == tw[grp]["attrs"], f"{what}/{grp}: attributes differ"


def write_inputs(c, d):
    import h5mini
    for tag, fn in (("data", "data.h5"), ("rand", "rand.h5")):
        h5mini.write(os.path.join(d, fn), c.inputs[tag])
    with open(os.path.join(d, "params.txt"), "w") as f:
        f.write(str(c.g["params"]))


@pytest.mark.parametrize("method", ["fp64", "fp32"])
@pytest.mark.parametrize("name", CASES)
def test_driver_against_reference_files(name, method, tmp_path):
    import h5mini
    if h5mini.lib() is None:
        pytest.skip("no libhdf5")
    c = InvpCase(name)
    print(f"[gpu] twopcf_invp golden {name} {method}", flush=True)
    write_inputs(c, str(tmp_path))
    env = dict(os.environ, TWOPCF_METHOD=method, TWOPCF_DITHER="1" if c.dither else "0")
    r = subprocess.run([DRIVER_INVP, "params.txt"], cwd=tmp_path, env=env, capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr
    files = [k[5:] for k in c.g if k.startswith("file_")]
    assert files
    want_dir = tmp_path / "reference"
    want_dir.mkdir()
    for fn in files:
        data = c.g["file_" + fn].tobytes()
        if fn.endswith(".txt"):
            got = (tmp_path / fn).read_bytes()
            assert got == data, f"{name}/{fn} differs from the reference's file"
        else:
            (want_dir / fn).write_bytes(data)
            h5_same(str(tmp_path / fn), str(want_dir / fn), f"{name}/{fn}")


@pytest.mark.parametrize("method", ["fp64", "fp32"])
def test_driver_hdf5_output_against_reference_files(method, tmp_path):
    """The plain driver with hdf5 output on the cart_w_jk3 golden case: the
    files equal those of the reference built with _HAVE_HDF5 (io.cpp:653-975):
    groups, datasets, shapes, element sizes, attributes and values."""
    import h5mini
    from conftest import GOLDEN, load_golden
    if h5mini.lib() is None:
        pytest.skip("no libhdf5")
    g = load_golden("hdf5_out")
    print(f"[gpu] twopcf hdf5 output golden {method}", flush=True)
    src = os.path.join(GOLDEN, "driver", "cart_w_jk3")
    for fn in ("data.txt", "rand.txt"):
        shutil.copy(os.path.join(src, fn), tmp_path)
    (tmp_path / "params.txt").write_text(str(g["params"]))
    env = dict(os.environ, TWOPCF_METHOD=method)
    r = subprocess.run([DRIVER, "params.txt"], cwd=tmp_path, env=env, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    want_dir = tmp_path / "reference"
    want_dir.mkdir()
    for k in g:
        if not k.startswith("file_"):
            continue
        fn = k[5:]
        (want_dir / fn).write_bytes(g[k].tobytes())
        h5_same(str(tmp_path / fn), str(want_dir / fn), fn)
