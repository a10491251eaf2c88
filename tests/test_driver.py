"""The parameter-file driver (two_pcf_amd/twopcf, main.cpp of the reference
restated on the GPU library) against the reference binary's own output files:
same parameter file, same ascii catalogues, byte-identical output."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT

DRIVER = os.path.join(ROOT, "two_pcf_amd", "twopcf")
CASES = sorted(os.listdir(os.path.join(GOLDEN, "driver")))


def run_driver(case, tmp_path, method):
    src = os.path.join(GOLDEN, "driver", case)
    for fn in ("params.txt", "data.txt", "rand.txt"):
        shutil.copy(os.path.join(src, fn), tmp_path)
    env = dict(os.environ, TWOPCF_METHOD=method)
    return subprocess.run([DRIVER, "params.txt"], cwd=tmp_path, env=env, capture_output=True, text=True, timeout=300)


@pytest.mark.gpu
@pytest.mark.parametrize("method", ["fp64", "fp32"])
@pytest.mark.parametrize("case", CASES)
def test_driver_outputs_identical(case, method, tmp_path):
    print(f"[gpu] driver {case} {method}", flush=True)
    r = run_driver(case, tmp_path, method)
    assert r.returncode == 0, r.stderr
    exp = os.path.join(GOLDEN, "driver", case, "expected")
    for fn in sorted(os.listdir(exp)):
        with open(os.path.join(exp, fn)) as f:
            want = f.read()
        with open(os.path.join(tmp_path, fn)) as f:
            got = f.read()
        assert got == want, f"{case}/{fn} differs"


def test_driver_fails_loudly_without_gpu(tmp_path):
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    r = run_driver(CASES[0], tmp_path, "fp32")
    assert r.returncode != 0
    assert "no usable GPU" in r.stderr


# ---------------------------------------------------------------- inverse-p --

DRIVER_INVP = os.path.join(ROOT, "two_pcf_amd", "twopcf_invp")
N_RUNS = 128


def _invp_inputs(seed=41, dither=False):
    """Equatorial clustered data with 2-word bit masks and weights, randoms
    with all-ones masks (their IIP weight is unchanged), 2 regions; with
    dither, sparse 2-word dither masks for both (many data pairs share no
    dither bit)."""
    import numpy as np
    from two_pcf_amd import synth
    rng = np.random.Generator(np.random.PCG64(seed))
    ra_l, dec_l, z_l = (180.0, 190.0), (10.0, 20.0), (0.1, 0.15)
    ra, dec, z = synth.neyman_scott_equatorial(60, 25.0, 4.0, ra_l, dec_l, z_l, seed=seed)
    rra, rdec, rz = synth.randoms_equatorial(3000, ra_l, dec_l, z, seed=seed + 1)
    out = {}
    for tag, (a, b, c) in (("data", (ra, dec, z)), ("rand", (rra, rdec, rz))):
        n = a.size
        if tag == "data":
            p = rng.uniform(0.3, 0.95, n)
            bits = rng.random((n, N_RUNS)) < p[:, None]
            masks = np.packbits(bits.astype(np.uint8), axis=1, bitorder="little").view(np.uint64).reshape(n, 2)
        else:
            masks = np.full((n, 2), np.uint64(0xFFFFFFFFFFFFFFFF), dtype=np.uint64)
        w = rng.uniform(0.5, 1.5, n).astype(np.float32)
        jk = (a > 185.0).astype(np.int32)
        out[tag] = dict(ra=a, dec=b, z=c, w=w, jk=jk, masks=masks)
        if dither:
            p = rng.uniform(0.0, 0.4, n)
            bits = rng.random((n, N_RUNS)) < p[:, None]
            out[tag]["dither"] = np.packbits(bits.astype(np.uint8), axis=1,
                                             bitorder="little").view(np.uint64).reshape(n, 2)
    return out


INVP_PARAMS = dict(
    data_file_type="ascii", random_file_type="ascii", coord_system="equatorial",
    use_weights=1, n_jk_regions=2, omega_m=0.25, h=1.0, z_min=0.0, z_max=1.0,
    plot_monopole=1, monopole_filename="mono.txt", monopole_output_type="ascii",
    monopole_log_base=1.3, monopole_min=0.0, monopole_max=40.0, monopole_n_bins=10,
    plot_sigma_pi=0, plot_s_mu=1, s_mu_filename="smu.txt", s_mu_output_type="ascii",
    s_log_base=1.3, s_min=0.0, s_max=40.0, s_n_bins=10, mu_n_bins=5,
    calculate_angular_dd=1, calculate_angular_dd_invpweights=1,
    calculate_angular_dr=1, calculate_angular_dr_invpweights=1,
    angular_dd_filename="add.hdf5", angular_dd_invpweights_filename="addinv.hdf5",
    angular_dr_filename="adr.hdf5", angular_dr_invpweights_filename="adrinv.hdf5",
    theta_max=0.03, theta_n_bins=6, theta_log_base=1.0, healpix_order=5,
    n_bitwise_runs=N_RUNS, bitwise_weight_dataset_name="BITWEIGHT",
    data_filename="data.txt", random_filename="rand.txt",
)


def _expected_invp(o, inp, angular=None):
    """The inverse-p pipeline from the oracle: positions (cosmology table,
    Galaxy::set), IIP weights (io.cpp:578-582, with dither masks
    io.cpp:566-572), angular DD/DR with parent and
    inverse-p weights, up-weights DD/DDinvp and DR/DRinvp (divide_by_hist),
    3-D DD (inverse-p), DR (up-weight), RR (plain), invert/x2, normalise and
    the LS estimator, formatted as the reference's ascii writer does."""
    import numpy as np
    p = INVP_PARAMS
    table = o.cosmo_table(p["omega_m"], p["h"], p["z_min"], p["z_max"])
    cats = {}
    for tag, a in inp.items():
        d = o.comov_dist(table, p["z_min"], a["z"])
        x, y, z, nx, ny, nz = o.eq_to_cart(a["ra"], a["dec"], d)
        tot = np.array([sum(int(int(v)).bit_count() for v in row) for row in a["masks"]])
        tot[tot == 0] = 1
        num = N_RUNS
        if "dither" in a:
            num = np.array([sum(int(int(v)).bit_count() for v in row) for row in a["dither"]])
        w = np.where(tot != N_RUNS, (a["w"].astype(np.float64) * (num / tot)).astype(np.float32), a["w"])
        cats[tag] = o.Catalog(x, y, z, nx, ny, nz, w, a["jk"], iw=a["w"], masks=a["masks"], dither=a.get("dither"))
    D, R = cats["data"], cats["rand"]
    dith = int("dither" in inp["data"])
    n_jk = p["n_jk_regions"]
    ang = o.axis(0.0, float(1.0 - np.cos(p["theta_max"])), p["theta_n_bins"], p["theta_log_base"])
    none = o.Axis()

    def fin(kind, a, b, ax1, ax2, wo):
        full, sub, _ = o.count(kind, a, b, ax1, ax2, n_jk, wopt=wo)
        return o.finish(b is None, full, sub)

    add = fin(3, D, None, ang, None, o.WOpt(3, N_RUNS, None, none))[0]
    addi = fin(3, D, None, ang, None, o.WOpt(4, N_RUNS, None, none, 0, dith))[0]
    adr = fin(3, D, R, ang, None, o.WOpt(3, N_RUNS, None, none))[0]
    adri = fin(3, D, R, ang, None, o.WOpt(5, N_RUNS, None, none))[0]
    if angular is not None:
        angular.update({"add.hdf5": add, "addinv.hdf5": addi, "adr.hdf5": adr, "adrinv.hdf5": adri})
    up_dd = np.ascontiguousarray(add / addi)
    up_dr = np.ascontiguousarray(adr / adri)
    files = {}
    for name, kind, axes in (("mono.txt", 0, [(0.0, 40.0, 10, 1.3)]),
                             ("smu.txt", 2, [(0.0, 40.0, 10, 1.3), (0.0, 1.0, 5, 0.0)])):
        ax1 = o.axis(*axes[0])
        ax2 = o.axis(*axes[1]) if len(axes) > 1 else None
        dd = fin(kind, D, None, ax1, ax2, o.WOpt(1, N_RUNS, up_dd.ctypes.data, ang, 0, dith))
        dr = fin(kind, D, R, ax1, ax2, o.WOpt(2, N_RUNS, up_dr.ctypes.data, ang))
        rr = fin(kind, R, None, ax1, ax2, None)
        dd = o.normalise_auto(*dd, D)
        dr = o.normalise_cross(*dr, D, R)
        rr = o.normalise_auto(*rr, R)
        L = o.lib()
        c1 = [L.or_axis_bin_centre(o.C.byref(ax1), k) for k in range(ax1.n_bins)]
        w1 = [L.or_axis_bin_width(o.C.byref(ax1), k) for k in range(ax1.n_bins)]
        if ax2 is not None:
            c2 = [L.or_axis_bin_centre(o.C.byref(ax2), k) for k in range(ax2.n_bins)]
            w2 = [L.or_axis_bin_width(o.C.byref(ax2), k) for k in range(ax2.n_bins)]

        def text(DD, DR, RR):
            rows = []
            for i in range(ax1.n_bins):
                js = range(ax2.n_bins) if ax2 is not None else [None]
                for j in js:
                    b = i if j is None else i * ax2.n_bins + j
                    xi = 0.0 if RR[b] == 0 else DD[b] / RR[b] - 2.0 * DR[b] / RR[b] + 1.0
                    if j is None:
                        vals = [i, c1[i], w1[i], xi, DD[b], DR[b], RR[b]]
                    else:
                        vals = [i, j, c1[i], c2[j], w1[i], w2[j], xi, DD[b], DR[b], RR[b]]
                    rows.append(vals)
            return rows

        files[name] = text(dd[0], dr[0], rr[0])
        stem, ext = name.rsplit(".", 1)
        for r in range(n_jk):
            files[f"{stem}{r}.{ext}"] = text(dd[1][r], dr[1][r], rr[1][r])
    return files


def _write_invp_inputs(tmp_path, inp):
    for tag, a in inp.items():
        with open(os.path.join(tmp_path, tag + ".txt"), "w") as f:
            for i in range(a["ra"].size):
                m = " ".join(str(int(v)) for v in a["masks"][i])
                if "dither" in a:
                    m += " " + " ".join(str(int(v)) for v in a["dither"][i])
                f.write("%.17g %.17g %.17g %.9g %d %s\n" % (a["ra"][i], a["dec"][i], a["z"][i], a["w"][i],
                                                          a["jk"][i], m))
    with open(os.path.join(tmp_path, "params.txt"), "w") as f:
        f.write("".join(f"{k} = {v}\n" for k, v in INVP_PARAMS.items()))


@pytest.mark.gpu
@pytest.mark.parametrize("dither", [False, True], ids=["bitwise", "dither"])
@pytest.mark.parametrize("method", ["fp64", "fp32"])
def test_driver_invp_pipeline(method, dither, tmp_path, oracle_lib):
    """twopcf_invp (the inverse-p build of the driver) on ascii catalogues with
    bit-mask columns (and dither-mask columns under TWOPCF_DITHER=1): every
    output file equals the oracle's pipeline to the 6 significant digits of
    the ascii writer (a last-digit difference is allowed: weighted sums are
    added in another order)."""
    import numpy as np
    print(f"[gpu] driver inverse-p {method} dither={dither}", flush=True)
    inp = _invp_inputs(dither=dither)
    _write_invp_inputs(tmp_path, inp)
    env = dict(os.environ, TWOPCF_METHOD=method, TWOPCF_DITHER="1" if dither else "0")
    r = subprocess.run([DRIVER_INVP, "params.txt"], cwd=tmp_path, env=env, capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr
    want = _expected_invp(oracle_lib, inp)
    for fn, rows in want.items():
        with open(os.path.join(tmp_path, fn)) as f:
            got = [[float(v) for v in line.split("\t")] for line in f.read().splitlines()]
        assert len(got) == len(rows), fn
        for g, w in zip(got, rows):
            np.testing.assert_allclose(g, w, rtol=1.5e-5, atol=0, err_msg=fn)


# --------------------------------------------------------------------- hdf5 --

def _fmt(v):
    """One value as the reference's ascii writer prints it (ostream default:
    %g with 6 significant digits; NaN with its sign)."""
    import math
    if isinstance(v, int):
        return str(v)
    if math.isnan(v):
        return "-nan" if math.copysign(1.0, v) < 0 else "nan"
    return "%g" % v


def _h5_as_text(path, group, two_d):
    """The rows output_xi_1D/2D would write in ascii, rebuilt from the hdf5
    file's datasets (axes from full_result, counts and xi from the group)."""
    import h5mini
    with h5mini.File(path) as f:
        g = {t: f.dataset(group, t)[0].reshape(-1) for t in ("DD", "DR", "RR", "xi")}
        if not two_d:
            c, w = f.dataset("full_result", "Bin_centre")[0], f.dataset("full_result", "Bin_width")[0]
            rows = [[i, c[i], w[i]] + [g[t][i] for t in ("xi", "DD", "DR", "RR")] for i in range(c.size)]
        else:
            c1, w1 = f.dataset("full_result", "Axis_1_bin_centre")[0], f.dataset("full_result", "Axis_1_bin_width")[0]
            c2, w2 = f.dataset("full_result", "Axis_2_bin_centre")[0], f.dataset("full_result", "Axis_2_bin_width")[0]
            rows = [[i, j, c1[i], c2[j], w1[i], w2[j]] + [g[t][i * c2.size + j] for t in ("xi", "DD", "DR", "RR")]
                    for i in range(c1.size) for j in range(c2.size)]
    return "".join("\t".join(_fmt(float(v)) if k >= (2 if two_d else 1) else str(int(v)) for k, v in enumerate(r))
                   + "\n" for r in rows)


def _fields(line):
    """The fields of a catalogue line as the ascii reader splits it: on ','
    if the line has one, else on tab if it has one, else on space; empty
    parts skipped, spaces removed."""
    line = line.rstrip("\r\n")
    sep = "," if "," in line else ("\t" if "\t" in line else " ")
    return [f.replace(" ", "") for f in line.split(sep) if f.replace(" ", "")]


def _params(path):
    out = {}
    for line in open(path):
        if "=" in line and not line.startswith("#"):
            k, v = line.split("#")[0].split("=", 1)
            out[k.strip()] = v.strip()
    return out


H5_NAMES = dict(ra_x_dataset_name="RA", dec_y_dataset_name="DEC", z_z_dataset_name="Z", weight_dataset_name="W",
                jk_dataset_name="JK", bitwise_weight_dataset_name="BITWEIGHT", dither_weight_dataset_name="DITHER")


@pytest.mark.gpu
@pytest.mark.parametrize("method", ["fp64", "fp32"])
@pytest.mark.parametrize("case", CASES)
def test_driver_hdf5_in_and_out(case, method, tmp_path):
    """The driver with hdf5 catalogues (io.cpp:417-520) and hdf5 output
    (io.cpp:653-975): the same catalogues as the golden case, converted to
    datasets (weights as the float values strtof gives), and every output
    group, printed as the ascii writer would, equals the reference binary's
    ascii file of that group byte for byte."""
    import h5mini
    from test_host_reader import strtof
    if h5mini.lib() is None:
        pytest.skip("no libhdf5")
    print(f"[gpu] driver hdf5 {case} {method}", flush=True)
    src = os.path.join(GOLDEN, "driver", case)
    p = _params(os.path.join(src, "params.txt"))
    use_w, n_jk = int(p["use_weights"]) != 0, int(p["n_jk_regions"])
    for tag in ("data", "rand"):
        rows = [_fields(l) for l in open(os.path.join(src, tag + ".txt")) if l.strip() and not l.startswith("#")]
        cols = {n: np.array([float(r[k]) for r in rows]) for k, n in enumerate(("RA", "DEC", "Z"))}
        k = 3
        if use_w:
            cols["W"] = np.array([strtof(r[k]) for r in rows])
            k += 1
        if n_jk > 1:
            cols["JK"] = np.array([int(r[k]) for r in rows], np.int32)
        h5mini.write(str(tmp_path / (tag + ".h5")), cols)
    q = dict(p, **H5_NAMES, data_file_type="hdf5", random_file_type="hdf5", data_filename="data.h5",
             random_filename="rand.h5")
    outs = {}
    for key, stem, two_d in (("monopole", "mono", False), ("sigma_pi", "sp", True), ("s_mu", "smu", True)):
        q[f"{key}_output_type"] = "hdf5"
        q[f"{key}_filename"] = stem + ".h5"
        outs[stem] = two_d
    (tmp_path / "params.txt").write_text("".join(f"{k} = {v}\n" for k, v in q.items()))
    env = dict(os.environ, TWOPCF_METHOD=method)
    r = subprocess.run([DRIVER, "params.txt"], cwd=tmp_path, env=env, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    exp = os.path.join(src, "expected")
    names = sorted(os.listdir(exp))
    assert names
    for fn in names:
        stem = fn[:-4].rstrip("0123456789")
        sub = fn[len(stem):-4]
        group = "full_result" if sub == "" else f"jk_reg{sub}"
        got = _h5_as_text(str(tmp_path / (stem + ".h5")), group, outs[stem])
        with open(os.path.join(exp, fn)) as f:
            assert got == f.read(), f"{case}/{fn} ({group}) differs"


@pytest.mark.gpu
@pytest.mark.parametrize("dither", [False, True], ids=["bitwise", "dither"])
def test_driver_invp_hdf5_handoff(dither, tmp_path, oracle_lib):
    """twopcf_invp on hdf5 catalogues with the bit masks in BITWEIGHT0, 1
    (and dither masks in DITHER0, 1), as the reference reads them
    (io.cpp:521-588).  Run 1 computes the four angular histograms, writes
    them to their HDF5 files (JKHists::output_hdf5) and the 3-D outputs as
    hdf5; run 2 computes no angular histogram and reads the files of run 1
    (read_hdf5_hist).  The angular files and both runs' outputs equal the
    oracle's pipeline to 1e-11 (weighted sums added in another order)."""
    import h5mini
    if h5mini.lib() is None:
        pytest.skip("no libhdf5")
    print(f"[gpu] driver inverse-p hdf5 dither={dither}", flush=True)
    inp = _invp_inputs(dither=dither)
    for tag, a in inp.items():
        cols = dict(RA=a["ra"], DEC=a["dec"], Z=a["z"], W=a["w"].astype(np.float64), JK=a["jk"].astype(np.int32))
        for k in range(a["masks"].shape[1]):
            cols[f"BITWEIGHT{k}"] = np.ascontiguousarray(a["masks"][:, k]).view(np.int64)
            if dither:
                cols[f"DITHER{k}"] = np.ascontiguousarray(a["dither"][:, k]).view(np.int64)
        h5mini.write(str(tmp_path / (tag + ".h5")), cols)
    angular = {}
    want = _expected_invp(oracle_lib, inp, angular)
    env = dict(os.environ, TWOPCF_METHOD="fp32", TWOPCF_DITHER="1" if dither else "0")
    base = dict(INVP_PARAMS, **H5_NAMES, data_file_type="hdf5", random_file_type="hdf5", data_filename="data.h5",
                random_filename="rand.h5", monopole_output_type="hdf5", s_mu_output_type="hdf5")
    for run, calc in ((1, 1), (2, 0)):
        q = dict(base, monopole_filename=f"mono_r{run}.h5", s_mu_filename=f"smu_r{run}.h5",
                 calculate_angular_dd=calc, calculate_angular_dd_invpweights=calc, calculate_angular_dr=calc,
                 calculate_angular_dr_invpweights=calc)
        (tmp_path / "params.txt").write_text("".join(f"{k} = {v}\n" for k, v in q.items()))
        r = subprocess.run([DRIVER_INVP, "params.txt"], cwd=tmp_path, env=env, capture_output=True, text=True,
                           timeout=300)
        assert r.returncode == 0, r.stderr
        if run == 1:
            for fn, full in angular.items():
                with h5mini.File(str(tmp_path / fn)) as f:
                    np.testing.assert_allclose(f.dataset("full_result", "hist")[0], full, rtol=1e-11, atol=0,
                                               err_msg=fn)
        else:
            assert "Reading file: add.hdf5" in r.stdout
        for fn, rows in want.items():
            stem = fn[:-4].rstrip("0123456789")
            sub = fn[len(stem):-4]
            group = "full_result" if sub == "" else f"jk_reg{sub}"
            with h5mini.File(str(tmp_path / f"{stem}_r{run}.h5")) as f:
                g = {t: f.dataset(group, t)[0].reshape(-1) for t in ("xi", "DD", "DR", "RR")}
            k0 = 6 if stem == "smu" else 3
            w = np.array(rows, dtype=np.float64)
            for k, t in enumerate(("xi", "DD", "DR", "RR")):
                np.testing.assert_allclose(g[t], w[:, k0 + k], rtol=1e-11, atol=1e-13, err_msg=f"run {run} {fn} {t}")
