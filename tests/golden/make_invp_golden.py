"""Write the inverse-p, dither and hdf5-output golden fixtures in tests/golden/
from the reference compiled here.

Run in the build container only (it needs oracle/_ref/ref_invp2,
ref_invp32, ref_dither3 and TWOPCF_h5, compiled from /root/reference/src by
oracle/Makefile with _HAVE_HDF5 and, for the first three, _USE_INV_WEIGHTS):

    python tests/golden/make_invp_golden.py [case ...]

For every case it generates seeded catalogues, writes them as hdf5 files the
way the reference's hdf5 reader takes them (io.cpp:417-590: RA, DEC, Z as
doubles, W as double, JK as int, the bit masks as BITWEIGHT0..k-1 and the
dither masks as DITHER0..k-1, one int64 dataset per mask word), writes a
complete parameter file, and runs, single-threaded:

  ref_<variant> params out.json probes  -> the catalogue as read (IIP weights),
        the four angular histograms, the up-weight tables, JKHists::interp at
        the probe values, and the raw and normalised 3-D DD (inverse-p), DR
        (up-weight) and RR histograms;
  ref_<variant> params --main           -> the reference's own output files
        (main.cpp:49-62): the angular hdf5 files and the correlation-function
        files, stored byte for byte.

<case>.npz keys:
  params                      parameter file text (file names relative)
  in_sha256                   SHA-256 of the generated catalogue columns
  in_<cat>_<col>              catalogue columns (small cases only)
  cat_<cat>_<f>               x y z nx ny nz d w iw as the reference read them
  ang_<DD|DDinvp|DR|DRinvp>_<full|sub>
  up_<dd|dr>_bins, up_<dd|dr>_interp, probes
  <corr>_<DD|DR|RR>_<full|sub>, <corr>_norm_<DD|DR|RR>_<full|sub>
  file_<name>                 bytes of each output file of the --main run

hdf5_out.npz holds the plain build's (TWOPCF_h5) hdf5 output files for the
driver golden case cart_w_jk3, for the layout of the hdf5 writers.
"""
import hashlib
import json
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))
from two_pcf_amd import synth  # noqa: E402
import h5mini  # noqa: E402

REF_DIR = os.path.join(ROOT, "oracle", "_ref")
F = np.float32

BASE = dict(
    data_file_type="hdf5", random_file_type="hdf5", coord_system="equatorial",
    ra_x_dataset_name="RA", dec_y_dataset_name="DEC", z_z_dataset_name="Z",
    weight_dataset_name="W", jk_dataset_name="JK", use_weights=1, n_threads=1,
    n_jk_regions=3, omega_m=0.25, h=1.0, z_min=0.0, z_max=1.0,
    plot_monopole=1, monopole_filename="mono.txt", monopole_output_type="ascii",
    monopole_log_base=1.3, monopole_min=0.0, monopole_max=40.0, monopole_n_bins=10,
    plot_sigma_pi=1, sigma_pi_filename="sp.txt", sigma_pi_output_type="ascii",
    sigma_log_base=1.0, sigma_min=0.0, sigma_max=30.0, sigma_n_bins=12,
    pi_log_base=1.0, pi_min=0.0, pi_max=30.0, pi_n_bins=10,
    plot_s_mu=1, s_mu_filename="smu.txt", s_mu_output_type="ascii",
    s_log_base=1.3, s_min=0.0, s_max=40.0, s_n_bins=10, mu_n_bins=5,
    calculate_angular_dd=1, calculate_angular_dd_invpweights=1,
    calculate_angular_dr=1, calculate_angular_dr_invpweights=1,
    angular_dd_filename="add.hdf5", angular_dd_invpweights_filename="addinv.hdf5",
    angular_dr_filename="adr.hdf5", angular_dr_invpweights_filename="adrinv.hdf5",
    theta_max=0.05, theta_n_bins=10, theta_log_base=1.6, healpix_order=5,
    n_bitwise_runs=128, bitwise_weight_dataset_name="BITWEIGHT", dither_weight_dataset_name="DITHER",
    data_filename="data.h5", random_filename="rand.h5",
)


def masks(rng, n, words, lo, hi):
    p = rng.uniform(lo, hi, n)
    bits = rng.random((n, words * 64)) < p[:, None]
    return np.packbits(bits.astype(np.uint8), axis=1, bitorder="little").view(np.uint64).reshape(n, words)


def catalogues(seed, n_par, mean, sigma, ra_l, dec_l, z_l, n_rand, words, n_jk, rand_masks, dither):
    """Clustered data and uniform randoms on an ra/dec/z footprint; data bit
    masks with per-object bit probability U[0.3, 0.95]; randoms' masks as
    rand_masks says ("ones": all bits set, the IIP weight unchanged, or
    "random": probability U[0.8, 1.0], some all-zero and some all-ones rows);
    sparse dither masks (probability U[0, 0.4]) when dither is set; weights
    U[0.5, 1.5] as doubles; regions by ra stripes."""
    ra, dec, z = synth.neyman_scott_equatorial(n_par, mean, sigma, ra_l, dec_l, z_l, seed=seed)
    rra, rdec, rz = synth.randoms_equatorial(n_rand, ra_l, dec_l, z, seed=seed + 1)
    rng = np.random.Generator(np.random.PCG64(seed + 2))
    out = {}
    for tag, (a, b, c) in (("data", (ra, dec, z)), ("rand", (rra, rdec, rz))):
        n = a.size
        if tag == "data":
            m = masks(rng, n, words, 0.3, 0.95)
            m[:3] = 0                              # total popcount 0: IIP weight from total = 1
        elif rand_masks == "ones":
            m = np.full((n, words), np.uint64(0xFFFFFFFFFFFFFFFF), dtype=np.uint64)
        else:
            m = masks(rng, n, words, 0.8, 1.0)
            m[:5] = np.uint64(0xFFFFFFFFFFFFFFFF)
            m[5:7] = 0
        cols = dict(RA=a, DEC=b, Z=c, W=rng.uniform(0.5, 1.5, n),
                    JK=np.minimum((n_jk * (a - ra_l[0]) / (ra_l[1] - ra_l[0])).astype(np.int32), n_jk - 1))
        for k in range(words):
            cols[f"BITWEIGHT{k}"] = m[:, k].view(np.int64)
        if dither:
            dm = masks(rng, n, words, 0.0, 0.4)
            for k in range(words):
                cols[f"DITHER{k}"] = dm[:, k].view(np.int64)
        out[tag] = cols
    return out


def probes_for(theta_max, n_bins, base):
    """Values at which JKHists::interp is evaluated: below, at and above the
    axis ends, every arithmetic bin centre and edge, and the values around
    each segment switch (where Axis::bin_num(float v) - 0.5 crosses an
    integer): float neighbours of the switch point and doubles between
    them."""
    vmax = 1.0 - np.cos(theta_max)
    vals = [-1e-3, 0.0, vmax, vmax * (1 + 1e-9), 1.0]
    if base > 1.02:
        ilb = float(F(1.0 / np.log(F(base))))
        ils = float(F((float(F(base)) ** n_bins - 1.0) / float(F(vmax))))
        edge = lambda b: (np.exp(b / ilb) - 1.0) / ils
    else:
        step = float(F(vmax)) / n_bins
        edge = lambda b: b * step
    for k in range(n_bins + 1):
        vals.append(edge(k))
        if k < n_bins:
            vals.append(0.5 * (edge(k) + edge(k + 1)))
            sw = edge(k + 0.5)
            f = F(sw)
            for d in range(-4, 5):
                g = f
                for _ in range(abs(d)):
                    g = np.nextafter(g, F(np.inf) if d > 0 else F(0))
                vals.append(float(g))
                vals.append(float(g) * (1 + 2e-8))
    return np.array([v for v in vals if np.isfinite(v)])


def run(variant, params, cats, name, probes, keep_inputs, extra_main_outputs=()):
    binary = os.path.join(REF_DIR, "ref_" + variant)
    p = dict(BASE)
    p.update(params)
    out = {}
    with tempfile.TemporaryDirectory() as td:
        for tag, fn in (("data", "data.h5"), ("rand", "rand.h5")):
            h5mini.write(os.path.join(td, fn), cats[tag])
        text = "".join(f"{k} = {v}\n" for k, v in p.items())
        with open(os.path.join(td, "params.txt"), "w") as f:
            f.write(text)
        np.savetxt(os.path.join(td, "probes.txt"), probes, fmt="%.17g")
        subprocess.run([binary, "params.txt", "out.json", "probes.txt"], check=True, cwd=td,
                       stdout=subprocess.DEVNULL)
        with open(os.path.join(td, "out.json")) as f:
            res = json.load(f)
        # the reference's own pipeline and output files (main.cpp:49-62)
        main_dir = os.path.join(td, "main")
        os.mkdir(main_dir)
        for fn in ("data.h5", "rand.h5", "params.txt"):
            shutil.copy(os.path.join(td, fn), main_dir)
        subprocess.run([binary, "params.txt", "--main"], check=True, cwd=main_dir, stdout=subprocess.DEVNULL)
        for fn in sorted(os.listdir(main_dir)):
            if fn in ("data.h5", "rand.h5", "params.txt"):
                continue
            with open(os.path.join(main_dir, fn), "rb") as f:
                out["file_" + fn] = np.frombuffer(f.read(), dtype=np.uint8)
    out["params"] = np.array(text)
    for tag in ("data", "rand"):
        for k in sorted(cats[tag]):
            if keep_inputs:
                out[f"in_{tag}_{k}"] = np.ascontiguousarray(cats[tag][k])
    out["in_sha256"] = np.array(sha256_of(cats))
    for tag, key in (("data", "data_cat"), ("rand", "randoms_cat")):
        for fld, v in res[key].items():
            out[f"cat_{tag}_{fld}"] = np.array([float.fromhex(s) for s in v])
    n_jk = max(1, int(p["n_jk_regions"]))
    for k, v in res["angular"].items():
        out[f"ang_{k}_full"] = np.array(v["full"], dtype=np.float64)
        out[f"ang_{k}_sub"] = np.array(v["sub"], dtype=np.float64).reshape(n_jk, -1)
    out["probes"] = np.array([float.fromhex(s) for s in res["probes"]])
    for k in ("up_dd", "up_dr"):
        out[f"{k}_bins"] = np.array(res[k]["bins"], dtype=np.float64)
        out[f"{k}_interp"] = np.array(res[k]["interp"], dtype=np.float64)
    for corr in ("monopole", "sigma_pi", "s_mu"):
        if corr not in res:
            continue
        r = res[corr]
        out[f"{corr}_n_used"] = np.array([r["n_data_used"], r["n_randoms_used"]])
        out[f"{corr}_seconds"] = np.array(r["seconds"])
        for t in ("DD", "DR", "RR"):
            out[f"{corr}_{t}_full"] = np.array(r[t]["full"], dtype=np.float64)
            out[f"{corr}_{t}_sub"] = np.array(r[t]["sub"], dtype=np.float64).reshape(n_jk, -1)
            out[f"{corr}_norm_{t}_full"] = np.array(r["norm"][t]["full"], dtype=np.float64)
            out[f"{corr}_norm_{t}_sub"] = np.array(r["norm"][t]["sub"], dtype=np.float64).reshape(n_jk, -1)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
    print(name, "angular", {k: v.sum() for k, v in out.items() if k.startswith("ang_") and k.endswith("_full")},
          "files", [k for k in out if k.startswith("file_")])


# ------------------------------------------------------------------ cases --

# the catalogues of each case, regenerated from their seeds by the tests too
CATS = dict(
    invp_w2=lambda: catalogues(41, 60, 25.0, 4.0, (180.0, 190.0), (10.0, 20.0), (0.1, 0.15), 3000, 2, 3,
                               "random", False),
    invp_w32_axis=lambda: catalogues(43, 40, 40.0, 5.0, (150.0, 170.0), (0.0, 20.0), (0.2, 0.3), 4000, 32, 1,
                                     "ones", False),
    dither_w3=lambda: catalogues(47, 50, 25.0, 4.0, (200.0, 210.0), (-5.0, 5.0), (0.1, 0.15), 2500, 3, 2,
                                 "random", True),
)


def sha256_of(cats):
    h = hashlib.sha256()
    for tag in ("data", "rand"):
        for k in sorted(cats[tag]):
            h.update(k.encode())
            h.update(np.ascontiguousarray(cats[tag][k]).tobytes())
    return h.hexdigest()


def case_invp_w2():
    """2 mask words (n_bitwise_runs 128), clustered data, randoms with random
    masks (their IIP weights change, and DR inverse-p uses them), 3 regions,
    all three 3-D correlators, log angular axis."""
    run("invp2", dict(), CATS["invp_w2"](), "invp_w2", probes_for(0.05, 10, 1.6), True)


def case_invp_w32_axis():
    """32 mask words (n_bitwise_runs 2048) and the angular axis of the
    bitwise inverse-p workload: theta < 0.175 rad, 40 log bins base 1.6,
    healpix order 5; s-mu 40 log bins base 1.3 to 100 x 50 mu and the
    monopole 30 log bins base 1.3 to 100; randoms with all-ones masks; no
    regions."""
    run("invp32", dict(n_jk_regions=0, theta_max=0.175, theta_n_bins=40, theta_log_base=1.6,
                       n_bitwise_runs=2048, plot_sigma_pi=0, monopole_max=100.0, monopole_n_bins=30,
                       s_max=100.0, s_n_bins=40, mu_n_bins=50, s_mu_output_type="hdf5", s_mu_filename="smu.h5",
                       monopole_output_type="hdf5", monopole_filename="mono.h5"),
        CATS["invp_w32_axis"](), "invp_w32_axis", probes_for(0.175, 40, 1.6), False)


def case_dither_w3():
    """_USE_DITHER_WEIGHTS with 3 mask words (n_bitwise_runs 192): sparse
    dither masks, so many pairs share no dither bit (0 counts as 1); linear
    angular axis; 2 regions."""
    run("dither3", dict(n_jk_regions=2, n_bitwise_runs=192, theta_log_base=1.0, theta_n_bins=8, theta_max=0.04,
                        plot_sigma_pi=0),
        CATS["dither_w3"](), "dither_w3", probes_for(0.04, 8, 1.0), True)


def case_hdf5_out():
    """The plain build with _HAVE_HDF5 (TWOPCF_h5) on the cart_w_jk3 driver
    case with every output in hdf5: the files the reference's hdf5 writers
    produce (io.cpp:653-975)."""
    src = os.path.join(HERE, "driver", "cart_w_jk3")
    text = open(os.path.join(src, "params.txt")).read()
    lines = []
    for line in text.splitlines():
        k = line.split("=", 1)[0].strip()
        if k in ("monopole_output_type", "sigma_pi_output_type", "s_mu_output_type"):
            line = f"{k} = hdf5"
        elif k in ("monopole_filename", "sigma_pi_filename", "s_mu_filename"):
            line = f"{k} = " + {"monopole_filename": "mono.h5", "sigma_pi_filename": "sp.h5",
                                "s_mu_filename": "smu.h5"}[k]
        lines.append(line)
    out = {"params": np.array("\n".join(lines) + "\n")}
    with tempfile.TemporaryDirectory() as td:
        for fn in ("data.txt", "rand.txt"):
            shutil.copy(os.path.join(src, fn), td)
        with open(os.path.join(td, "params.txt"), "w") as f:
            f.write(str(out["params"]))
        subprocess.run([os.path.join(REF_DIR, "TWOPCF_h5"), "params.txt"], check=True, cwd=td,
                       stdout=subprocess.DEVNULL)
        for fn in ("mono.h5", "sp.h5", "smu.h5"):
            with open(os.path.join(td, fn), "rb") as f:
                out["file_" + fn] = np.frombuffer(f.read(), dtype=np.uint8)
    np.savez_compressed(os.path.join(HERE, "hdf5_out.npz"), **out)
    print("hdf5_out", [k for k in out if k.startswith("file_")])


CASES = dict(invp_w2=case_invp_w2, invp_w32_axis=case_invp_w32_axis, dither_w3=case_dither_w3,
             hdf5_out=case_hdf5_out)

if __name__ == "__main__":
    for b in ("ref_invp2", "ref_invp32", "ref_dither3", "TWOPCF_h5"):
        if not os.path.exists(os.path.join(REF_DIR, b)):
            sys.exit(f"{b} missing: run make -C oracle ref")
    for name in sys.argv[1:] or list(CASES):
        CASES[name]()
