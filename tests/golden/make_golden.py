"""Write the golden fixtures in tests/golden/ from the reference binary.

Run in the build container only (it needs oracle/_ref/ref_counts, which is
compiled from /root/reference/src by oracle/Makefile):

    python tests/golden/make_golden.py

For every case it generates seeded catalogues, writes them as ascii with
round-trip precision (the reference's own reader, io.cpp:240-415, parses
them), writes a complete parameter file (io.cpp:16-237 leaves missing keys
uninitialised, so every key is set), runs ref_counts single-threaded and
stores inputs and outputs in <case>.npz:

  in_<cat>_<col>          catalogue columns as read (x/y/z or ra/dec/z, w, jk)
  cat_<cat>_<field>       positions / unit vectors as the reference computed them
  <corr>_<DD|DR|RR>_full  histograms after invert_jk_hists / multiply_hists
  <corr>_<DD|DR|RR>_sub   (n_jk, nb)
  <corr>_norm_<..>_*      the same after normalise()
  <corr>_axis<k>          [min, max, step, inv_step, ilb, ils] as float32
  <corr>_axis<k>_meta     [n_bins, log_binning]
  params                  the parameter file text
"""
import hashlib
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
from two_pcf_amd import synth  # noqa: E402

REF = os.path.join(ROOT, "oracle", "_ref", "ref_counts")
# the reference built without _ANGULAR_PI_DEFN (pair-midpoint line of sight)
REF_MID = os.path.join(ROOT, "oracle", "_ref", "ref_counts_mid")

BASE_PARAMS = dict(
    data_file_type="ascii", random_file_type="ascii", coord_system="cartesian",
    ra_x_dataset_name="x", dec_y_dataset_name="y", z_z_dataset_name="z",
    weight_dataset_name="w", jk_dataset_name="jk", use_weights=0, n_threads=1,
    n_jk_regions=0, omega_m=0.25, h=1.0, z_min=0.0, z_max=1.0,
    plot_monopole=0, monopole_filename="mono.txt", monopole_output_type="ascii",
    monopole_log_base=1.3, monopole_min=0.0, monopole_max=100.0, monopole_n_bins=30,
    plot_sigma_pi=0, sigma_pi_filename="sp.txt", sigma_pi_output_type="ascii",
    sigma_log_base=1.0, sigma_min=0.0, sigma_max=50.0, sigma_n_bins=50,
    pi_log_base=1.0, pi_min=0.0, pi_max=50.0, pi_n_bins=50,
    plot_s_mu=0, s_mu_filename="smu.txt", s_mu_output_type="ascii",
    s_log_base=1.3, s_min=0.0, s_max=100.0, s_n_bins=40, mu_n_bins=50,
    calculate_angular_dd=0, calculate_angular_dd_invpweights=0,
    calculate_angular_dr=0, calculate_angular_dr_invpweights=0,
    angular_dd_filename="a.h5", angular_dd_invpweights_filename="b.h5",
    angular_dr_filename="c.h5", angular_dr_invpweights_filename="d.h5",
    theta_max=0.175, theta_n_bins=40, theta_log_base=1.6, healpix_order=5,
    n_bitwise_runs=2048, bitwise_weight_dataset_name="BW", dither_weight_dataset_name="DM",
)


def write_cat(path, cols, w, jk, use_w, n_jk):
    with open(path, "w") as f:
        f.write("# generated by make_golden.py\n")
        for i in range(len(cols[0])):
            parts = ["%.17g" % float(c[i]) for c in cols]
            if use_w:
                parts.append("%.9g" % float(w[i]))
            if n_jk > 1:
                parts.append("%d" % int(jk[i]))
            f.write(" ".join(parts) + "\n")


def run_case(name, params, data, rand, coord, dump_cat=False, store_inputs=True, norm_sub=("monopole",),
             midpoint=False):
    """data/rand: dict with three position columns (x,y,z or ra,dec,z), w, jk.
    store_inputs=False keeps only a SHA-256 of the inputs (for catalogues
    regenerated from a seed); norm_sub lists the correlators whose
    normalised jackknife blocks are kept."""
    cols = ("x", "y", "z") if coord == "cartesian" else ("ra", "dec", "z")
    p = dict(BASE_PARAMS)
    p.update(params)
    p["coord_system"] = coord
    out = {}
    with tempfile.TemporaryDirectory() as td:
        df, rf = os.path.join(td, "data.txt"), os.path.join(td, "rand.txt")
        p["data_filename"], p["random_filename"] = df, rf
        for path, cat in ((df, data), (rf, rand)):
            write_cat(path, [cat[c] for c in cols], cat["w"], cat["jk"], p["use_weights"], p["n_jk_regions"])
        pf = os.path.join(td, "params.txt")
        text = "".join(f"{k} = {v}\n" for k, v in p.items())
        with open(pf, "w") as f:
            f.write(text)
        js = os.path.join(td, "out.json")
        subprocess.run([REF_MID if midpoint else REF, pf, js, "1" if dump_cat else "0", "1"], check=True,
                       stdout=subprocess.DEVNULL, cwd=td)
        with open(js) as f:
            res = json.load(f)
    text = text.replace(df, "data.txt").replace(rf, "rand.txt")
    out["params"] = np.array(text)
    h = hashlib.sha256()
    for tag, cat in (("data", data), ("randoms", rand)):
        arrs = {f"in_{tag}_{c}": np.asarray(cat[c], dtype=np.float64) for c in cols}
        arrs[f"in_{tag}_w"] = np.asarray(cat["w"], dtype=np.float32)
        arrs[f"in_{tag}_jk"] = np.asarray(cat["jk"], dtype=np.int32)
        for k, v in arrs.items():
            h.update(np.ascontiguousarray(v).tobytes())
        if store_inputs:
            out.update(arrs)
        if dump_cat:
            for fld, v in res[f"{tag}_cat"].items():
                out[f"cat_{tag}_{fld}"] = np.array([float.fromhex(s) for s in v])
    out["in_sha256"] = np.array(h.hexdigest())
    out["pi_def"] = np.array(1 if midpoint else 0)
    n_jk = max(1, int(p["n_jk_regions"]))
    for corr in ("monopole", "sigma_pi", "s_mu"):
        if corr not in res:
            continue
        r = res[corr]
        axes = ["axis"] if corr == "monopole" else ["axis1", "axis2"]
        for k, an in enumerate(axes):
            a = r[an]
            out[f"{corr}_axis{k + 1}"] = np.array(
                [float.fromhex(a[f]) for f in ("min", "max", "step", "inv_step", "ilb", "ils")], dtype=np.float32)
            out[f"{corr}_axis{k + 1}_meta"] = np.array([a["n_bins"], a["log"]], dtype=np.int32)
        out[f"{corr}_n_used"] = np.array([r["n_data_used"], r["n_randoms_used"]])
        for h in ("DD", "DR", "RR"):
            out[f"{corr}_{h}_full"] = np.array(r[h]["full"], dtype=np.float64)
            out[f"{corr}_{h}_sub"] = np.array(r[h]["sub"], dtype=np.float64).reshape(n_jk, -1)
            out[f"{corr}_norm_{h}_full"] = np.array(r["norm"][h]["full"], dtype=np.float64)
            if corr in norm_sub:
                out[f"{corr}_norm_{h}_sub"] = np.array(r["norm"][h]["sub"], dtype=np.float64).reshape(n_jk, -1)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
    print(name, {k: v.shape for k, v in out.items() if k.endswith("_full")})


def cat_dict(cols, names, w, jk):
    d = {n: c for n, c in zip(names, cols)}
    d["w"], d["jk"] = w, jk
    return d


def case_cart_uniform():
    rng = np.random.Generator(np.random.PCG64(11))
    nd, nr = 1500, 2500
    d = rng.uniform(0, 160, (nd, 3))
    d[-20:] = d[:20]                      # exact duplicates: zero separation
    r = rng.uniform(0, 160, (nr, 3))
    wd = rng.uniform(0.5, 1.5, nd).astype(np.float32)
    wr = rng.uniform(0.5, 1.5, nr).astype(np.float32)
    jd = rng.integers(0, 4, nd).astype(np.int32)
    jr = rng.integers(0, 4, nr).astype(np.int32)
    run_case("cart_uniform",
             dict(use_weights=1, n_jk_regions=4, plot_monopole=1, plot_sigma_pi=1, plot_s_mu=1),
             cat_dict(d.T, "xyz", wd, jd), cat_dict(r.T, "xyz", wr, jr), "cartesian")


def case_eq_clustered():
    ra_l, dec_l, z_l = (180.0, 195.0), (5.0, 20.0), (0.10, 0.16)
    ra, dec, z = synth.neyman_scott_equatorial(70, 25.0, 5.0, ra_l, dec_l, z_l, seed=12)
    rra, rdec, rz = synth.randoms_equatorial(3000, ra_l, dec_l, z, seed=13)
    rng = np.random.Generator(np.random.PCG64(14))
    wd = rng.uniform(0.5, 1.5, ra.size).astype(np.float32)
    wr = rng.uniform(0.5, 1.5, rra.size).astype(np.float32)
    jd = synth.jk_grid_labels(ra, dec, ra_l, dec_l, 3, 3)
    jr = synth.jk_grid_labels(rra, rdec, ra_l, dec_l, 3, 3)
    run_case("eq_clustered",
             dict(use_weights=1, n_jk_regions=9, plot_monopole=1, plot_sigma_pi=1, plot_s_mu=1),
             cat_dict((ra, dec, z), ("ra", "dec", "z"), wd, jd),
             cat_dict((rra, rdec, rz), ("ra", "dec", "z"), wr, jr), "equatorial", dump_cat=True,
             norm_sub=("monopole", "s_mu"))


def _edge_axis_values():
    """Separations placed exactly on the s-axis edges (40 bins base 1.3 over
    [0,100]) computed as hist.cpp:55-60 does, plus their float neighbours."""
    ilb = np.float32(1.0 / np.log(np.float32(1.3)))
    ils = np.float32((1.3 ** 40 - 1.0) / 100.0)
    k = np.arange(1, 41)
    e = (np.exp(k / np.float64(ilb)) - 1.0) / np.float64(ils)
    vals = []
    for v in e:
        f = np.float32(v)
        vals += [float(v), float(f), float(np.nextafter(f, np.float32(0))), float(np.nextafter(f, np.float32(1e9)))]
    return np.array(vals + [100.0, 99.99999, 50.0, 49.99999])


def case_edge():
    rng = np.random.Generator(np.random.PCG64(15))
    pts = []
    # Pairs separated along x by exact edge values (each pair far from the others).
    for i, s in enumerate(_edge_axis_values()):
        base = np.array([1000.0 + 400.0 * (i % 12), 1000.0 + 400.0 * (i // 12), 1200.0])
        pts += [base, base + np.array([s, 0.0, 0.0])]
    # Radial pairs along rays through the origin (mu near or at 1).
    for i in range(40):
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        a = 300.0 + 37.0 * i
        pts += [a * u + 6000.0, (a + rng.uniform(0.1, 90.0)) * u + 6000.0]
    # Integer lattice: many exactly equal separations (sqrt of integers).
    g = np.stack(np.meshgrid(np.arange(6), np.arange(6), np.arange(6), indexing="ij"), -1).reshape(-1, 3) * 7.0
    pts += list(g + 9000.0)
    # Duplicates.
    pts += [pts[0], pts[5], pts[5]]
    d = np.array(pts)
    r = rng.uniform(0, 1, (600, 3)) * 400.0 + 9000.0
    nd = len(d)
    run_case("edge",
             dict(use_weights=0, n_jk_regions=0, plot_monopole=1, plot_sigma_pi=1, plot_s_mu=1,
                  sigma_log_base=1.5, sigma_min=0.5, sigma_max=50.0, sigma_n_bins=20,
                  pi_log_base=1.4, pi_min=1.0, pi_max=40.0, pi_n_bins=15),
             cat_dict(d.T, "xyz", np.ones(nd, np.float32), np.zeros(nd, np.int32)),
             cat_dict(r.T, "xyz", np.ones(len(r), np.float32), np.zeros(len(r), np.int32)),
             "cartesian")


def case_uniform_monopole():
    c = synth.uniform_monopole()
    d, r = c["data"], c["randoms"]
    run_case("uniform_monopole_50k",
             dict(use_weights=0, n_jk_regions=0, plot_monopole=1, n_threads=8),
             cat_dict(d.T, "xyz", np.ones(len(d), np.float32), np.zeros(len(d), np.int32)),
             cat_dict(r.T, "xyz", np.ones(len(r), np.float32), np.zeros(len(r), np.int32)),
             "cartesian", store_inputs=False)


def case_cart_midpoint():
    """Pair-midpoint line of sight (the reference built without
    _ANGULAR_PI_DEFN), cartesian points around the origin: pairs on opposite
    sides have a tiny or zero x1 + x2."""
    rng = np.random.Generator(np.random.PCG64(16))
    nd, nr = 1200, 2000
    d = rng.uniform(-70, 70, (nd, 3))
    d[:10] = -d[10:20]                    # x1 + x2 == 0 exactly
    r = rng.uniform(-70, 70, (nr, 3))
    wd = rng.uniform(0.5, 1.5, nd).astype(np.float32)
    wr = rng.uniform(0.5, 1.5, nr).astype(np.float32)
    jd = rng.integers(0, 3, nd).astype(np.int32)
    jr = rng.integers(0, 3, nr).astype(np.int32)
    run_case("cart_midpoint",
             dict(use_weights=1, n_jk_regions=3, plot_sigma_pi=1, plot_s_mu=1, sigma_log_base=1.25,
                  sigma_min=1.0, sigma_max=40.0, sigma_n_bins=20, pi_max=40.0, pi_n_bins=20),
             cat_dict(d.T, "xyz", wd, jd), cat_dict(r.T, "xyz", wr, jr), "cartesian", midpoint=True)


def case_eq_midpoint():
    ra_l, dec_l, z_l = (180.0, 195.0), (5.0, 20.0), (0.10, 0.16)
    ra, dec, z = synth.neyman_scott_equatorial(70, 25.0, 5.0, ra_l, dec_l, z_l, seed=12)
    rra, rdec, rz = synth.randoms_equatorial(3000, ra_l, dec_l, z, seed=13)
    rng = np.random.Generator(np.random.PCG64(14))
    wd = rng.uniform(0.5, 1.5, ra.size).astype(np.float32)
    wr = rng.uniform(0.5, 1.5, rra.size).astype(np.float32)
    jd = synth.jk_grid_labels(ra, dec, ra_l, dec_l, 3, 3)
    jr = synth.jk_grid_labels(rra, rdec, ra_l, dec_l, 3, 3)
    run_case("eq_midpoint",
             dict(use_weights=1, n_jk_regions=9, plot_sigma_pi=1, plot_s_mu=1),
             cat_dict((ra, dec, z), ("ra", "dec", "z"), wd, jd),
             cat_dict((rra, rdec, rz), ("ra", "dec", "z"), wr, jr), "equatorial", midpoint=True)


CASES = dict(cart_uniform=case_cart_uniform, eq_clustered=case_eq_clustered, edge=case_edge,
             uniform_monopole=case_uniform_monopole, cart_midpoint=case_cart_midpoint,
             eq_midpoint=case_eq_midpoint)

if __name__ == "__main__":
    # python make_golden.py [case ...]   (all cases by default)
    for ref in (REF, REF_MID):
        if not os.path.exists(ref):
            sys.exit(f"{ref} missing: run make -C oracle ref")
    for name in sys.argv[1:] or list(CASES):
        CASES[name]()
