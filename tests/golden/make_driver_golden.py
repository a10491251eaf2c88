"""Golden output files of the reference's parameter-file driver.

Run in the build container (needs oracle/_ref/TWOPCF, the reference compiled
from its own sources by oracle/Makefile):

    python tests/golden/make_driver_golden.py

Each case directory under tests/golden/driver/ holds a parameter file, the
ascii catalogues it names (space, tab and comma separated variants) and, under
expected/, every output file the reference wrote (full result plus one file
per jackknife region, ascii, default stream precision).
"""
import os
import shutil
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)
from make_golden import BASE_PARAMS  # noqa: E402
from two_pcf_amd import synth  # noqa: E402

REF = os.path.join(ROOT, "oracle", "_ref", "TWOPCF")
OUT = os.path.join(HERE, "driver")


def write(path, cols, sep, w=None, jk=None):
    with open(path, "w") as f:
        f.write("# x y z [w] [jk]\n")
        for i in range(len(cols[0])):
            parts = ["%.17g" % float(c[i]) for c in cols]
            if w is not None:
                parts.append("%.9g" % float(w[i]))
            if jk is not None:
                parts.append("%d" % int(jk[i]))
            f.write(sep.join(parts) + "\n")


def run_case(name, params, data, rand):
    d = os.path.join(OUT, name)
    if os.path.isdir(d):
        shutil.rmtree(d)
    os.makedirs(os.path.join(d, "expected"))
    write(os.path.join(d, "data.txt"), *data)
    write(os.path.join(d, "rand.txt"), *rand)
    p = dict(BASE_PARAMS)
    p.update(params)
    p["data_filename"], p["random_filename"] = "data.txt", "rand.txt"
    with open(os.path.join(d, "params.txt"), "w") as f:
        f.write("# driver golden case %s\n" % name)
        f.write("".join(f"{k} = {v}\n" for k, v in p.items()))
    work = os.path.join(d, "expected")
    for fn in ("data.txt", "rand.txt", "params.txt"):
        shutil.copy(os.path.join(d, fn), work)
    subprocess.run([REF, "params.txt"], cwd=work, check=True, stdout=subprocess.DEVNULL)
    for fn in ("data.txt", "rand.txt", "params.txt"):
        os.remove(os.path.join(work, fn))
    print(name, sorted(os.listdir(work)))


def case_cart():
    rng = np.random.Generator(np.random.PCG64(21))
    d = rng.uniform(0, 120, (800, 3))
    r = rng.uniform(0, 120, (1500, 3))
    wd = rng.uniform(0.5, 1.5, 800).astype(np.float32)
    wr = rng.uniform(0.5, 1.5, 1500).astype(np.float32)
    jd = rng.integers(0, 3, 800)
    jr = rng.integers(0, 3, 1500)
    run_case("cart_w_jk3",
             dict(use_weights=1, n_jk_regions=3, plot_monopole=1, monopole_n_bins=12, monopole_max=60.0,
                  monopole_log_base=1.3, plot_sigma_pi=1, sigma_n_bins=8, pi_n_bins=8, sigma_max=40.0, pi_max=40.0,
                  plot_s_mu=1, s_n_bins=10, s_max=60.0, s_log_base=1.4, mu_n_bins=5),
             (d.T, "\t", wd, jd), (r.T, ",", wr, jr))


def case_eq():
    ra_l, dec_l, z_l = (170.0, 180.0), (10.0, 20.0), (0.08, 0.12)
    ra, dec, z = synth.neyman_scott_equatorial(30, 20.0, 4.0, ra_l, dec_l, z_l, seed=22)
    rra, rdec, rz = synth.randoms_equatorial(1200, ra_l, dec_l, z, seed=23)
    run_case("eq_nojk",
             dict(coord_system="equatorial", use_weights=0, n_jk_regions=0, plot_monopole=1, monopole_n_bins=8,
                  monopole_max=50.0, plot_s_mu=1, s_n_bins=8, s_max=50.0, mu_n_bins=4),
             ((ra, dec, z), " ", None, None), ((rra, rdec, rz), " ", None, None))


if __name__ == "__main__":
    if not os.path.exists(REF):
        sys.exit(f"{REF} missing: run make -C oracle ref")
    case_cart()
    case_eq()
