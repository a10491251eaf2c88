"""The inverse-p golden cases (tests/golden/make_invp_golden.py: the reference
built with _USE_INV_WEIGHTS, and _USE_DITHER_WEIGHTS for dither_w3) as
catalogues and axes for the oracle and for the GPU library.  Shared by the CPU
and GPU test modules.

The catalogue inputs come from the fixture (small cases) or are regenerated
from their seeds and checked against the fixture's SHA-256."""
import importlib.util
import os

import numpy as np

from conftest import GOLDEN, load_golden, parse_params

CASES = ("invp_w2", "invp_w32_axis", "dither_w3")
CORRS = ("monopole", "sigma_pi", "s_mu")
ANGULAR = ("DD", "DDinvp", "DR", "DRinvp")

_gen = None


def generator():
    """tests/golden/make_invp_golden.py as a module (its seeded catalogues)."""
    global _gen
    if _gen is None:
        spec = importlib.util.spec_from_file_location("make_invp_golden",
                                                      os.path.join(GOLDEN, "make_invp_golden.py"))
        _gen = importlib.util.module_from_spec(spec)
        spec.loader.exec_module(_gen)
    return _gen


def popcount_rows(m):
    """Set bits per row of a uint64 array, by Python's int.bit_count."""
    return np.array([sum(int(v).bit_count() for v in row) for row in m.astype(np.uint64)], dtype=np.int64)


class InvpCase:
    def __init__(self, name):
        self.name = name
        self.g = load_golden(name)
        self.p = parse_params(self.g["params"])
        self.n_runs = int(self.p["n_bitwise_runs"])
        self.words = (self.n_runs + 63) // 64
        self.n_jk = max(1, int(self.p["n_jk_regions"]))
        if "in_data_RA" in self.g:
            self.inputs = {tag: {k[len(f"in_{tag}_"):]: v for k, v in self.g.items() if k.startswith(f"in_{tag}_")}
                           for tag in ("data", "rand")}
        else:
            self.inputs = generator().CATS[name]()
            assert generator().sha256_of(self.inputs) == str(self.g["in_sha256"]), \
                "regenerated catalogues differ from the fixture's"
        self.dither = "DITHER0" in self.inputs["data"]

    def fp(self, k):
        return float(self.p[k])

    def columns(self, tag):
        """(ra, dec, z, w as float, jk, masks[n, words], dither masks or None)
        as the reader takes them: the weight dataset is double, stored in a
        float (io.cpp:489-499); regions only when n_jk_regions > 1."""
        c = self.inputs[tag]
        m = np.stack([c[f"BITWEIGHT{k}"].view(np.uint64) for k in range(self.words)], axis=1)
        dm = None
        if self.dither:
            dm = np.stack([c[f"DITHER{k}"].view(np.uint64) for k in range(self.words)], axis=1)
        jk = c["JK"].astype(np.int32) if int(self.p["n_jk_regions"]) > 1 else np.zeros(c["RA"].size, np.int32)
        return c["RA"], c["DEC"], c["Z"], c["W"].astype(np.float32), jk, m, dm

    def iip_weights(self, tag):
        """(input weight, weight after the IIP rule) computed here, not by the
        code under test: total = popcount of the own masks (0 counts as 1);
        where total != n_runs the weight is multiplied by n_runs / total, or
        with dither masks by dither_total / total (io.cpp:521-588)."""
        _, _, _, w, _, m, dm = self.columns(tag)
        tot = popcount_rows(m)
        tot[tot == 0] = 1
        num = popcount_rows(dm).astype(np.float64) if self.dither else float(self.n_runs)
        w_iip = np.where(tot != self.n_runs, (w.astype(np.float64) * (num / tot)).astype(np.float32), w)
        return w, w_iip

    def angular_axis(self):
        return (0.0, float(1.0 - np.cos(self.fp("theta_max"))), int(self.p["theta_n_bins"]),
                self.fp("theta_log_base"))

    def axes(self, corr):
        p = self.p
        if corr == "monopole":
            return [(self.fp("monopole_min"), self.fp("monopole_max"), int(p["monopole_n_bins"]),
                     self.fp("monopole_log_base"))]
        if corr == "sigma_pi":
            return [(self.fp("sigma_min"), self.fp("sigma_max"), int(p["sigma_n_bins"]), self.fp("sigma_log_base")),
                    (self.fp("pi_min"), self.fp("pi_max"), int(p["pi_n_bins"]), self.fp("pi_log_base"))]
        return [(self.fp("s_min"), self.fp("s_max"), int(p["s_n_bins"]), self.fp("s_log_base")),
                (0.0, 1.0, int(p["mu_n_bins"]), 0.0)]

    def corrs(self):
        return [c for c in CORRS if f"{c}_DD_full" in self.g]

    def oracle_cats(self, o):
        """Oracle catalogues: positions by the oracle's cosmology table and
        Galaxy::set, weights from iip_weights."""
        table = o.cosmo_table(self.fp("omega_m"), self.fp("h"), self.fp("z_min"), self.fp("z_max"))
        out = {}
        for tag in ("data", "rand"):
            ra, dec, z, _, jk, m, dm = self.columns(tag)
            d = o.comov_dist(table, self.fp("z_min"), z)
            x, y, zz, nx, ny, nz = o.eq_to_cart(ra, dec, d)
            iw, w = self.iip_weights(tag)
            out[tag] = o.Catalog(x, y, zz, nx, ny, nz, w, jk, iw=iw, masks=m, dither=dm)
        return out["data"], out["rand"]
