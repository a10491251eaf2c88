"""Helpers that turn a golden fixture into catalogues and axes for the oracle
and for the GPU library.  Shared by the CPU and GPU test modules."""
import hashlib

import numpy as np

from conftest import load_golden, parse_params

CORRS = ("monopole", "sigma_pi", "s_mu")


class Case:
    def __init__(self, name, regen=None):
        self.name = name
        self.g = load_golden(name)
        self.p = parse_params(self.g["params"])
        self.coord = self.p["coord_system"]
        self.n_jk = max(1, int(self.p["n_jk_regions"]))
        # line of sight: 0 angular (the reference's default build), 1 pair midpoint
        self.pi_def = int(self.g["pi_def"]) if "pi_def" in self.g else 0
        self.inputs = {}
        cols = ("x", "y", "z") if self.coord == "cartesian" else ("ra", "dec", "z")
        if "in_data_w" in self.g:
            for tag in ("data", "randoms"):
                self.inputs[tag] = {c: self.g[f"in_{tag}_{c}"] for c in cols + ("w", "jk")}
        else:
            self.inputs = regen()
            h = hashlib.sha256()
            for tag in ("data", "randoms"):
                for c in cols:
                    h.update(np.ascontiguousarray(self.inputs[tag][c], dtype=np.float64).tobytes())
                h.update(np.ascontiguousarray(self.inputs[tag]["w"], dtype=np.float32).tobytes())
                h.update(np.ascontiguousarray(self.inputs[tag]["jk"], dtype=np.int32).tobytes())
            assert h.hexdigest() == str(self.g["in_sha256"]), "regenerated inputs differ from the fixture's"

    def fp(self, k):
        return float(self.p[k])

    def axes(self, corr):
        """(min, max, n, base) per axis, as the correlators build them
        (correlators.cpp:131, 284-285, 436-437)."""
        if corr == "monopole":
            return [(self.fp("monopole_min"), self.fp("monopole_max"), int(self.p["monopole_n_bins"]),
                     self.fp("monopole_log_base"))]
        if corr == "sigma_pi":
            return [(self.fp("sigma_min"), self.fp("sigma_max"), int(self.p["sigma_n_bins"]), self.fp("sigma_log_base")),
                    (self.fp("pi_min"), self.fp("pi_max"), int(self.p["pi_n_bins"]), self.fp("pi_log_base"))]
        return [(self.fp("s_min"), self.fp("s_max"), int(self.p["s_n_bins"]), self.fp("s_log_base")),
                (0.0, 1.0, int(self.p["mu_n_bins"]), 0.0)]

    def corrs(self):
        return [c for c in CORRS if f"{c}_DD_full" in self.g]

    def golden(self, corr, h, norm=False):
        pre = f"{corr}_norm_{h}" if norm else f"{corr}_{h}"
        sub = self.g.get(pre + "_sub")
        return self.g[pre + "_full"], sub


def uniform_monopole_inputs():
    from two_pcf_amd import synth
    c = synth.uniform_monopole()
    out = {}
    for tag in ("data", "randoms"):
        a = c[tag]
        out[tag] = dict(x=a[:, 0].copy(), y=a[:, 1].copy(), z=a[:, 2].copy(),
                        w=np.ones(len(a), np.float32), jk=np.zeros(len(a), np.int32))
    return out


def load_case(name):
    return Case(name, regen=uniform_monopole_inputs if name == "uniform_monopole_50k" else None)
