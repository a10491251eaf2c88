"""ctypes front-end of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker or the CPU baseline.  The product
package (two_pcf_amd) never imports it.

The C library restates the reference's pair arithmetic (correlators.cpp,
hist.cpp, cosmology.cpp, galaxy.cpp); see the header of twopcf_oracle.c for
what pins each function.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")

MONO, SIGMA_PI, S_MU, ANGULAR = 0, 1, 2, 3
W_PLAIN, W_INVP_3D, W_ANGUP_3D, W_ANG_IW, W_ANG_INVP, W_ANG_W = range(6)


class Axis(C.Structure):
    _fields_ = [("min", C.c_float), ("minus_min", C.c_float), ("max", C.c_float),
                ("step", C.c_float), ("inv_step", C.c_float), ("base", C.c_float),
                ("ilb", C.c_float), ("ils", C.c_float), ("n_bins", C.c_int),
                ("log_binning", C.c_int)]


class Cat(C.Structure):
    _fields_ = [("n", C.c_long),
                ("x", C.c_void_p), ("y", C.c_void_p), ("z", C.c_void_p),
                ("nx", C.c_void_p), ("ny", C.c_void_p), ("nz", C.c_void_p),
                ("w", C.c_void_p), ("jk", C.c_void_p), ("iw", C.c_void_p),
                ("masks", C.c_void_p), ("n_words", C.c_int), ("dither", C.c_void_p)]


class WOpt(C.Structure):
    _fields_ = [("mode", C.c_int), ("n_runs", C.c_int), ("up_bins", C.c_void_p),
                ("up_axis", Axis), ("pi_def", C.c_int), ("dither", C.c_int)]


PI_ANGULAR, PI_MIDPOINT = 0, 1


def build():
    """Compile liboracle.so (gcc) — and, when the reference sources are
    present in this container, oracle/_ref from them."""
    subprocess.run(["make", "-s", "-C", HERE, "oracle"], check=True)
    if os.path.isdir("/root/reference/src"):
        subprocess.run(["make", "-s", "-j8", "-C", HERE, "ref"], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        L.or_axis_init.argtypes = [C.POINTER(Axis), C.c_float, C.c_float, C.c_int, C.c_float]
        L.or_axis_bin_num.argtypes = [C.POINTER(Axis), C.c_float]
        L.or_axis_bin_num.restype = C.c_float
        for f in ("or_axis_value_from_bin",):
            getattr(L, f).argtypes = [C.POINTER(Axis), C.c_double]
            getattr(L, f).restype = C.c_double
        for f in ("or_axis_bin_centre", "or_axis_bin_width"):
            getattr(L, f).argtypes = [C.POINTER(Axis), C.c_int]
            getattr(L, f).restype = C.c_double
        L.or_interp.argtypes = [C.POINTER(Axis), C.c_void_p, C.c_double]
        L.or_interp.restype = C.c_double
        L.or_count.argtypes = [C.c_int, C.POINTER(Cat), C.POINTER(Cat), C.POINTER(Axis),
                               C.POINTER(Axis), C.c_int, C.POINTER(WOpt), C.c_int, C.c_int,
                               C.c_void_p, C.c_void_p, C.POINTER(C.c_ulonglong)]
        L.or_count_sample.argtypes = [C.c_int, C.POINTER(Cat), C.POINTER(Cat), C.POINTER(Axis),
                                      C.POINTER(Axis), C.c_int, C.c_int, C.c_double,
                                      C.c_ulonglong, C.c_void_p, C.c_void_p,
                                      C.POINTER(C.c_ulonglong), C.POINTER(C.c_double)]
        L.or_range_sq.argtypes = [C.c_int, C.POINTER(Axis), C.POINTER(Axis),
                                  C.POINTER(C.c_float), C.POINTER(C.c_float)]
        L.or_finish.argtypes = [C.c_int, C.c_long, C.c_int, C.c_void_p, C.c_void_p]
        L.or_normalise_auto.argtypes = [C.c_long, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                        C.c_void_p, C.c_long]
        L.or_normalise_cross.argtypes = [C.c_long, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                         C.c_void_p, C.c_long, C.c_void_p, C.c_void_p, C.c_long]
        L.or_ls.argtypes = [C.c_double, C.c_double, C.c_double]
        L.or_ls.restype = C.c_double
        L.or_cosmo_len.argtypes = [C.c_double, C.c_double, C.c_double]
        L.or_cosmo_len.restype = C.c_int
        L.or_cosmo_table.argtypes = [C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                     C.c_int, C.c_void_p]
        L.or_comov_dist.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_double]
        L.or_comov_dist.restype = C.c_double
        L.or_eq_to_cart.argtypes = [C.c_long] + [C.c_void_p] * 9
        L.or_cart_norm.argtypes = [C.c_long] + [C.c_void_p] * 7
        L.or_scan_s2_table.argtypes = [C.c_int, C.POINTER(Axis), C.POINTER(Axis), C.c_void_p,
                                       C.c_float, C.c_float, C.c_int]
        L.or_scan_s2_table.restype = C.c_long
        L.or_scan_value_table.argtypes = [C.POINTER(Axis), C.c_void_p, C.c_float, C.c_float, C.c_int]
        L.or_scan_value_table.restype = C.c_long
        _lib = L
    return _lib


def axis(mn, mx, n, base):
    a = Axis()
    lib().or_axis_init(C.byref(a), mn, mx, n, base)
    return a


def _p(a):
    return a.ctypes.data if a is not None else None


class Catalog:
    """SoA catalogue as the oracle wants it (keeps the numpy arrays alive)."""

    def __init__(self, x, y, z, nx, ny, nz, w=None, jk=None, iw=None, masks=None, dither=None):
        n = len(x)
        f8 = lambda a: np.ascontiguousarray(a, dtype=np.float64)
        self.x, self.y, self.z = f8(x), f8(y), f8(z)
        self.nx, self.ny, self.nz = f8(nx), f8(ny), f8(nz)
        self.w = np.ascontiguousarray(np.ones(n) if w is None else w, dtype=np.float32)
        self.jk = np.ascontiguousarray(np.zeros(n) if jk is None else jk, dtype=np.int32)
        self.iw = None if iw is None else np.ascontiguousarray(iw, dtype=np.float32)
        self.masks = None if masks is None else np.ascontiguousarray(masks, dtype=np.uint64)
        self.n_words = 0 if masks is None else self.masks.shape[1]
        self.dither = None if dither is None else np.ascontiguousarray(dither, dtype=np.uint64)
        if self.dither is not None and self.dither.shape != self.masks.shape:
            raise ValueError("dither masks must have the shape of the bit masks")
        self.c = Cat(n, _p(self.x), _p(self.y), _p(self.z), _p(self.nx), _p(self.ny), _p(self.nz),
                     _p(self.w), _p(self.jk), _p(self.iw), _p(self.masks), self.n_words, _p(self.dither))

    def __len__(self):
        return int(self.c.n)


def count(kind, a, b, ax1, ax2=None, n_jk=1, method=1, nthreads=0, wopt=None):
    """Raw sums (full[nb], sub[n_jk, nb]) before invert/x2, plus the number of
    accepted pairs."""
    ax2 = ax2 if ax2 is not None else ax1
    nb = ax1.n_bins * (ax2.n_bins if kind in (SIGMA_PI, S_MU) else 1)
    full = np.zeros(nb)
    sub = np.zeros((n_jk, nb))
    acc = C.c_ulonglong(0)
    rc = lib().or_count(kind, C.byref(a.c), C.byref(b.c) if b is not None else None,
                        C.byref(ax1), C.byref(ax2), n_jk,
                        C.byref(wopt) if wopt is not None else None, method, nthreads,
                        full.ctypes.data, sub.ctypes.data, C.byref(acc))
    if rc != 0:
        raise ValueError(f"or_count failed: {rc}")
    return full, sub, int(acc.value)


def count_sample(kind, a, b, ax1, ax2, n_jk, frac, seed=1, nthreads=0):
    nb = ax1.n_bins * (ax2.n_bins if kind in (SIGMA_PI, S_MU) else 1)
    full = np.zeros(nb)
    sub = np.zeros((n_jk, nb))
    acc = C.c_ulonglong(0)
    fr = C.c_double(0)
    rc = lib().or_count_sample(kind, C.byref(a.c), C.byref(b.c) if b is not None else None,
                               C.byref(ax1), C.byref(ax2), n_jk, nthreads, frac, seed,
                               full.ctypes.data, sub.ctypes.data, C.byref(acc), C.byref(fr))
    if rc != 0:
        raise ValueError(f"or_count_sample failed: {rc}")
    return full, sub, int(acc.value), fr.value


def finish(is_auto, full, sub):
    full = full.copy()
    sub = sub.copy()
    lib().or_finish(int(is_auto), full.size, sub.shape[0], full.ctypes.data, sub.ctypes.data)
    return full, sub


def normalise_auto(full, sub, cat):
    full = full.copy()
    sub = sub.copy()
    lib().or_normalise_auto(full.size, sub.shape[0], full.ctypes.data, sub.ctypes.data,
                            cat.w.ctypes.data, cat.jk.ctypes.data, len(cat))
    return full, sub


def normalise_cross(full, sub, c1, c2):
    full = full.copy()
    sub = sub.copy()
    lib().or_normalise_cross(full.size, sub.shape[0], full.ctypes.data, sub.ctypes.data,
                             c1.w.ctypes.data, c1.jk.ctypes.data, len(c1),
                             c2.w.ctypes.data, c2.jk.ctypes.data, len(c2))
    return full, sub


def range_sq(kind, ax1, ax2=None):
    mn, mx = C.c_float(0), C.c_float(0)
    lib().or_range_sq(kind, C.byref(ax1), C.byref(ax2 if ax2 is not None else ax1),
                      C.byref(mn), C.byref(mx))
    return mn.value, mx.value


def cosmo_table(om, h, z_min, z_max, step=1e-4, overs=100):
    n = lib().or_cosmo_len(z_min, z_max, step)
    y = np.zeros(n)
    lib().or_cosmo_table(om, h, z_min, z_max, step, overs, y.ctypes.data)
    return y


def comov_dist(table, z_min, z, step=1e-4):
    L = lib()
    return np.array([L.or_comov_dist(table.ctypes.data, table.size, z_min, step, float(v))
                     for v in np.atleast_1d(z)])


def eq_to_cart(ra, dec, d):
    n = len(ra)
    ra, dec, d = (np.ascontiguousarray(v, dtype=np.float64) for v in (ra, dec, d))
    out = [np.zeros(n) for _ in range(6)]
    lib().or_eq_to_cart(n, ra.ctypes.data, dec.ctypes.data, d.ctypes.data,
                        *[o.ctypes.data for o in out])
    return out


def cart_norm(x, y, z):
    n = len(x)
    x, y, z = (np.ascontiguousarray(v, dtype=np.float64) for v in (x, y, z))
    out = [np.zeros(n) for _ in range(4)]
    lib().or_cart_norm(n, x.ctypes.data, y.ctypes.data, z.ctypes.data, *[o.ctypes.data for o in out])
    return out  # d, nx, ny, nz
