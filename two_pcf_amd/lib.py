"""ctypes binding of libtwopcf_hip.so (include/twopcf_hip.h).

Loading fails loudly if the library is not built: there is no fallback path.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# the in-tree build, or another build of the same ABI named by $TWOPCF_LIB
LIB_PATH = os.environ.get("TWOPCF_LIB") or os.path.join(HERE, "libtwopcf_hip.so")

MONOPOLE, SIGMA_PI, S_MU, ANGULAR = 0, 1, 2, 3
PLAIN_FP64, FP32_RECHECK = 0, 1
W_PLAIN, W_INVP_3D, W_ANGUP_3D, W_ANG_IW, W_ANG_INVP, W_ANG_W = range(6)
PI_ANGULAR, PI_MIDPOINT = 0, 1
ABI_VERSION = 5   # TWOPCF_ABI_VERSION of the structures below

EXPORTS = [
    "twopcf_abi_version", "twopcf_ctx_create", "twopcf_ctx_destroy", "twopcf_last_error",
    "twopcf_device_info", "twopcf_catalog_upload", "twopcf_catalog_free", "twopcf_catalog_size",
    "twopcf_count", "twopcf_s2_thresholds", "twopcf_axis_thresholds", "twopcf_comoving_distance",
    "twopcf_eq_to_cart", "twopcf_autotune", "twopcf_count_noalloc",
]


class CatDesc(C.Structure):
    _fields_ = [("n", C.c_int64),
                ("x", C.c_void_p), ("y", C.c_void_p), ("z", C.c_void_p),
                ("nx", C.c_void_p), ("ny", C.c_void_p), ("nz", C.c_void_p),
                ("w", C.c_void_p), ("jk", C.c_void_p), ("iw", C.c_void_p),
                ("masks", C.c_void_p), ("n_words", C.c_int32), ("dither", C.c_void_p)]


class AxisDesc(C.Structure):
    _fields_ = [("min", C.c_double), ("max", C.c_double), ("n_bins", C.c_int32), ("log_base", C.c_double)]


class JobDesc(C.Structure):
    _fields_ = [("kind", C.c_int32), ("a", C.c_void_p), ("b", C.c_void_p), ("ax", AxisDesc * 2),
                ("n_jk", C.c_int32), ("weight_mode", C.c_int32), ("n_runs", C.c_int32),
                ("up_bins", C.c_void_p), ("up_axis", AxisDesc), ("pi_def", C.c_int32),
                ("dither", C.c_int32), ("full", C.c_void_p), ("jk", C.c_void_p), ("n_pairs", C.c_void_p),
                ("n_rechecked", C.c_void_p)]


class Timing(C.Structure):
    _fields_ = [("kernel_ms", C.c_double), ("prep_ms", C.c_double), ("total_ms", C.c_double),
                ("n_launches", C.c_int32), ("n_work_items", C.c_int32),
                ("n_pairs", C.c_uint64), ("n_rechecked", C.c_uint64), ("n_staged", C.c_uint64)]


class Tuning(C.Structure):
    _fields_ = [("cell_subdiv", C.c_int32), ("max_slots", C.c_int32), ("i_chunk", C.c_int32)]


class TwopcfError(RuntimeError):
    pass


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built (run `python -m two_pcf_amd.build`); "
                               "there is no CPU fallback")
        L = C.CDLL(LIB_PATH)
        if L.twopcf_abi_version() != ABI_VERSION:
            raise RuntimeError(f"{LIB_PATH} has ABI version {L.twopcf_abi_version()}, this binding {ABI_VERSION}; "
                               "rebuild it (python -m two_pcf_amd.build)")
        L.twopcf_ctx_create.argtypes = [C.c_int, C.POINTER(C.c_void_p)]
        L.twopcf_ctx_destroy.argtypes = [C.c_void_p]
        L.twopcf_last_error.argtypes = [C.c_void_p]
        L.twopcf_last_error.restype = C.c_char_p
        L.twopcf_device_info.argtypes = [C.c_void_p, C.c_char_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.twopcf_catalog_upload.argtypes = [C.c_void_p, C.POINTER(CatDesc), C.POINTER(C.c_void_p)]
        L.twopcf_catalog_free.argtypes = [C.c_void_p]
        L.twopcf_catalog_size.argtypes = [C.c_void_p]
        L.twopcf_catalog_size.restype = C.c_int64
        L.twopcf_count.argtypes = [C.c_void_p, C.POINTER(JobDesc), C.c_int, C.c_int, C.POINTER(Tuning),
                                   C.POINTER(Timing)]
        L.twopcf_count_noalloc.argtypes = L.twopcf_count.argtypes
        L.twopcf_s2_thresholds.argtypes = [C.c_int, C.POINTER(AxisDesc), C.POINTER(AxisDesc), C.c_void_p]
        L.twopcf_axis_thresholds.argtypes = [C.POINTER(AxisDesc), C.c_void_p]
        L.twopcf_comoving_distance.argtypes = [C.c_double, C.c_double, C.c_double, C.c_double, C.c_int64,
                                               C.c_void_p, C.c_void_p]
        L.twopcf_eq_to_cart.argtypes = [C.c_int64] + [C.c_void_p] * 9
        L.twopcf_autotune.argtypes = [C.c_void_p, C.POINTER(JobDesc), C.c_int, C.POINTER(Tuning),
                                      C.POINTER(C.c_double)]
        _lib = L
    return _lib


def axis_desc(mn, mx, n, base):
    return AxisDesc(float(mn), float(mx), int(n), float(base))


# ----------------------------------------------------------- host helpers --

def s2_thresholds(kind, ax1, ax2=None):
    n = ax1[2]
    t = np.zeros(n + 1 if kind != SIGMA_PI else 2, np.float32)
    a1 = axis_desc(*ax1)
    a2 = axis_desc(*ax2) if ax2 is not None else a1
    rc = lib().twopcf_s2_thresholds(kind, C.byref(a1), C.byref(a2), t.ctypes.data)
    if rc:
        raise TwopcfError(f"twopcf_s2_thresholds: {rc}")
    return t


def axis_thresholds(ax):
    t = np.zeros(ax[2] + 1, np.float32)
    rc = lib().twopcf_axis_thresholds(C.byref(axis_desc(*ax)), t.ctypes.data)
    if rc:
        raise TwopcfError(f"twopcf_axis_thresholds: {rc}")
    return t


def comoving_distance(omega_m, h, z_min, z_max, z):
    z = np.ascontiguousarray(z, dtype=np.float64)
    d = np.empty_like(z)
    rc = lib().twopcf_comoving_distance(omega_m, h, z_min, z_max, z.size, z.ctypes.data, d.ctypes.data)
    if rc:
        raise TwopcfError(f"twopcf_comoving_distance: {rc}")
    return d


def eq_to_cart(ra, dec, d):
    ra, dec, d = (np.ascontiguousarray(v, dtype=np.float64) for v in (ra, dec, d))
    out = [np.empty_like(ra) for _ in range(6)]
    rc = lib().twopcf_eq_to_cart(ra.size, ra.ctypes.data, dec.ctypes.data, d.ctypes.data,
                                 *[o.ctypes.data for o in out])
    if rc:
        raise TwopcfError(f"twopcf_eq_to_cart: {rc}")
    return out  # x, y, z, nx, ny, nz


def cart_norm(x, y, z):
    """Cartesian reader geometry (io.cpp:393-399): d = |x|, n = x / d."""
    x, y, z = (np.ascontiguousarray(v, dtype=np.float64) for v in (x, y, z))
    d = np.sqrt((x * x + y * y) + z * z)
    with np.errstate(invalid="ignore", divide="ignore"):
        return d, x / d, y / d, z / d


# ---------------------------------------------------------------- device --

class Catalog:
    def __init__(self, ctx, handle, n, keep):
        self.ctx, self.handle, self.n, self._keep = ctx, handle, n, keep

    def free(self):
        if self.handle:
            lib().twopcf_catalog_free(self.handle)
            self.handle = None

    def __len__(self):
        return self.n


class Context:
    def __init__(self, device=0):
        h = C.c_void_p()
        rc = lib().twopcf_ctx_create(device, C.byref(h))
        if rc:
            raise TwopcfError(f"twopcf_ctx_create({device}) failed: {rc}")
        self.h = h
        self.cats = []

    def err(self):
        return lib().twopcf_last_error(self.h).decode()

    def check(self, rc, what):
        if rc:
            raise TwopcfError(f"{what}: {rc} ({self.err()})")

    def info(self):
        name = C.create_string_buffer(64)
        ncu, lds = C.c_int(), C.c_int()
        self.check(lib().twopcf_device_info(self.h, name, 64, C.byref(ncu), C.byref(lds)), "device_info")
        return dict(arch=name.value.decode(), n_cu=ncu.value, lds_per_block=lds.value)

    def upload(self, x, y, z, nx, ny, nz, w=None, jk=None, iw=None, masks=None, dither=None):
        f8 = lambda a: np.ascontiguousarray(a, dtype=np.float64)
        keep = [f8(x), f8(y), f8(z), f8(nx), f8(ny), f8(nz)]
        n = keep[0].size
        wa = None if w is None else np.ascontiguousarray(w, dtype=np.float32)
        ja = None if jk is None else np.ascontiguousarray(jk, dtype=np.int32)
        ia = None if iw is None else np.ascontiguousarray(iw, dtype=np.float32)
        ma = None if masks is None else np.ascontiguousarray(masks, dtype=np.uint64)
        da = None if dither is None else np.ascontiguousarray(dither, dtype=np.uint64)
        if da is not None and (ma is None or da.shape != ma.shape):
            raise ValueError("dither masks must have the shape of the bit masks")
        keep += [wa, ja, ia, ma, da]
        p = lambda a: a.ctypes.data if a is not None else None
        d = CatDesc(n, *[p(a) for a in keep[:6]], p(wa), p(ja), p(ia), p(ma),
                    0 if ma is None else ma.shape[1], p(da))
        h = C.c_void_p()
        self.check(lib().twopcf_catalog_upload(self.h, C.byref(d), C.byref(h)), "catalog_upload")
        cat = Catalog(self, h, n, None)
        self.cats.append(cat)
        return cat

    def count(self, jobs, method=FP32_RECHECK, tuning=None, noalloc=False):
        """jobs: list of dicts(kind, a, b, axes=[(min,max,n,base), ...], n_jk).
        Fills and returns, per job, dict(full, jk, n_pairs, n_rechecked);
        plus the timing record.  tuning None: the autotuned choice for this
        job set if autotune() ran, else the defaults.  noalloc: through
        twopcf_count_noalloc (no device allocation; the catalogues must be
        gridded for the job set by an earlier count)."""
        descs, outs = self._descs(jobs)
        t = Timing()
        tu = None
        if tuning is not None:
            tu = Tuning(int(tuning.get("cell_subdiv", 0)), int(tuning.get("max_slots", 0)),
                        int(tuning.get("i_chunk", 0)))
        fn = lib().twopcf_count_noalloc if noalloc else lib().twopcf_count
        self.check(fn(self.h, descs, len(jobs), method, C.byref(tu) if tu is not None else None, C.byref(t)),
                   "count_noalloc" if noalloc else "count")
        res = []
        for o in outs:
            o.pop("_up", None)
            o["n_pairs"] = int(o["n_pairs"][0])
            o["n_rechecked"] = int(o["n_rechecked"][0])
            res.append(o)
        timing = {f: getattr(t, f) for f, _ in Timing._fields_}
        return res, timing

    def autotune(self, jobs):
        """Choose the batched method's tuning for this job set (kept in the
        context for later count() calls without a tuning).  Returns the
        tuning dict and the estimated full-batch kernel time (ms)."""
        descs, outs = self._descs(jobs)
        best = Tuning()
        ms = C.c_double()
        self.check(lib().twopcf_autotune(self.h, descs, len(jobs), C.byref(best), C.byref(ms)), "autotune")
        return dict(cell_subdiv=best.cell_subdiv, max_slots=best.max_slots, i_chunk=best.i_chunk), ms.value

    def _descs(self, jobs):
        descs = (JobDesc * len(jobs))()
        outs = []
        for k, j in enumerate(jobs):
            kind = j["kind"]
            axes = list(j["axes"])
            if len(axes) == 1:
                axes.append((0.0, 1.0, 1, 0.0))
            nb = axes[0][2] * (axes[1][2] if kind in (SIGMA_PI, S_MU) else 1)
            n_jk = int(j.get("n_jk", 1))
            o = dict(full=np.zeros(nb), jk=np.zeros((n_jk, nb)), n_pairs=np.zeros(1, np.uint64),
                     n_rechecked=np.zeros(1, np.uint64))
            outs.append(o)
            D = descs[k]
            D.kind = kind
            D.a = j["a"].handle
            D.b = j["b"].handle if j.get("b") is not None else None
            D.ax[0] = axis_desc(*axes[0])
            D.ax[1] = axis_desc(*axes[1])
            D.n_jk = n_jk
            D.weight_mode = int(j.get("weight_mode", W_PLAIN))
            D.n_runs = int(j.get("n_runs", 0))
            D.pi_def = int(j.get("pi_def", PI_ANGULAR))
            D.dither = int(bool(j.get("dither", False)))
            if j.get("up_bins") is not None:
                ub = np.ascontiguousarray(j["up_bins"], dtype=np.float64)
                o["_up"] = ub
                D.up_bins = ub.ctypes.data
                D.up_axis = axis_desc(*j["up_axis"])
            D.full = o["full"].ctypes.data
            D.jk = o["jk"].ctypes.data
            D.n_pairs = o["n_pairs"].ctypes.data
            D.n_rechecked = o["n_rechecked"].ctypes.data
        return descs, outs

    def close(self):
        if self.h:
            for c in self.cats:
                c.handle = None
            lib().twopcf_ctx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def finish(is_auto, full, jk):
    """invert_jk_hists + multiply_hists(2.0) for auto counts
    (hist.cpp:464-482; correlators.cpp:391-392, 429)."""
    sub = full[None, :] - jk
    if is_auto:
        return full * 2.0, sub * 2.0
    return full.copy(), sub
