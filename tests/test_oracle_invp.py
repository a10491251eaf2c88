"""Pin the oracle's inverse-p pair weights and angular counts
(correlators_invpweight.cpp) without the reference binary: its inverse-p
build needs libhdf5, so these rules are checked against an independent
pure-Python evaluation — popcounts with int.bit_count, float32 roundings
with numpy, and JKHists::interp (hist.cpp:369-391) restated in Python.  The
binning itself (or_pair_bin) is pinned by the golden tests."""
import ctypes as C

import numpy as np
import pytest

F = np.float32


def rand_masks(rng, n, words, lo, hi):
    p = rng.uniform(lo, hi, n)
    bits = rng.random((n, words * 64)) < p[:, None]
    return np.packbits(bits.astype(np.uint8), axis=1, bitorder="little").view(np.uint64).reshape(n, words)


def make_cat(o, n, words, seed, ones=False, dither=False):
    rng = np.random.Generator(np.random.PCG64(seed))
    ra = rng.uniform(180, 186, n)
    dec = rng.uniform(10, 16, n)
    d = rng.uniform(300, 360, n)
    x, y, z, nx, ny, nz = o.eq_to_cart(ra, dec, d)
    if ones:
        masks = np.full((n, words), np.uint64(0xFFFFFFFFFFFFFFFF), dtype=np.uint64)
    else:
        p = rng.uniform(0.3, 0.95, n)
        bits = rng.random((n, words * 64)) < p[:, None]
        masks = np.packbits(bits.astype(np.uint8), axis=1, bitorder="little").view(np.uint64).reshape(n, words)
    w = rng.uniform(0.5, 1.5, n).astype(F)
    iw = rng.uniform(0.5, 1.5, n).astype(F)
    jk = rng.integers(0, 2, n).astype(np.int32)
    # sparse dither masks: many pairs share no dither bit (the 0 -> 1 rule)
    dm = rand_masks(rng, n, words, 0.0, 0.3) if dither else None
    return o.Catalog(x, y, z, nx, ny, nz, w, jk, iw=iw, masks=masks, dither=dm)


def py_interp(o, ax, bins, v):
    """hist.cpp:369-391 in Python (index from the oracle's Axis::bin_num)."""
    L = o.lib()
    if v < float(F(ax.min)):
        return bins[0]
    if v > float(F(ax.max)):
        return bins[-1]
    bn = float(F(float(L.or_axis_bin_num(C.byref(ax), float(F(v)))) - 0.5))
    n = ax.n_bins
    if bn < 0:
        return bins[0]
    if bn > n - 1:
        return bins[-1]
    k = int(bn)
    c = [L.or_axis_bin_centre(C.byref(ax), i) for i in (k, k + 1)]
    return bins[k] + (bins[k + 1] - bins[k]) * (v - c[0]) / (c[1] - c[0])


def py_popc(m1, m2):
    """float(popcount(m1 & m2)), 0 counting as 1 (correlators_invpweight.cpp:55-59)."""
    pc = F(sum(int(int(x) & int(y)).bit_count() for x, y in zip(m1, m2)))
    return F(pc + F(1.0)) if pc < 0.5 else pc


def py_weight(o, mode, a, i, b, j, n_runs, up_ax=None, up_bins=None, dither=False):
    n1 = (a.nx[i], a.ny[i], a.nz[i])
    n2 = (b.nx[j], b.ny[j], b.nz[j])
    omc = 1.0
    for k in range(3):
        omc -= n1[k] * n2[k]
    pc = py_popc(a.masks[i], b.masks[j])
    # numerator: n_runs, or the dither popcount (correlators_invpweight.cpp:62-70, 749-757)
    num = py_popc(a.dither[i], b.dither[j]) if dither else n_runs
    if mode == 1:    # (iw1*iw2) * float(interp * num / pc)
        w = F(py_interp(o, up_ax, up_bins, omc) * float(num) / float(pc))
        return F(F(a.iw[i] * b.iw[j]) * w)
    if mode == 2:    # float(float(w1*w2) * interp)
        return F(float(F(a.w[i] * b.w[j])) * py_interp(o, up_ax, up_bins, omc))
    if mode == 3:
        return F(a.iw[i] * b.iw[j])
    if mode == 4:    # (iw1*iw2) * ((float)num / pc)
        return F(F(a.iw[i] * b.iw[j]) * F(F(num) / pc))
    return F(a.w[i] * b.w[j])


def brute(o, kind, mode, a, b, ax1, ax2, n_runs, up_ax=None, up_bins=None, dither=False):
    L = o.lib()
    nb = ax1.n_bins * (ax2.n_bins if kind in (1, 2) else 1)
    full = np.zeros(nb)
    mn, mx = o.range_sq(kind, ax1, ax2) if kind != 3 else (0.0, 0.0)
    bb = a if b is None else b
    binv = C.c_int()
    for i in range(len(a)):
        for j in range(i + 1 if b is None else 0, len(bb)):
            x1 = (C.c_double * 3)(a.x[i], a.y[i], a.z[i])
            m1 = (C.c_double * 3)(a.nx[i], a.ny[i], a.nz[i])
            x2 = (C.c_double * 3)(bb.x[j], bb.y[j], bb.z[j])
            m2 = (C.c_double * 3)(bb.nx[j], bb.ny[j], bb.nz[j])
            if L.or_pair_bin(kind, x1, m1, x2, m2, C.c_float(mn), C.c_float(mx), C.byref(ax1),
                             C.byref(ax2 if ax2 is not None else ax1), C.byref(binv)):
                full[binv.value] += float(py_weight(o, mode, a, i, bb, j, n_runs, up_ax, up_bins, dither))
    return full


@pytest.fixture(scope="module")
def cats(oracle_lib):
    o = oracle_lib
    o.lib().or_pair_bin.argtypes = [C.c_int] + [C.c_void_p] * 4 + [C.c_float, C.c_float, C.POINTER(o.Axis),
                                                                   C.POINTER(o.Axis), C.POINTER(C.c_int)]
    return make_cat(o, 70, 2, 1), make_cat(o, 90, 2, 2, ones=True)


@pytest.mark.parametrize("mode", [3, 4, 5])
def test_angular_weights_brute_force(oracle_lib, cats, mode):
    o = oracle_lib
    d, r = cats
    ax = o.axis(0.0, float(1.0 - np.cos(0.02)), 12, 1.6)
    for b in (None, r):
        wo = o.WOpt(mode, 128, None, o.Axis())
        full, _, _ = o.count(3, d, b, ax, None, 2, method=0, wopt=wo)
        want = brute(o, 3, mode, d, b, ax, None, 128)
        np.testing.assert_allclose(full, want, rtol=1e-13, atol=0)


@pytest.mark.parametrize("mode", [1, 2])
def test_3d_weights_with_upweight_brute_force(oracle_lib, cats, mode):
    o = oracle_lib
    d, r = cats
    ax1, ax2 = o.axis(0.0, 40.0, 10, 1.3), o.axis(0.0, 1.0, 5, 0.0)
    up_ax = o.axis(0.0, float(1.0 - np.cos(0.05)), 8, 1.6)
    up_bins = np.linspace(1.5, 0.9, 8)
    wo = o.WOpt(mode, 128, up_bins.ctypes.data, up_ax)
    b = None if mode == 1 else r
    full, sub, acc = o.count(2, d, b, ax1, ax2, 2, method=0, wopt=wo)
    want = brute(o, 2, mode, d, b, ax1, ax2, 128, up_ax, list(up_bins))
    np.testing.assert_allclose(full, want, rtol=1e-13, atol=0)
    # cells and brute force agree
    full1, sub1, acc1 = o.count(2, d, b, ax1, ax2, 2, method=1, wopt=wo)
    np.testing.assert_allclose(full1, full, rtol=1e-13)
    assert acc1 == acc


@pytest.fixture(scope="module")
def dcats(oracle_lib, cats):
    o = oracle_lib
    return make_cat(o, 70, 2, 11, dither=True), make_cat(o, 90, 2, 12, ones=True, dither=True)


@pytest.mark.parametrize("mode", [1, 4])
def test_dither_weights_brute_force(oracle_lib, dcats, mode):
    """_USE_DITHER_WEIGHTS: the inverse-p numerator is the dither popcount."""
    o = oracle_lib
    d, r = dcats
    # the sample really exercises the 0 -> 1 rule and popcounts above 1
    pcs = [sum(int(int(x) & int(y)).bit_count() for x, y in zip(d.dither[i], d.dither[j]))
           for i in range(len(d)) for j in range(i + 1, len(d))]
    assert min(pcs) == 0 and max(pcs) > 1
    if mode == 1:
        kind, ax1, ax2 = 2, o.axis(0.0, 40.0, 10, 1.3), o.axis(0.0, 1.0, 5, 0.0)
        up_ax, up_bins = o.axis(0.0, float(1.0 - np.cos(0.05)), 8, 1.6), np.linspace(1.5, 0.9, 8)
        wo = o.WOpt(mode, 128, up_bins.ctypes.data, up_ax, 0, 1)
        args = (up_ax, list(up_bins))
    else:
        kind, ax1, ax2 = 3, o.axis(0.0, float(1.0 - np.cos(0.02)), 12, 1.6), None
        wo = o.WOpt(mode, 128, None, o.Axis(), 0, 1)
        args = (None, None)
    for b in (None, r):
        full, _, acc = o.count(kind, d, b, ax1, ax2, 2, method=0, wopt=wo)
        want = brute(o, kind, mode, d, b, ax1, ax2, 128, *args, dither=True)
        np.testing.assert_allclose(full, want, rtol=1e-13, atol=0)
        # the dither rule changes the sums (a dropped flag would pass otherwise)
        plain = brute(o, kind, mode, d, b, ax1, ax2, 128, *args, dither=False)
        assert np.abs(full - plain).max() > 1e-3 * np.abs(plain).max()
        full1, _, acc1 = o.count(kind, d, b, ax1, ax2, 2, method=1, wopt=wo)
        np.testing.assert_allclose(full1, full, rtol=1e-13)
        assert acc1 == acc
    # a catalogue without dither masks is refused
    with pytest.raises(Exception):
        o.count(kind, make_cat(o, 20, 2, 13), None, ax1, ax2, 2, method=0, wopt=wo)


def test_interp_closed_form(oracle_lib):
    """Away from segment edges the up-weight is the straight line between the
    two neighbouring bin centres (hist.cpp:386-387); below the first centre it
    is the first bin and above the last, the last."""
    o = oracle_lib
    L = o.lib()
    ax = o.axis(0.0, 0.0152, 10, 1.6)
    bins = np.array([2.0, 1.8, 1.6, 1.5, 1.3, 1.2, 1.1, 1.05, 1.0, 0.9])
    c = [L.or_axis_bin_centre(C.byref(ax), k) for k in range(10)]
    for k in range(9):
        v = c[k] + 0.3 * (c[k + 1] - c[k])
        got = L.or_interp(C.byref(ax), bins.ctypes.data, v)
        assert abs(got - (bins[k] + 0.3 * (bins[k + 1] - bins[k]))) < 1e-12
    assert L.or_interp(C.byref(ax), bins.ctypes.data, -1.0) == bins[0]
    assert L.or_interp(C.byref(ax), bins.ctypes.data, 0.5 * c[0]) == bins[0]
    assert L.or_interp(C.byref(ax), bins.ctypes.data, 1.0) == bins[-1]
