"""Seeded synthetic catalogues for the benchmark workloads and the tests.

The reference ships no data, so every workload is generated here from a fixed
seed (numpy PCG64).  Galaxy positions come out as the reference's readers take
them: (ra, dec, z) in degrees/redshift for equatorial catalogues
(io.cpp:267-336), (x, y, z) in Mpc/h for cartesian ones (io.cpp:337-405).

The comoving-distance relation used *here* only places the objects; the
pipeline itself converts (ra, dec, z) with the reference's own table
(two_pcf_amd/csrc/host/cosmology.cpp), so generation accuracy does not enter
any parity check.
"""
import numpy as np

C_KMS = 299792.458


def _dc_grid(om, h, z_max=1.2, n=24001):
    z = np.linspace(0.0, z_max, n)
    f = 1.0 / np.sqrt(om * (1 + z) ** 3 + (1 - om))
    dc = np.concatenate([[0.0], np.cumsum(0.5 * (f[1:] + f[:-1]) * np.diff(z))])
    return z, dc * C_KMS / (100.0 * h)


def uniform_cube(n, side, rng):
    return rng.uniform(0.0, side, size=(n, 3))


def _sky_uniform(n, ra_lim, dec_lim, rng):
    ra = rng.uniform(ra_lim[0], ra_lim[1], n)
    s0, s1 = np.sin(np.radians(dec_lim[0])), np.sin(np.radians(dec_lim[1]))
    dec = np.degrees(np.arcsin(rng.uniform(s0, s1, n)))
    return ra, dec


def _sky_to_cart(ra, dec, d):
    r, t = np.radians(ra), np.radians(dec)
    return np.stack([d * np.cos(t) * np.cos(r), d * np.cos(t) * np.sin(r), d * np.sin(t)], axis=1)


def neyman_scott_equatorial(n_parents, mean_children, sigma, ra_lim, dec_lim, z_lim,
                            om=0.25, h=1.0, seed=0):
    """Neyman-Scott field: parents uniform in comoving volume over the
    footprint, Poisson(mean_children) children with N(0, sigma) offsets per
    cartesian axis, then cut back to the footprint.  Returns ra, dec, z."""
    rng = np.random.Generator(np.random.PCG64(seed))
    zg, dg = _dc_grid(om, h)
    d0, d1 = np.interp(z_lim, zg, dg)
    ra, dec = _sky_uniform(n_parents, ra_lim, dec_lim, rng)
    dp = np.cbrt(rng.uniform(d0 ** 3, d1 ** 3, n_parents))
    par = _sky_to_cart(ra, dec, dp)
    nc = rng.poisson(mean_children, n_parents)
    pts = np.repeat(par, nc, axis=0) + rng.normal(0.0, sigma, size=(int(nc.sum()), 3))
    d = np.sqrt((pts ** 2).sum(axis=1))
    ra_c = np.degrees(np.arctan2(pts[:, 1], pts[:, 0])) % 360.0
    dec_c = np.degrees(np.arcsin(np.clip(pts[:, 2] / d, -1, 1)))
    z_c = np.interp(d, dg, zg)
    keep = ((ra_c >= ra_lim[0]) & (ra_c < ra_lim[1]) & (dec_c >= dec_lim[0]) & (dec_c < dec_lim[1])
            & (z_c >= z_lim[0]) & (z_c < z_lim[1]))
    return ra_c[keep], dec_c[keep], z_c[keep]


def randoms_equatorial(n, ra_lim, dec_lim, z_data, seed=0):
    """Uniform on the footprint, redshifts drawn from the data n(z)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    ra, dec = _sky_uniform(n, ra_lim, dec_lim, rng)
    z = rng.choice(z_data, size=n, replace=True)
    return ra, dec, z


def neyman_scott_box(n_target, mean_children, sigma, side, seed=0):
    """Neyman-Scott field in a periodic cube [0, side)^3 (cartesian)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    n_par = max(1, int(round(n_target / mean_children)))
    par = rng.uniform(0.0, side, size=(n_par, 3))
    nc = rng.poisson(mean_children, n_par)
    pts = np.repeat(par, nc, axis=0) + rng.normal(0.0, sigma, size=(int(nc.sum()), 3))
    return np.mod(pts, side)


def jk_grid_labels(ra, dec, ra_lim, dec_lim, n_ra, n_dec):
    """Region label n_dec*floor((ra-ra0)/dra) + floor((dec-dec0)/ddec)."""
    dra = (ra_lim[1] - ra_lim[0]) / n_ra
    ddec = (dec_lim[1] - dec_lim[0]) / n_dec
    a = np.clip(np.floor((ra - ra_lim[0]) / dra), 0, n_ra - 1).astype(np.int32)
    b = np.clip(np.floor((dec - dec_lim[0]) / ddec), 0, n_dec - 1).astype(np.int32)
    return (a * n_dec + b).astype(np.int32)


# --------------------------------------------------------------- workloads --

FOOT_RA, FOOT_DEC, FOOT_Z = (150.0, 250.0), (0.0, 60.0), (0.2, 0.6)


def jackknife_batch(scale=1.0, seed=20260102):
    """The 2M clustered data + 10M randoms jackknife batch: Neyman-Scott with
    40k parents x Poisson(50) at 5 Mpc/h over ra 150-250, dec 0-60,
    z 0.2-0.6; 10x10 ra/dec grid of 100 regions; unit weights.
    `scale` < 1 thins the parents and randoms proportionally (test sizes)."""
    n_par = max(1, int(round(40000 * scale)))
    ra, dec, z = neyman_scott_equatorial(n_par, 50.0, 5.0, FOOT_RA, FOOT_DEC, FOOT_Z, seed=seed)
    n_r = max(1, int(round(10_000_000 * scale)))
    rra, rdec, rz = randoms_equatorial(n_r, FOOT_RA, FOOT_DEC, z, seed=seed + 1)
    return dict(
        data=dict(ra=ra, dec=dec, z=z, w=np.ones(ra.size, np.float32),
                  jk=jk_grid_labels(ra, dec, FOOT_RA, FOOT_DEC, 10, 10)),
        randoms=dict(ra=rra, dec=rdec, z=rz, w=np.ones(rra.size, np.float32),
                     jk=jk_grid_labels(rra, rdec, FOOT_RA, FOOT_DEC, 10, 10)),
        n_jk=100, om=0.25, h=1.0, z_min=0.0, z_max=1.0)


def equatorial_smu_sigma_pi(scale=1.0, seed=20260101):
    """1M clustered galaxies (20k parents) + 5M randoms, weights U[0.5,1.5]."""
    n_par = max(1, int(round(20000 * scale)))
    ra, dec, z = neyman_scott_equatorial(n_par, 50.0, 5.0, FOOT_RA, FOOT_DEC, FOOT_Z, seed=seed)
    n_r = max(1, int(round(5_000_000 * scale)))
    rra, rdec, rz = randoms_equatorial(n_r, FOOT_RA, FOOT_DEC, z, seed=seed + 1)
    rng = np.random.Generator(np.random.PCG64(seed + 2))
    return dict(
        data=dict(ra=ra, dec=dec, z=z, w=rng.uniform(0.5, 1.5, ra.size).astype(np.float32),
                  jk=np.zeros(ra.size, np.int32)),
        randoms=dict(ra=rra, dec=rdec, z=rz, w=rng.uniform(0.5, 1.5, rra.size).astype(np.float32),
                     jk=np.zeros(rra.size, np.int32)),
        n_jk=1, om=0.25, h=1.0, z_min=0.0, z_max=1.0)


def uniform_monopole(seed=20260101):
    """50k data + 100k randoms uniform in [0,500)^3 Mpc/h, unit weights."""
    rng = np.random.Generator(np.random.PCG64(seed))
    d = uniform_cube(50_000, 500.0, rng)
    r = uniform_cube(100_000, 500.0, rng)
    return dict(data=d, randoms=r)


def uneven_mocks(n_mocks=64, n_min=20_000, n_max=2_000_000, density=3e-4, seed=20260103):
    """Independent cartesian mocks of uneven size: N_k log-uniform in
    [n_min, n_max], Neyman-Scott (N_k / 50 parents, Poisson(50) children,
    5 Mpc/h) in a cube of side (N_k / density)^(1/3), each with 5 N_k uniform
    randoms in the same cube; unit weights.  Returns a list of dicts with
    `data` and `randoms` position arrays (n, 3) and the cube side."""
    rng = np.random.Generator(np.random.PCG64(seed))
    sizes = np.round(np.exp(rng.uniform(np.log(n_min), np.log(n_max), n_mocks))).astype(np.int64)
    out = []
    for k, n in enumerate(sizes):
        side = float((n / density) ** (1.0 / 3.0))
        d = neyman_scott_box(int(n), 50.0, 5.0, side, seed=seed + 17 * (k + 1))
        r = uniform_cube(5 * int(n), side, np.random.Generator(np.random.PCG64(seed + 17 * (k + 1) + 1)))
        out.append(dict(data=d, randoms=r, side=side))
    return out


def random_masks(n, n_words, p, seed):
    """n x n_words uint64 masks, bit k of object i set with probability p[i]
    (built one 64-bit word at a time to bound memory)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = np.empty((n, n_words), dtype=np.uint64)
    for k in range(n_words):
        bits = rng.random((n, 64)) < p[:, None]
        out[:, k] = np.packbits(bits.astype(np.uint8), axis=1, bitorder="little").view(np.uint64)[:, 0]
    return out


def popcount_rows(masks):
    """Number of set bits per row of a uint64 mask array."""
    m = masks.copy()
    m = m - ((m >> np.uint64(1)) & np.uint64(0x5555555555555555))
    m = (m & np.uint64(0x3333333333333333)) + ((m >> np.uint64(2)) & np.uint64(0x3333333333333333))
    m = (m + (m >> np.uint64(4))) & np.uint64(0x0F0F0F0F0F0F0F0F)
    m = (m * np.uint64(0x0101010101010101)) >> np.uint64(56)
    return m.sum(axis=1).astype(np.int64)


def inverse_p_catalogs(scale=1.0, n_words=32, seed=20260104):
    """500k clustered equatorial galaxies (10k parents x Poisson(50), 5 Mpc/h,
    ra 150-250, dec 0-60, z 0.2-0.6) with n_words x 64-bit masks, each bit set
    with a per-galaxy probability p ~ U[0.3, 0.95]; 2.5M randoms from the same
    footprint and n(z) with all-ones masks; weights U[0.5, 1.5].  Returns the
    input weights `iw` and the weights after the inverse-probability rule at
    read time (io.cpp:578-582) as `w`."""
    n_runs = 64 * n_words
    n_par = max(1, int(round(10000 * scale)))
    ra, dec, z = neyman_scott_equatorial(n_par, 50.0, 5.0, FOOT_RA, FOOT_DEC, FOOT_Z, seed=seed)
    n_r = max(1, int(round(2_500_000 * scale)))
    rra, rdec, rz = randoms_equatorial(n_r, FOOT_RA, FOOT_DEC, z, seed=seed + 1)
    rng = np.random.Generator(np.random.PCG64(seed + 2))
    p = rng.uniform(0.3, 0.95, ra.size)
    dmask = random_masks(ra.size, n_words, p, seed + 3)
    rmask = np.full((rra.size, n_words), np.uint64(0xFFFFFFFFFFFFFFFF), dtype=np.uint64)
    out = {}
    for tag, (a, b, c, m) in (("data", (ra, dec, z, dmask)), ("randoms", (rra, rdec, rz, rmask))):
        iw = rng.uniform(0.5, 1.5, a.size).astype(np.float32)
        tot = popcount_rows(m)
        tot[tot == 0] = 1
        w = np.where(tot != n_runs, (iw.astype(np.float64) * (n_runs / tot)).astype(np.float32), iw)
        out[tag] = dict(ra=a, dec=b, z=c, w=w, iw=iw, masks=m, jk=np.zeros(a.size, np.int32))
    out.update(n_runs=n_runs, om=0.25, h=1.0, z_min=0.0, z_max=1.0)
    return out
