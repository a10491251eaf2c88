"""Benchmark: s-mu DD+DR+RR pair counting on the jackknife batch.

Workload: 2M Neyman-Scott galaxies (40k parents x Poisson(50), 5 Mpc/h)
+ 10M randoms over ra 150-250, dec 0-60, z 0.2-0.6 (omega_m 0.25, h 1),
100 jackknife regions (10x10 ra/dec grid), unit weights, s-mu 40 log bins
(base 1.3, s < 100 Mpc/h) x 50 mu bins.  One step = DD + DR + RR counted
with the batched FP32 + edge-recheck method in one persistent launch, inputs
already resident in HBM.  Value = accepted (in-range, binned) pairs per
second of step wall time.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--scale S]

Prints one JSON line.  Also measures the plain FP64 kernel on the same
workload, the FP32 roofline of the counting kernel (26 FP32 flop per
accepted s-mu pair, HIP-event timing on the library's stream) and the
reference's OpenMP path on a bounded sample.
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FLOP_PER_PAIR = 26            # FP32 flop per accepted s-mu pair
# per accepted pair and correlator kind: monopole 15 (d 3, s^2 5, range 2,
# sqrt 1, log index 3, weight 1), s-mu 26 (adds n1+n2 3, pi dot 5, x1/2 1,
# mu divide 1, mu index 1), sigma-pi 29 (monopole's without sqrt and log
# index, plus n1+n2 3, pi dot 5, x1/2 1, s^2 - pi^2 2, sqrt 1, sigma log
# index 3, |pi| index 3), angular 15 (chord 3, chord^2 5, range 2, x1/2 1,
# log index 3, weight 1)
FLOP_BY_KIND = {0: 15, 1: 29, 2: 26, 3: 15}
# HBM bytes per launch of the counting kernel on the default workload (scale
# 1, 100 regions, tuning (2, 12, 256) with 8 region slots fitting), from two
# rocprofv3 --pmc runs: FETCH_SIZE x 2 + WRITE_SIZE x 1 (factors from a 1 GiB
# copy), KiB -> B.  Keyed by the first 16 hex digits of the SHA-256 of the
# kernel source they were measured on: for any other source the line reports
# no traffic.  Source: profiles/k_count_fp32_jackknife_batch_hbm.md
KERNEL_SRC = os.path.join(ROOT, "two_pcf_amd", "csrc", "twopcf_hip.hip")
TRAFFIC_BYTES = {("3786eef705ed7a8e", 1.0, 100, 2, 8, 256): (4.1820e7 * 2 + 2.5050e6) * 1024}
# algorithmic HBM bytes of the counting kernel: every object a work item
# reads (its first-side objects and its stencil rows' second-side objects,
# the library's n_staged) once, at the bytes one object takes in the SoA:
# x, y, z, nx, ny, nz (f64), plus the region (i32) when there are regions
# and the weight (f32) when weighted
OBJ_BYTES_POS = 48
# region slots that fit in LDS on the default workload (2000-bin u32
# histograms next to the 1024-entry tile, queues, tables and two copies of
# the full histogram)
SLOTS_FIT = 8
PEAK_FP32_TFLOPS = 157.3      # MI355X dense FP32 peak (matrix cores; the vector peak is the same figure)
AXES = [(0.0, 100.0, 40, 1.3), (0.0, 1.0, 50, 0.0)]


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def make_catalogs(scale):
    from two_pcf_amd import lib as L
    from two_pcf_amd import synth
    t0 = time.time()
    c = synth.jackknife_batch(scale=scale)
    out = {}
    for tag in ("data", "randoms"):
        a = c[tag]
        d = L.comoving_distance(c["om"], c["h"], c["z_min"], c["z_max"], a["z"])
        x, y, z, nx, ny, nz = L.eq_to_cart(a["ra"], a["dec"], d)
        out[tag] = dict(x=x, y=y, z=z, nx=nx, ny=ny, nz=nz, w=a["w"], jk=a["jk"],
                        ra=a["ra"], dec=a["dec"], zz=a["z"])
    log(f"catalogues: {out['data']['x'].size} data, {out['randoms']['x'].size} randoms "
        f"({time.time() - t0:.1f} s)")
    return out, c["n_jk"]


def cpu_threads():
    """Host threads for the reference's OpenMP path: every core this process
    may run on, within OMP_NUM_THREADS when that is set (a GPU host's share)."""
    n = len(os.sched_getaffinity(0))
    omp = os.environ.get("OMP_NUM_THREADS", "")
    return min(n, int(omp)) if omp.isdigit() and int(omp) > 0 else n


def host_cpu():
    """lscpu's model name, sockets, cores per socket and threads per core."""
    want = {"Model name": "model", "Socket(s)": "sockets", "Core(s) per socket": "cores_per_socket",
            "Thread(s) per core": "threads_per_core", "CPU(s)": "cpus"}
    out = {}
    try:
        text = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=30).stdout
    except (OSError, subprocess.SubprocessError):
        return out
    for line in text.splitlines():
        k, _, v = line.partition(":")
        if k.strip() in want:
            out[want[k.strip()]] = v.strip()
    out["affinity_cpus"] = len(os.sched_getaffinity(0))
    out["omp_num_threads_env"] = os.environ.get("OMP_NUM_THREADS")
    return out


REF_PARAMS = dict(data_file_type="ascii", random_file_type="ascii", coord_system="equatorial",
                  ra_x_dataset_name="x", dec_y_dataset_name="y", z_z_dataset_name="z", weight_dataset_name="w",
                  jk_dataset_name="jk", use_weights=0, n_jk_regions=0, omega_m=0.25, h=1.0, z_min=0.0, z_max=1.0,
                  plot_monopole=0, monopole_filename="mono.txt", monopole_output_type="ascii", monopole_log_base=1.3,
                  monopole_min=0.0, monopole_max=100.0, monopole_n_bins=30,
                  plot_sigma_pi=0, sigma_pi_filename="sp.txt", sigma_pi_output_type="ascii", sigma_log_base=1.0,
                  sigma_min=0.0, sigma_max=50.0, sigma_n_bins=50, pi_log_base=1.0, pi_min=0.0, pi_max=50.0,
                  pi_n_bins=50, plot_s_mu=0, s_mu_filename="smu.txt", s_mu_output_type="ascii", s_log_base=1.3,
                  s_min=0.0, s_max=100.0, s_n_bins=40, mu_n_bins=50,
                  calculate_angular_dd=0, calculate_angular_dd_invpweights=0, calculate_angular_dr=0,
                  calculate_angular_dr_invpweights=0, angular_dd_filename="add.h5",
                  angular_dd_invpweights_filename="addinv.h5", angular_dr_filename="adr.h5",
                  angular_dr_invpweights_filename="adrinv.h5", theta_max=0.175, theta_n_bins=40,
                  theta_log_base=1.6, healpix_order=5, n_bitwise_runs=2048, bitwise_weight_dataset_name="BITWEIGHT",
                  dither_weight_dataset_name="DITHER")


def run_reference(binary, params, cols, td, extra_args=("0", "0")):
    """Run a reference-linked driver (oracle/_ref, compiled from the
    reference's sources) on catalogues written as ascii (cols: tag -> list of
    columns; formats as the reference's reader takes them) or given as
    files; returns (its JSON, counting seconds, end-to-end seconds)."""
    p = dict(REF_PARAMS)
    p.update(params, n_threads=cpu_threads())
    for tag, key in (("data", "data_filename"), ("randoms", "random_filename")):
        c = cols[tag]
        if isinstance(c, str):
            p[key] = c
            continue
        fn = os.path.join(td, tag + ".txt")
        fmt = " ".join("%d" if np.issubdtype(np.asarray(v).dtype, np.integer) else
                       ("%.9g" if np.asarray(v).dtype == np.float32 else "%.17g") for v in c) + "\n"
        with open(fn, "w") as f:
            for row in zip(*[np.asarray(v).tolist() for v in c]):
                f.write(fmt % row)
        p[key] = fn
    pf = os.path.join(td, "params.txt")
    with open(pf, "w") as f:
        f.write("".join(f"{k} = {v}\n" for k, v in p.items()))
    js = os.path.join(td, "out.json")
    t0 = time.perf_counter()
    subprocess.run([binary, pf, js, *extra_args], check=True, stdout=subprocess.DEVNULL, cwd=td)
    wall = time.perf_counter() - t0
    with open(js) as f:
        r = json.load(f)
    secs = sum(sum(v["seconds"]) for k, v in r.items() if isinstance(v, dict) and "seconds" in v)
    secs += r.get("angular_seconds", 0.0)
    return r, secs, wall


def ref_binary(name):
    b = os.path.join(ROOT, "oracle", "_ref", name)
    if not os.path.exists(b):
        raise FileNotFoundError(f"{b} is not built")
    return b


def subsample(n, frac, rng):
    return rng.random(n) < frac


def baseline_record(pairs, secs, wall, what):
    return dict(value=pairs / secs, unit="pairs/s", cores=cpu_threads(), kind="reference", sample=what,
                counting_s=secs, end_to_end_s=wall, host=host_cpu())


def uniform_monopole_full():
    """The uniform monopole case in full on the reference: 50k data + 100k
    randoms uniform in [0, 500)^3, 30 log bins base 1.3 to 100, DD, DR, RR."""
    from two_pcf_amd import synth
    c = synth.uniform_monopole()
    with tempfile.TemporaryDirectory() as td:
        r, secs, wall = run_reference(ref_binary("ref_counts"), dict(coord_system="cartesian", plot_monopole=1),
                                      {t: [c[t][:, 0], c[t][:, 1], c[t][:, 2]] for t in ("data", "randoms")}, td)
    m = r["monopole"]
    pairs = sum(m["DD"]["full"]) / 2 + sum(m["DR"]["full"]) + sum(m["RR"]["full"]) / 2
    return dict(workload="uniform monopole 50k + 100k in [0,500)^3, 30 log bins to 100, DD+DR+RR, full run",
                pairs=pairs, counting_s=secs, end_to_end_s=wall, pairs_per_s=pairs / secs, cores=cpu_threads())


def gpu_pairs(L, ctx, jobs_of_sample):
    """Accepted pairs of the sample, counted by the library (the reference's
    weighted sums do not give them)."""
    res, t = ctx.count(jobs_of_sample, method=L.FP32_RECHECK)
    return t["n_pairs"]


def cpu_baseline_jackknife(cats, n_jk, frac, seed=5):
    """The reference's OpenMP s-mu path (oracle/_ref/ref_counts) on a seeded
    subsample of both catalogues (frac of the objects, frac^2 of the pairs);
    rate = accepted pairs / counting seconds.  Also the uniform monopole case
    in full."""
    rng = np.random.Generator(np.random.PCG64(seed))
    keep = {t: subsample(cats[t]["x"].size, frac, rng) for t in ("data", "randoms")}
    cols = {t: [cats[t]["ra"][keep[t]], cats[t]["dec"][keep[t]], cats[t]["zz"][keep[t]], cats[t]["jk"][keep[t]]]
            for t in ("data", "randoms")}
    with tempfile.TemporaryDirectory() as td:
        r, secs, wall = run_reference(ref_binary("ref_counts"), dict(plot_s_mu=1, n_jk_regions=n_jk), cols, td)
    m = r["s_mu"]
    pairs = sum(m["DD"]["full"]) / 2 + sum(m["DR"]["full"]) + sum(m["RR"]["full"]) / 2
    out = baseline_record(pairs, secs, wall, f"seeded {frac:.0%} subsample of data and randoms "
                          f"({int(keep['data'].sum())} + {int(keep['randoms'].sum())} objects, {pairs:.3g} pairs, "
                          f"{frac * frac:.2%} of the pairs)")
    out["uniform_monopole_full"] = uniform_monopole_full()
    return out


def copy_calibration(mib):
    """One device-to-device copy of `mib` MiB (HIP's blit kernel): the
    reference dispatch whose FETCH_SIZE / WRITE_SIZE counters calibrate the
    HBM-traffic figures of the counting kernel."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    n = mib << 20
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(n)) == 0 and hip.hipMalloc(C.byref(b), C.c_size_t(n)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(n)) == 0
    assert hip.hipDeviceSynchronize() == 0
    assert hip.hipMemcpy(b, a, C.c_size_t(n), 3) == 0   # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_calibration_mib": mib}), flush=True)


def workload_jackknife(args, L, ctx):
    cats, n_jk = make_catalogs(args.scale)
    if args.n_jk:
        n_jk = args.n_jk
    t_up = time.perf_counter()
    dc = {t: ctx.upload(a["x"], a["y"], a["z"], a["nx"], a["ny"], a["nz"], a["w"], a["jk"])
          for t, a in cats.items()}
    upload_ms = 1000.0 * (time.perf_counter() - t_up)
    jobs = [dict(kind=L.S_MU, a=dc["data"], b=None, axes=AXES, n_jk=n_jk),
            dict(kind=L.S_MU, a=dc["data"], b=dc["randoms"], axes=AXES, n_jk=n_jk),
            dict(kind=L.S_MU, a=dc["randoms"], b=None, axes=AXES, n_jk=n_jk)]
    config = {"workload": "jackknife batch: 2M clustered data + 10M randoms, 100 jk regions, "
                          "s-mu 40 log bins (base 1.3, s<100) x 50 mu, DD+DR+RR in one launch",
              "n_data": int(cats["data"]["x"].size), "n_randoms": int(cats["randoms"]["x"].size),
              "scale": args.scale, "n_jk": n_jk}
    return dict(jobs=jobs, upload_ms=upload_ms, config=config, traffic_key=(args.scale, n_jk),
                obj_bytes=OBJ_BYTES_POS + (4 if n_jk > 1 else 0),
                cpu=lambda: cpu_baseline_jackknife(cats, n_jk, args.cpu_frac))


MONO_AXES = [(0.0, 100.0, 30, 1.3)]


def workload_multi(args, L, ctx):
    """64 independent cartesian mocks, sizes log-uniform in 20k-2M (scaled by
    --scale), Neyman-Scott at density 3e-4 (Mpc/h)^-3, 5x uniform randoms
    each; monopole (30 log bins base 1.3 to 100) and s-mu (40 x 50) DD, DR,
    RR of every mock in one call: each monopole is binned in the pass of the
    s-mu job on the same catalogues, one persistent launch captured in a
    hipGraph."""
    from two_pcf_amd import synth
    t0 = time.time()
    mocks = synth.uneven_mocks(n_mocks=64, n_min=max(50, int(20_000 * args.scale)),
                               n_max=max(100, int(2_000_000 * args.scale)))
    nd = sum(m["data"].shape[0] for m in mocks)
    nr = sum(m["randoms"].shape[0] for m in mocks)
    log(f"mocks: 64, {nd} data + {nr} randoms ({time.time() - t0:.1f} s)")
    up_ms = 0.0
    jobs = []
    for m in mocks:
        dc = {}
        for tag in ("data", "randoms"):
            p = m[tag]
            x, y, z = (np.ascontiguousarray(p[:, k]) for k in range(3))
            _, nx, ny, nz = L.cart_norm(x, y, z)
            t_up = time.perf_counter()
            dc[tag] = ctx.upload(x, y, z, nx, ny, nz)
            up_ms += 1000.0 * (time.perf_counter() - t_up)
        for kind, axes in ((L.MONOPOLE, MONO_AXES), (L.S_MU, AXES)):
            jobs += [dict(kind=kind, a=dc["data"], b=None, axes=axes, n_jk=1),
                     dict(kind=kind, a=dc["data"], b=dc["randoms"], axes=axes, n_jk=1),
                     dict(kind=kind, a=dc["randoms"], b=None, axes=axes, n_jk=1)]
    config = {"workload": "uneven multi-catalogue batch: 64 cartesian Neyman-Scott mocks, sizes log-uniform "
                          "in 20k-2M, density 3e-4, 5x randoms; monopole (30 log bins) + s-mu (40x50) "
                          "DD+DR+RR of every mock in one call",
              "n_data": nd, "n_randoms": nr, "scale": args.scale, "n_jobs": len(jobs)}

    def cpu(frac=args.cpu_frac, seed=6):
        """The reference (ref_counts: monopole then s-mu, DD, DR, RR) on every
        8th mock by size, a seeded frac of each mock's objects."""
        rng = np.random.Generator(np.random.PCG64(seed))
        order = sorted(range(len(mocks)), key=lambda k: mocks[k]["data"].shape[0])
        pick = order[::8]
        secs = wall = 0.0
        sample_jobs, nobj = [], 0
        for k in pick:
            cols, dc = {}, {}
            for tag in ("data", "randoms"):
                p = mocks[k][tag][subsample(mocks[k][tag].shape[0], frac, rng)]
                nobj += p.shape[0]
                cols[tag] = [p[:, 0], p[:, 1], p[:, 2]]
                x, y, z = (np.ascontiguousarray(p[:, i]) for i in range(3))
                _, nx, ny, nz = L.cart_norm(x, y, z)
                dc[tag] = ctx.upload(x, y, z, nx, ny, nz)
            with tempfile.TemporaryDirectory() as td:
                _, s_, w_ = run_reference(ref_binary("ref_counts"),
                                          dict(coord_system="cartesian", plot_monopole=1, plot_s_mu=1), cols, td)
            secs += s_
            wall += w_
            for kind, axes in ((L.MONOPOLE, MONO_AXES), (L.S_MU, AXES)):
                sample_jobs += [dict(kind=kind, a=dc["data"], b=None, axes=axes, n_jk=1),
                                dict(kind=kind, a=dc["data"], b=dc["randoms"], axes=axes, n_jk=1),
                                dict(kind=kind, a=dc["randoms"], b=None, axes=axes, n_jk=1)]
        pairs = gpu_pairs(L, ctx, sample_jobs)
        return baseline_record(pairs, secs, wall, f"{len(pick)} of the 64 mocks (every 8th by size), seeded "
                               f"{frac:.0%} of their objects ({nobj} objects, {pairs:.3g} monopole + s-mu pairs)")

    return dict(jobs=jobs, upload_ms=up_ms, config=config, traffic_key=None, cpu=cpu)


def _eq_upload(L, ctx, a, c, masks=False):
    d = L.comoving_distance(c["om"], c["h"], c["z_min"], c["z_max"], a["z"])
    x, y, z, nx, ny, nz = L.eq_to_cart(a["ra"], a["dec"], d)
    return ctx.upload(x, y, z, nx, ny, nz, a["w"], a["jk"], iw=a.get("iw"), masks=a.get("masks") if masks else None)


SP_AXES = [(0.0, 50.0, 50, 1.0), (0.0, 50.0, 50, 1.0)]


def workload_equatorial(args, L, ctx):
    """1M clustered galaxies (20k parents x Poisson(50), 5 Mpc/h) + 5M randoms
    over ra 150-250, dec 0-60, z 0.2-0.6, weights U[0.5, 1.5]: s-mu (40 log
    bins x 50 mu) and sigma-pi (50 x 50 linear to 50 Mpc/h) DD, DR and RR in
    one call."""
    from two_pcf_amd import synth
    c = synth.equatorial_smu_sigma_pi(scale=args.scale)
    t_up = time.perf_counter()
    dc = {t: _eq_upload(L, ctx, c[t], c) for t in ("data", "randoms")}
    upload_ms = 1000.0 * (time.perf_counter() - t_up)
    jobs = []
    for kind, axes in ((L.S_MU, AXES), (L.SIGMA_PI, SP_AXES)):
        jobs += [dict(kind=kind, a=dc["data"], b=None, axes=axes, n_jk=1),
                 dict(kind=kind, a=dc["data"], b=dc["randoms"], axes=axes, n_jk=1),
                 dict(kind=kind, a=dc["randoms"], b=None, axes=axes, n_jk=1)]
    config = {"workload": "equatorial s-mu + sigma-pi: 1M clustered galaxies + 5M randoms, weights "
                          "U[0.5,1.5]; s-mu 40 log bins x 50 mu, sigma-pi 50x50 linear to 50; DD+DR+RR in one call",
              "n_data": int(c["data"]["ra"].size), "n_randoms": int(c["randoms"]["ra"].size), "scale": args.scale}

    def cpu(frac=args.cpu_frac, seed=7):
        """The reference (ref_counts: sigma-pi then s-mu, weighted) on a seeded
        frac of the objects."""
        rng = np.random.Generator(np.random.PCG64(seed))
        sub, cols = {}, {}
        for tag in ("data", "randoms"):
            a = c[tag]
            k = subsample(a["ra"].size, frac, rng)
            sub[tag] = {f: a[f][k] for f in ("ra", "dec", "z", "w", "jk")}
            cols[tag] = [sub[tag]["ra"], sub[tag]["dec"], sub[tag]["z"], sub[tag]["w"]]
        with tempfile.TemporaryDirectory() as td:
            _, secs, wall = run_reference(ref_binary("ref_counts"),
                                          dict(use_weights=1, plot_sigma_pi=1, plot_s_mu=1), cols, td)
        sd = {t: _eq_upload(L, ctx, sub[t], c) for t in ("data", "randoms")}
        sample_jobs = []
        for kind, axes in ((L.S_MU, AXES), (L.SIGMA_PI, SP_AXES)):
            sample_jobs += [dict(kind=kind, a=sd["data"], b=None, axes=axes, n_jk=1),
                            dict(kind=kind, a=sd["data"], b=sd["randoms"], axes=axes, n_jk=1),
                            dict(kind=kind, a=sd["randoms"], b=None, axes=axes, n_jk=1)]
        pairs = gpu_pairs(L, ctx, sample_jobs)
        return baseline_record(pairs, secs, wall, f"seeded {frac:.0%} subsample of data and randoms "
                               f"({sub['data']['ra'].size} + {sub['randoms']['ra'].size} objects, {pairs:.3g} "
                               f"s-mu + sigma-pi pairs, {frac * frac:.2%} of the pairs)")

    return dict(jobs=jobs, upload_ms=upload_ms, config=config, traffic_key=None, cpu=cpu)


ANG_AXIS = (0.0, float(1.0 - np.cos(0.175)), 40, 1.6)


def workload_invp(args, L, ctx):
    """Bitwise inverse-p weights: 500k clustered galaxies with 2048-bit masks
    (32 words, bit probability U[0.3, 0.95] per galaxy) + 2.5M randoms with
    all-ones masks.  A step is the pipeline of the reference's inverse-p build:
    angular DD and DR with parent and inverse-p weights (theta < 0.175 rad, 40
    log bins base 1.6), the up-weights DD/DDinvp and DR/DRinvp, then s-mu DD
    (inverse-p weights), DR (angular up-weight) and RR."""
    from two_pcf_amd import synth
    t0 = time.time()
    c = synth.inverse_p_catalogs(scale=args.scale)
    log(f"inverse-p catalogues ({time.time() - t0:.1f} s)")
    n_runs = c["n_runs"]
    t_up = time.perf_counter()
    dc = {t: _eq_upload(L, ctx, c[t], c, masks=True) for t in ("data", "randoms")}
    upload_ms = 1000.0 * (time.perf_counter() - t_up)
    def angular(D, R):
        return [dict(kind=L.ANGULAR, a=D, b=None, axes=[ANG_AXIS], weight_mode=L.W_ANG_IW, n_runs=n_runs),
                dict(kind=L.ANGULAR, a=D, b=None, axes=[ANG_AXIS], weight_mode=L.W_ANG_INVP, n_runs=n_runs),
                dict(kind=L.ANGULAR, a=D, b=R, axes=[ANG_AXIS], weight_mode=L.W_ANG_IW, n_runs=n_runs),
                dict(kind=L.ANGULAR, a=D, b=R, axes=[ANG_AXIS], weight_mode=L.W_ANG_W, n_runs=n_runs)]

    def threed_of(D, R, up_dd, up_dr):
        return [dict(kind=L.S_MU, a=D, b=None, axes=AXES, weight_mode=L.W_INVP_3D, n_runs=n_runs,
                     up_bins=up_dd, up_axis=ANG_AXIS),
                dict(kind=L.S_MU, a=D, b=R, axes=AXES, weight_mode=L.W_ANGUP_3D, n_runs=n_runs,
                     up_bins=up_dr, up_axis=ANG_AXIS),
                dict(kind=L.S_MU, a=R, b=None, axes=AXES)]

    D, R = dc["data"], dc["randoms"]
    ang = angular(D, R)

    def threed(up_dd, up_dr):
        return threed_of(D, R, up_dd, up_dr)

    def step(ang, threed, method, tu):
        r1, t1 = ctx.count(ang, method=method, tuning=tu)
        fin = [L.finish(j.get("b") is None, r["full"], r["jk"])[0] for r, j in zip(r1, ang)]
        with np.errstate(invalid="ignore", divide="ignore"):
            up_dd, up_dr = fin[0] / fin[1], fin[2] / fin[3]
        r2, t2 = ctx.count(threed(up_dd, up_dr), method=method, tuning=tu)
        t = {k: t1[k] + t2[k] for k in ("kernel_ms", "total_ms", "prep_ms", "n_work_items", "n_rechecked", "n_pairs")}
        t["angular_kernel_ms"], t["threed_kernel_ms"] = t1["kernel_ms"], t2["kernel_ms"]
        log(f"  angular {t1['kernel_ms']:.1f} ms ({t1['n_pairs']} pairs), 3-D {t2['kernel_ms']:.1f} ms "
            f"({t2['n_pairs']} pairs)")
        return r1 + r2, t

    def run(method, tu):
        return step(ang, threed, method, tu)

    def cpu(frac=min(args.cpu_frac, 0.1), seed=8):
        """The reference's inverse-p build (ref_invp32: _USE_INV_WEIGHTS,
        _N_MASK_INTS=32, hdf5 catalogues with BITWEIGHT0..31) on a seeded frac
        of the objects: calc_angular_upweight, then s-mu DD (inverse-p), DR
        (up-weight) and RR."""
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import h5mini
        rng = np.random.Generator(np.random.PCG64(seed))
        sub = {}
        with tempfile.TemporaryDirectory() as td:
            files = {}
            for tag in ("data", "randoms"):
                a = c[tag]
                k = subsample(a["ra"].size, frac, rng)
                sub[tag] = {f: a[f][k] for f in ("ra", "dec", "z", "w", "iw", "jk", "masks")}
                cols = dict(RA=sub[tag]["ra"], DEC=sub[tag]["dec"], Z=sub[tag]["z"],
                            W=sub[tag]["iw"].astype(np.float64))
                for w_ in range(sub[tag]["masks"].shape[1]):
                    cols[f"BITWEIGHT{w_}"] = np.ascontiguousarray(sub[tag]["masks"][:, w_]).view(np.int64)
                files[tag] = os.path.join(td, tag + ".h5")
                h5mini.write(files[tag], cols)
            _, secs, wall = run_reference(
                ref_binary("ref_invp32"),
                dict(data_file_type="hdf5", random_file_type="hdf5", ra_x_dataset_name="RA", dec_y_dataset_name="DEC",
                     z_z_dataset_name="Z", weight_dataset_name="W", use_weights=1, n_bitwise_runs=n_runs,
                     calculate_angular_dd=1, calculate_angular_dd_invpweights=1, calculate_angular_dr=1,
                     calculate_angular_dr_invpweights=1, plot_s_mu=1),
                files, td, extra_args=("none", "0"))
        sd = {t: _eq_upload(L, ctx, sub[t], c, masks=True) for t in ("data", "randoms")}
        _, t = step(angular(sd["data"], sd["randoms"]),
                    lambda u, v: threed_of(sd["data"], sd["randoms"], u, v), L.FP32_RECHECK, None)
        pairs = t["n_pairs"]
        return baseline_record(pairs, secs, wall, f"seeded {frac:.0%} subsample of data and randoms "
                               f"({sub['data']['ra'].size} + {sub['randoms']['ra'].size} objects, {pairs:.3g} "
                               f"angular + s-mu pairs, {frac * frac:.2%} of the pairs)")

    config = {"workload": "bitwise inverse-p: 500k clustered galaxies with 2048-bit masks + 2.5M randoms; "
                          "angular DD/DR (parent and inverse-p weights, theta<0.175, 40 log bins) then s-mu "
                          "DD (inverse-p) / DR (up-weight) / RR, as two calls per step",
              "n_data": int(c["data"]["ra"].size), "n_randoms": int(c["randoms"]["ra"].size), "scale": args.scale,
              "n_runs": n_runs}
    all_jobs = ang + threed(None, None)
    return dict(jobs=ang, all_jobs=all_jobs, upload_ms=upload_ms, config=config, traffic_key=None, cpu=cpu,
                run=run, tunable=False)


WORKLOADS = {"jackknife": lambda a, L, c: workload_jackknife(a, L, c),
             "multi": lambda a, L, c: workload_multi(a, L, c),
             "equatorial": lambda a, L, c: workload_equatorial(a, L, c),
             "invp": lambda a, L, c: workload_invp(a, L, c)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--plain-steps", type=int, default=1)
    ap.add_argument("--cpu-frac", type=float, default=0.2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--subdiv", type=int, default=0)
    ap.add_argument("--slots", type=int, default=0)
    ap.add_argument("--n-jk", type=int, default=0, help="override the number of jackknife regions (1: none)")
    ap.add_argument("--no-autotune", action="store_true", help="use the default tuning instead of the autotuner's")
    ap.add_argument("--workload", choices=["jackknife", "multi", "equatorial", "invp"], default="jackknife",
                    help="jackknife: the 2M+10M s-mu jackknife batch (default); multi: 64 uneven "
                         "cartesian mocks, monopole + s-mu of each in one call; equatorial: 1M+5M weighted "
                         "s-mu + sigma-pi; invp: 500k+2.5M with 2048-bit masks, angular DD/DR then s-mu "
                         "with inverse-p weights")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="only run a device copy of this size (counter calibration)")
    args = ap.parse_args()
    if args.calib_copy_mib:
        copy_calibration(args.calib_copy_mib)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: there is no multi-GPU path "
              "(--gpus must be 1)", file=sys.stderr)
        sys.exit(2)

    from two_pcf_amd import lib as L
    ctx = L.Context(0)
    log("device", ctx.info())
    W = WORKLOADS[args.workload](args, L, ctx)
    jobs, upload_ms = W["jobs"], W["upload_ms"]
    W.setdefault("all_jobs", jobs)
    tuning = {}
    if args.subdiv:
        tuning["cell_subdiv"] = args.subdiv
    if args.slots:
        tuning["max_slots"] = args.slots
    tuned = None
    run = W.get("run") or (lambda method, tu: ctx.count(jobs, method=method, tuning=tu))
    ref = None
    first_call_ms = None
    if not tuning and not args.no_autotune and W.get("tunable", True):
        # the first call as a single call makes it (default tuning, catalogues
        # not yet gridded), timed before the autotuner grids them
        _, t = run(L.FP32_RECHECK, None)
        first_call_ms = t["total_ms"]
        log(f"first call (default tuning): total {t['total_ms']:.1f} ms, prep {t['prep_ms']:.1f} ms")
        t0 = time.perf_counter()
        tu, est = ctx.autotune(jobs)
        tuned = dict(choice=tu, estimated_kernel_ms=est, autotune_s=time.perf_counter() - t0)
        log(f"autotune: {tu} (estimated kernel {est:.1f} ms, {tuned['autotune_s']:.1f} s)")

    for i in range(args.warmup):
        ref, t = run(L.FP32_RECHECK, tuning or None)
        if first_call_ms is None:
            first_call_ms = t["total_ms"]   # gridding + planning + graph capture + one run
        log(f"warmup {i}: kernel {t['kernel_ms']:.1f} ms, total {t['total_ms']:.1f} ms, "
            f"prep {t['prep_ms']:.1f} ms, items {t['n_work_items']}, rechecked {t['n_rechecked']}")
    kms = []
    t0 = time.perf_counter()
    for i in range(args.steps):
        res, t = run(L.FP32_RECHECK, tuning or None)
        kms.append(t["kernel_ms"])
        log(f"step {i}: kernel {t['kernel_ms']:.1f} ms, total {t['total_ms']:.1f} ms")
    wall = time.perf_counter() - t0
    pairs = t["n_pairs"]
    rechecked = t["n_rechecked"]
    if ref is not None:
        for a, b in zip(ref, res):
            assert a["n_pairs"] == b["n_pairs"], "steps differ"
            np.testing.assert_allclose(a["full"], b["full"], rtol=1e-12, atol=0, err_msg="steps differ")
    ms_step = 1000.0 * wall / args.steps
    value = pairs / (wall / args.steps)
    kernel_s = float(np.mean(kms)) / 1000.0

    plain = None
    if args.plain_steps > 0:
        pk = []
        for i in range(args.plain_steps):
            pres, pt = run(L.PLAIN_FP64, None)
            pk.append(pt["kernel_ms"])
            log(f"plain fp64 {i}: kernel {pt['kernel_ms']:.1f} ms, total {pt['total_ms']:.1f} ms")
        # counts must be identical; weighted sums (added in another order)
        # are compared by their largest relative difference
        same = all(a["n_pairs"] == b["n_pairs"] for a, b in zip(pres, res))
        rel = 0.0
        for a, b in zip(pres, res):
            for k in ("full", "jk"):
                fin = np.isfinite(a[k]) & np.isfinite(b[k])
                same = same and bool(np.array_equal(np.isfinite(a[k]), np.isfinite(b[k])))
                d = np.abs(a[k] - b[k])
                m = fin & (np.abs(a[k]) > 0)
                if m.any():
                    rel = max(rel, float(np.max(d[m] / np.abs(a[k][m]))))
                same = same and bool(np.all(d[fin & ~m] == 0))
        plain = dict(kernel_ms=float(np.mean(pk)), pairs_per_s=pairs / (float(np.mean(pk)) / 1000.0),
                     counts_identical=bool(same), max_rel_diff_sums=rel)

    flop = sum(r["n_pairs"] * FLOP_BY_KIND.get(j["kind"], 0) for r, j in zip(res, W["all_jobs"]))
    achieved = flop / kernel_s / 1e12
    tu_used = tuned["choice"] if tuned else dict(cell_subdiv=tuning.get("cell_subdiv", 2),
                                                   max_slots=tuning.get("max_slots", 8), i_chunk=256)
    # region slots in use: the LDS holds at most SLOTS_FIT per item on the
    # default workload, so larger settings run the same plan
    import hashlib
    src_sha = hashlib.sha256(open(KERNEL_SRC, "rb").read()).hexdigest()[:16] if os.path.exists(KERNEL_SRC) else None
    traffic_key = ((src_sha,) + W["traffic_key"] + (tu_used["cell_subdiv"], min(tu_used["max_slots"], SLOTS_FIT),
                                                   tu_used["i_chunk"])
                   if W["traffic_key"] else None)
    traffic = TRAFFIC_BYTES.get(traffic_key)
    config = dict(W["config"])
    config.update({"parallelism": "single GPU", "tuning": tuning or (tuned if tuned else "default")})
    line = {
        "metric": "pairs/s s-mu DD+DR+RR: FP32+edge-recheck batched vs FP64 kernel, exact counts",
        "value": value, "unit": "pairs/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64-recheck",
        "data": "synthetic (seeded Neyman-Scott + uniform randoms)",
        "config": config,
        "pairs_per_step": pairs,
        "kernel_ms": float(np.mean(kms)),
        "kernel_pairs_per_s": pairs / kernel_s,
        "rechecked_pairs": rechecked,
        "upload_ms": upload_ms,
        "first_call_ms": first_call_ms,
        "pcie_inclusive_pairs_per_s": (pairs / ((upload_ms + first_call_ms) / 1000.0)) if first_call_ms else None,
        "plain_fp64": plain,
        "speedup_vs_plain_fp64": (plain["kernel_ms"] / float(np.mean(kms))) if plain else None,
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_TFLOPS,
                     "traffic": traffic,
                     "traffic_source": (f"rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE of this kernel source (sha256 "
                                        f"{src_sha}), profiles/k_count_fp32_jackknife_batch_hbm.md" if traffic else
                                        f"not measured for this kernel source (sha256 {src_sha}) and workload"),
                     "traffic_bytes_per_pair": traffic / pairs if traffic else None,
                     "algorithmic_bytes_per_pair": (t["n_staged"] * W["obj_bytes"] / pairs
                                                    if W.get("obj_bytes") and t.get("n_staged") else None),
                     "algorithmic_bytes": t["n_staged"] * W["obj_bytes"] if W.get("obj_bytes") and t.get("n_staged")
                                          else None,
                     "note": "FP32 compute bound (pair screening on the matrix cores, binning on the vector "
                             "ALUs; both FP32 peaks are 157.3 TFLOP/s); algorithmic flop per accepted pair: "
                             "s-mu 26, sigma-pi 29, monopole and angular 15"},
    }
    if not args.no_cpu and W["cpu"] is not None:
        try:
            line["cpu_baseline"] = W["cpu"]()
        except Exception as e:  # reported, not fatal
            line["cpu_baseline"] = {"error": str(e)}
    else:
        line["cpu_baseline"] = None
    ctx.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
