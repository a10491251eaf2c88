"""The host layer's ascii reader (io.cpp:240-415 restated, parallel parse):
separators, empty fields, comments, blank lines, CRLF, weights (strtof),
regions, bit-mask columns and the IIP weights of the inverse-p build, against
an independent Python parse and libc's strtof.  CPU only."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT

LIBC = C.CDLL(None)
LIBC.strtof.restype = C.c_float
LIBC.strtof.argtypes = [C.c_char_p, C.POINTER(C.c_char_p)]


def strtof(s):
    return float(LIBC.strtof(s.encode(), None))


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    lib_dir = os.path.join(ROOT, "two_pcf_amd")
    if not os.path.exists(os.path.join(lib_dir, "libtwopcf_hip.so")):
        pytest.skip("libtwopcf_hip.so is not built")
    out = str(tmp_path_factory.mktemp("reader") / "host_reader")
    subprocess.run(["g++", "-O2", "-std=c++17", "-pthread", "-ffp-contract=off",
                    os.path.join(ROOT, "tests", "host_reader_main.cpp"),
                    os.path.join(ROOT, "two_pcf_amd", "csrc", "host", "twopcf_ref_api.cpp"),
                    os.path.join(ROOT, "two_pcf_amd", "csrc", "host", "h5lib.cpp"),
                    "-L", lib_dir, "-ltwopcf_hip", f"-Wl,-rpath,{lib_dir}", "-ldl", "-o", out], check=True)
    return out


def run(harness, tmp_path, params, text, invp=0):
    (tmp_path / "cat.txt").write_bytes(text.encode())
    params = dict(params, data_filename=str(tmp_path / "cat.txt"), data_file_type="ascii")
    (tmp_path / "p.txt").write_text("".join(f"{k} = {v}\n" for k, v in params.items()))
    r = subprocess.run([harness, str(tmp_path / "p.txt"), str(invp)], capture_output=True, text=True)
    return r


def parse(out):
    """The harness output after the reader's own messages: 'n <count> <words>'
    then one line per object."""
    lines = out.splitlines()
    k = max(i for i, l in enumerate(lines) if l.startswith("n "))
    head = lines[k].split()
    return int(head[1]), int(head[2]), [l.split() for l in lines[k + 1:]]


CART = dict(coord_system="cartesian", use_weights=1, n_jk_regions=3, omega_m=0.25, h=1.0, z_min=0.0, z_max=1.0)


def test_separators_comments_and_fields(harness, tmp_path):
    text = ("# x y z w jk\n"
            "1.5 2.25 -3.125 0.1 2\n"
            "\n"
            "4\t5 \t 6.5\t0.7\t1\n"
            "7, 8 ,9.75,1.3333333,0\n"
            "10  11   12 0.2   1\r\n"
            "#comment line\n"
            "1e2 -2E-3 3.0000000000000001 1.1 2 extra\n")
    r = run(harness, tmp_path, CART, text)
    assert r.returncode == 0, r.stderr
    n, nw, rows = parse(r.stdout)
    assert n == 5 and nw == 0
    want = [("1.5", "2.25", "-3.125", "0.1", 2), ("4", "5", "6.5", "0.7", 1), ("7", "8", "9.75", "1.3333333", 0),
            ("10", "11", "12", "0.2", 1), ("1e2", "-2E-3", "3.0000000000000001", "1.1", 2)]
    for row, (x, y, z, w, jk) in zip(rows, want):
        got = [float.fromhex(v) for v in row[:8]]
        assert got[:3] == [float(x), float(y), float(z)]
        d = np.sqrt(float(x) ** 2 + float(y) ** 2 + float(z) ** 2)
        assert got[3:6] == [float(x) / d, float(y) / d, float(z) / d]
        assert got[6] == strtof(w) and got[7] == strtof(w)
        assert int(row[8]) == jk
        assert float.fromhex(row[9]) == d


def test_equatorial_positions_match_library(harness, tmp_path):
    from two_pcf_amd import lib as L
    rng = np.random.Generator(np.random.PCG64(4))
    ra, dec, z = rng.uniform(150, 250, 300), rng.uniform(0, 60, 300), rng.uniform(0.2, 0.6, 300)
    z[:3] = [0.0, 1.0, -0.1]   # table ends and a negative redshift
    text = "".join("%.17g %.17g %.17g\n" % v for v in zip(ra, dec, z))
    p = dict(CART, coord_system="equatorial", use_weights=0, n_jk_regions=0)
    r = run(harness, tmp_path, p, text)
    assert r.returncode == 0, r.stderr
    n, _, rows = parse(r.stdout)
    assert n == 300
    d = L.comoving_distance(0.25, 1.0, 0.0, 1.0, z)
    x, y, zz, nx, ny, nz = L.eq_to_cart(ra, dec, d)
    got = np.array([[float.fromhex(v) for v in row[:6]] + [float.fromhex(row[9])] for row in rows])
    np.testing.assert_array_equal(got, np.stack([x, y, zz, nx, ny, nz, d], 1))


def test_bit_masks_and_iip_weights(harness, tmp_path):
    """Inverse-p build: 2 mask words (n_bitwise_runs 100), signed or unsigned
    decimal; weight *= n_runs / popcount unless the popcount is n_runs
    (0 counts as 1); input_weight keeps the read weight (io.cpp:521-588)."""
    masks = [(3, 0), (2 ** 64 - 1, 2 ** 36 - 1), (-1, 0), (0, 0)]
    lines = []
    for k, (a, b) in enumerate(masks):
        a_txt = str(a)
        lines.append(f"{k + 1}.0 2.0 3.0 0.75 {a_txt} {b}\n")
    p = dict(CART, n_jk_regions=0, n_bitwise_runs=100, bitwise_weight_dataset_name="BW")
    r = run(harness, tmp_path, p, "".join(lines), invp=1)
    assert r.returncode == 0, r.stderr
    n, nw, rows = parse(r.stdout)
    assert n == 4 and nw == 2
    for row, (a, b) in zip(rows, masks):
        ua, ub = a % 2 ** 64, b % 2 ** 64
        assert (int(row[10]), int(row[11])) == (ua, ub)
        tot = bin(ua).count("1") + bin(ub).count("1")
        tot = 1 if tot == 0 else tot
        w0 = strtof("0.75")
        want = w0 if tot == 100 else float(np.float32(w0 * (100 / tot)))
        assert float.fromhex(row[6]) == want
        assert float.fromhex(row[7]) == w0


def test_dither_masks_and_iip_weights(harness, tmp_path):
    """Dither weights: n_words dither-mask columns follow the bit masks; the
    read weight is scaled by dither_total / total unless total is n_runs
    (io.cpp:549-574); input_weight keeps the read weight."""
    rows_in = [((3, 0), (1, 1)), ((2 ** 64 - 1, 2 ** 36 - 1), (7, 0)), ((0, 0), (0, 0)), ((5, 6), (0, 2 ** 63))]
    text = "".join(f"{k + 1}.0 2.0 3.0 0.75 {m[0]} {m[1]} {d[0]} {d[1]}\n" for k, (m, d) in enumerate(rows_in))
    p = dict(CART, n_jk_regions=0, n_bitwise_runs=100, bitwise_weight_dataset_name="BW",
             dither_weight_dataset_name="DW")
    r = run(harness, tmp_path, p, text, invp=2)
    assert r.returncode == 0, r.stderr
    n, nw, rows = parse(r.stdout)
    assert n == 4 and nw == 2
    w0 = strtof("0.75")
    for row, (m, d) in zip(rows, rows_in):
        assert [int(v) for v in row[10:14]] == [*m, *d]
        tot = sum(bin(v).count("1") for v in m) or 1
        dtot = sum(bin(v).count("1") for v in d)
        want = w0 if tot == 100 else float(np.float32(w0 * (dtot / tot)))
        assert float.fromhex(row[6]) == want
        assert float.fromhex(row[7]) == w0
    # the dither columns are required
    r = run(harness, tmp_path, p, "1.0 2.0 3.0 0.75 3 0\n", invp=2)
    assert r.returncode != 0 and "too few columns" in r.stderr


def test_errors_are_reported(harness, tmp_path):
    r = run(harness, tmp_path, CART, "1 2 3 0.5\n")
    assert r.returncode != 0 and "too few columns" in r.stderr
    r = run(harness, tmp_path, CART, "1 2 x 0.5 1\n")
    assert r.returncode != 0 and "not a number" in r.stderr
    r = run(harness, tmp_path, CART, "1 2 3 0.5 7\n")
    assert r.returncode != 0 and "outside [0, 3)" in r.stderr
