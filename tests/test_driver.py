"""The parameter-file driver (two_pcf_amd/twopcf, main.cpp of the reference
restated on the GPU library) against the reference binary's own output files:
same parameter file, same ascii catalogues, byte-identical output."""
import os
import shutil
import subprocess

import pytest

from conftest import GOLDEN, ROOT

DRIVER = os.path.join(ROOT, "two_pcf_amd", "twopcf")
CASES = sorted(os.listdir(os.path.join(GOLDEN, "driver")))


def run_driver(case, tmp_path, method):
    src = os.path.join(GOLDEN, "driver", case)
    for fn in ("params.txt", "data.txt", "rand.txt"):
        shutil.copy(os.path.join(src, fn), tmp_path)
    env = dict(os.environ, TWOPCF_METHOD=method)
    return subprocess.run([DRIVER, "params.txt"], cwd=tmp_path, env=env, capture_output=True, text=True, timeout=300)


@pytest.mark.gpu
@pytest.mark.parametrize("method", ["fp64", "fp32"])
@pytest.mark.parametrize("case", CASES)
def test_driver_outputs_identical(case, method, tmp_path):
    print(f"[gpu] driver {case} {method}", flush=True)
    r = run_driver(case, tmp_path, method)
    assert r.returncode == 0, r.stderr
    exp = os.path.join(GOLDEN, "driver", case, "expected")
    for fn in sorted(os.listdir(exp)):
        with open(os.path.join(exp, fn)) as f:
            want = f.read()
        with open(os.path.join(tmp_path, fn)) as f:
            got = f.read()
        assert got == want, f"{case}/{fn} differs"


def test_driver_fails_loudly_without_gpu(tmp_path):
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    r = run_driver(CASES[0], tmp_path, "fp32")
    assert r.returncode != 0
    assert "no usable GPU" in r.stderr
