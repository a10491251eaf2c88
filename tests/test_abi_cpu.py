"""CPU checks of libtwopcf_hip: it loads, exports every symbol the C header
declares, and its host-side arithmetic (threshold tables, cosmology,
Galaxy::set) agrees with the oracle and with the reference's golden dumps.
No kernel is launched here."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from cases import load_case
from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "twopcf_hip.h")


@pytest.fixture(scope="module")
def hip():
    from two_pcf_amd import lib as L
    return L


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[a-z0-9_]+\s*\*?\s*(twopcf_[a-z0-9_]+)\s*\(", text, re.M)))


def test_exports_every_declared_symbol(hip):
    names = declared_functions()
    assert len(names) >= 10
    L = hip.lib()
    for n in names:
        assert hasattr(L, n), n
    assert sorted(hip.EXPORTS) == names
    assert L.twopcf_abi_version() == hip.ABI_VERSION == 5


def test_null_arguments_are_rejected(hip):
    L = hip.lib()
    assert L.twopcf_ctx_create(0, None) == -1
    assert L.twopcf_count(None, None, 0, 0, None, None) == -1
    assert L.twopcf_count_noalloc(None, None, 0, 0, None, None) == -1
    assert L.twopcf_catalog_free(None) == -1
    assert L.twopcf_s2_thresholds(0, None, None, None) == -1


S_AXES = [
    (2, (0.0, 100.0, 40, 1.3), None),          # s-mu s axis (configs 1-4)
    (0, (0.0, 100.0, 30, 1.3), None),          # monopole (configs 0, 3)
    (2, (0.5, 60.0, 25, 1.5), None),           # log axis with min > 0
    (0, (0.0, 50.0, 20, 1.0), None),           # linear s axis
]


@pytest.mark.parametrize("kind,ax,_", S_AXES)
def test_s2_thresholds_exhaustive(hip, oracle_lib, kind, ax, _):
    """Every float s^2 in [0, 1.5 * max^2] bins identically under the table
    and under the reference's sqrtf + log1pf pipeline."""
    o = oracle_lib
    t = hip.s2_thresholds(kind, ax, (0.0, 1.0, 50, 0.0))
    oa = o.axis(*ax)
    ob = o.axis(0.0, 1.0, 50, 0.0)
    hi = np.float32(1.5 * ax[1] ** 2)
    bad = o.lib().or_scan_s2_table(kind, C.byref(oa), C.byref(ob), t.ctypes.data, 0.0, float(hi), 0)
    assert bad == 0
    assert np.all(np.diff(t) >= 0)


@pytest.mark.parametrize("ax", [(0.5, 50.0, 20, 1.5), (1.0, 40.0, 15, 1.4), (0.0, 50.0, 50, 1.0),
                                (0.0, 1.0 - np.cos(0.175), 40, 1.6)])
def test_value_thresholds_exhaustive(hip, oracle_lib, ax):
    o = oracle_lib
    t = hip.axis_thresholds(ax)
    oa = o.axis(*ax)
    bad = o.lib().or_scan_value_table(C.byref(oa), t.ctypes.data, 0.0, float(np.float32(2 * ax[1])), 0)
    assert bad == 0


def test_geometry_matches_reference(hip):
    case = load_case("eq_clustered")
    for tag in ("data", "randoms"):
        a = case.inputs[tag]
        d = hip.comoving_distance(0.25, 1.0, 0.0, 1.0, a["z"])
        np.testing.assert_array_equal(d, case.g[f"cat_{tag}_d"])
        out = hip.eq_to_cart(a["ra"], a["dec"], d)
        for fld, v in zip(("x", "y", "z", "nx", "ny", "nz"), out):
            np.testing.assert_array_equal(v, case.g[f"cat_{tag}_{fld}"], err_msg=fld)


def test_no_gpu_context_fails_cleanly(hip):
    """Without a device the context call returns an error code, never crashes."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    h = C.c_void_p()
    assert hip.lib().twopcf_ctx_create(0, C.byref(h)) == -3
