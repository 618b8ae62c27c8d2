"""CPU checks of the C-ABI library: it loads, exports every symbol that
include/*.h declares, its struct layouts match the binding, its defaults
are the reference's parameters.dat, and parameter validation maps onto the
documented status codes.  No device compute is issued here."""
import ctypes as C
import json
import math
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = json.load(open(os.path.join(ROOT, "tests", "golden", "reference_vectors.json")))


def _declared_symbols():
    inc = os.path.join(ROOT, "include")
    txt = "".join(open(os.path.join(inc, f)).read() for f in sorted(os.listdir(inc)) if f.endswith(".h"))
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(smc_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    from supermc_amd import _native
    lib = _native.load()
    syms = _declared_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(lib, s), s
    assert set(syms) == set(_native.SIGNATURES), "binding and header disagree"


def test_binding_constants_match_header():
    """Array sizes the binding passes to the library follow the header."""
    from supermc_amd import _native
    txt = open(os.path.join(ROOT, "include", "smc.h")).read()
    want = {"SMC_TILE_VARIANTS": _native.TILE_VARIANTS, "SMC_EV_STRIDE": _native.EV_STRIDE,
            "SMC_MAX_A": _native.MAX_A}
    for name, val in want.items():
        m = re.search(r"#define\s+%s\s+(\d+)" % name, txt)
        assert m and int(m.group(1)) == val, name


def test_struct_layouts():
    from supermc_amd import _native
    lib = _native.load()
    assert lib.smc_params_sizeof() == C.sizeof(_native.SmcParams)
    assert lib.smc_derived_sizeof() == C.sizeof(_native.SmcDerived)


def test_defaults_are_reference_parameters_dat():
    from supermc_amd import default_params
    p = default_params()
    for k, v in GOLD["parameters_dat"].items():
        if k.startswith("_"):
            continue
        attr = "lambda_" if k == "lambda" else k
        assert getattr(p, attr) == pytest.approx(v), k
    assert math.isnan(p.proj_beta2) and p.sigma_nn_mb < 0


@pytest.mark.parametrize("override,code", [
    (dict(which_mc_model=3), 7),
    (dict(which_mc_model=5, sub_model=3, collision_criterion=1), 1),
    (dict(which_mc_model=5, sub_model=1, collision_criterion=3), 7),
    (dict(which_mc_model=5, sub_model=1, collision_criterion=4), 7),
    (dict(which_mc_model=5, sub_model=1, collision_criterion=1, Aproj=16), 7),
    (dict(which_mc_model=5, sub_model=1, collision_criterion=1, Aproj=300), 1),
    (dict(which_mc_model=5, sub_model=1, collision_criterion=1, ecc_from_order=4, ecc_to_order=2), 1),
    (dict(which_mc_model=5, sub_model=1, collision_criterion=1, cc_fluctuation_model=1), 7),
    (dict(which_mc_model=5, sub_model=1, collision_criterion=1, shape_of_nucleons=4, gauss_nucl_width=0.08,
          ecm=5020.0), 7),
])
def test_validation_status_codes(override, code):
    from supermc_amd import _native, default_params
    lib = _native.load()
    p = default_params(**override)
    h = C.c_void_p()
    rc = lib.smc_create(C.byref(h), C.byref(p))
    assert rc == code
    assert not h.value


@pytest.mark.parametrize("override", [dict(random_flag=4), dict(random_flag=-1), dict(grid_nx=0),
                                      dict(grid_deta=0.0), dict(ecm=1.5)])
def test_p3d_parameter_validation(override):
    # smc_p3d_create validates before touching the device
    from supermc_amd import _native
    lib = _native.load()
    args = dict(grid_nx=261, grid_ny=261, grid_neta=101, grid_dx=0.1, grid_dy=0.1, grid_deta=0.1, ecm=19.6,
                random_flag=1)
    args.update(override)
    p = _native.P3dParams(**args)
    h = C.c_void_p()
    assert lib.smc_p3d_create(C.byref(h), C.byref(p)) == 1
    assert not h.value
