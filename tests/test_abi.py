"""The C ABI library without a GPU: it loads, it exports every entry point
declared in include/fest_hip.h, and it rejects bad configurations with
FEST_E_CONFIG before touching a device (the reference's Fatal_error on an
unknown scheme or boundary, src/error.h:17)."""
import ctypes
import dataclasses
import os
import re

import pytest

from festgpu import cases as C
from festgpu import native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "fest_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fest_[a-z_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = native.lib()
    names = declared_functions()
    assert len(names) >= 12
    for n in names:
        assert hasattr(lib, n), n


def test_struct_sizes_match_header():
    # fest_extent is four int32; fest_scheme 12 int32, one double, then the
    # transition, pbs(3) and muscl_limiter int32 (padded to 8)
    assert ctypes.sizeof(native.Extent) == 16
    assert ctypes.sizeof(native.Scheme) == 80
    assert ctypes.sizeof(native.BC) == 24 + 8 * 72
    # 16 doubles, mu_sutherland (int32, then 4 bytes of padding), tv_inf, tu_inf, tkl_inf, tgm_inf
    assert ctypes.sizeof(native.Flow) == 8 * 16 + 8 + 8 + 8 + 8 + 8


@pytest.mark.parametrize("change,code", [
    (dict(flux="roe"), None),
    (dict(bc_id=(-12, -9, -5, -5, -9, -9)), -3),
    (dict(bc_id=(-9, -9, -5, -5, -7, -9), kmx=2), -3),  # a k pole on a one-cell k extent
    (dict(time_accuracy="RK2", bc_id=(-9, -9, -5, -5, -9, -9), imx=3), -3),
    (dict(kmx=3), -3),  # two cells in k: neither 3-D nor the reference's 2-D
    (dict(turb="kkl", transition="bc", bc_id=(-9, -9, -5, -5, -9, -9)), -3),
    # lctm2015 exists for SST and SST-2003 only, with RK
    (dict(turb="sa", transition="lctm2015", mu_ref=1e-5), -3),
    (dict(turb="kkl", transition="lctm2015", mu_ref=1e-5), -3),
    (dict(turb="sst", transition="lctm2015", mu_ref=1e-5, time_accuracy="implicit"), -3),
])
def test_config_errors(change, code):
    case, _, _ = C.laminar_box(8, 6, 5)
    if "flux" in change:
        with pytest.raises(KeyError):
            native.c_structs(dataclasses.replace(case, **change))
        return
    with pytest.raises(native.FestError) as e:
        native.GpuBlock(dataclasses.replace(case, **change))
    assert e.value.code == code


def test_precision_modes():
    case, _, _ = C.laminar_box(8, 6, 5)
    with pytest.raises(native.FestError) as e:
        native.GpuBlock(case, precision=7)
    assert e.value.code == -3
    # lctm2015 runs in FP64 only
    lc = dataclasses.replace(case, turb="sst", transition="lctm2015", mu_ref=1e-5)
    with pytest.raises(native.FestError) as e:
        native.GpuBlock(lc, precision=1)
    assert e.value.code == -3


def test_lctm2015_passes_the_configuration_check():
    """SST + lctm2015 with n_var 8 is accepted; without a GPU the handle
    fails on the device, after the check."""
    case, _, _ = C.sst_bump(8, 6, 5, transition="lctm2015")
    assert case.n_var == 8 and case.n_grad == 7
    try:
        native.GpuBlock(case).close()
    except native.FestError as e:
        assert e.code == -4, e


@pytest.mark.parametrize("ids", [(-7, -9, -5, -5, -9, -9), (-11, -4, -5, -5, -9, -9)])
def test_pole_and_total_pressure_pass_the_configuration_check(ids):
    """Without a GPU the handle fails on the device, after the check."""
    case, _, _ = C.laminar_box(8, 6, 5)
    try:
        native.GpuBlock(dataclasses.replace(case, bc_id=ids)).close()
    except native.FestError as e:
        assert e.code == -4, e


def test_two_dimensional_extent_passes_the_configuration_check():
    """One cell in k (kmx == 2) is the reference's 2-D mode; without a GPU
    the handle fails on the device, after the configuration check."""
    case, _, _ = C.laminar_box(8, 6, 1)
    try:
        native.GpuBlock(case).close()
    except native.FestError as e:
        assert e.code == -4, e
