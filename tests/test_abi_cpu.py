"""The C-ABI libraries load and export every entry point their headers
declare; argument errors come back as status codes. CPU only: no call here
reaches a HIP device (creation without a device must fail cleanly)."""
import ctypes as C
import os
import re

import pytest

from conftest import gpu_available
from meso_hip import capi
from meso_hip import meshgen as mg

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = re.sub(r"/\*.*?\*/", "", open(header).read(), flags=re.S)
    return sorted(set(re.findall(r"^[a-z_0-9 ]*?\**\s*\b((?:meso|oracle)_\w+)\s*\(", text, flags=re.M)))


def test_product_library_exports_the_header():
    names = declared(os.path.join(ROOT, "include", "meso_hip.h"))
    assert len(names) >= 27
    L = C.CDLL(capi.LIB_PATH)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_oracle_library_exports_its_header():
    from oracle import oracle

    names = declared(os.path.join(ROOT, "oracle", "meso_oracle.h"))
    assert "oracle_step" in names
    L = C.CDLL(oracle.LIB_PATH)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_product_library_does_not_link_the_oracle():
    import subprocess

    out = subprocess.run(["ldd", capi.LIB_PATH], capture_output=True, text=True).stdout
    assert "oracle" not in out
    syms = subprocess.run(["nm", "-D", "--defined-only", capi.LIB_PATH], capture_output=True, text=True).stdout
    assert "oracle_" not in syms


def test_null_arguments_are_rejected():
    L = capi.lib()
    rep = capi.StepReport()
    assert L.meso_hip_step(None, 1, C.byref(rep)) == capi.MESO_E_INVALID
    assert L.meso_hip_sync(None) == capi.MESO_E_INVALID
    assert L.meso_hip_get_macros(None, None, None, None, None) == capi.MESO_E_INVALID
    assert L.meso_hip_residual(None, None) == capi.MESO_E_INVALID
    assert L.meso_hip_set_params(None, None) == capi.MESO_E_INVALID
    assert L.meso_hip_device_bytes(None) == 0
    assert L.meso_hip_autotune(None, 1, None) == capi.MESO_E_INVALID
    assert L.meso_hip_set_graphs(None, 1) == capi.MESO_E_INVALID
    assert L.meso_hip_tiles(None, None) == capi.MESO_E_INVALID
    assert L.meso_hip_step_batch(None, 1, 1, None) == capi.MESO_E_INVALID
    arr = (C.c_void_p * 2)(None, None)
    assert L.meso_hip_step_batch(arr, 2, 1, None) == capi.MESO_E_INVALID  # a null context
    assert L.meso_hip_step_batch(arr, -1, 1, None) == capi.MESO_E_INVALID
    L.meso_hip_destroy(None)  # no-op
    h = C.c_void_p()
    assert L.meso_case_open(None, 0.0, C.byref(h)) == capi.MESO_E_INVALID
    assert L.meso_hip_create(None, None, None, 0, C.byref(h)) == capi.MESO_E_INVALID


@pytest.mark.skipif(gpu_available(), reason="checks the no-device error path")
def test_create_without_a_device_fails_cleanly(tmp_path):
    p = mg.shakhov_cavity_case(str(tmp_path), mg.quad_cavity(3), ("nc", 1, 4, 4.0), 0.1, 0.1)
    cs = capi.Case(p)
    with pytest.raises(capi.MesoError) as e:
        capi.Device(cs)
    assert e.value.code == capi.MESO_E_DEVICE


def test_create_rejects_bad_storage(tmp_path):
    p = mg.shakhov_cavity_case(str(tmp_path), mg.quad_cavity(3), ("nc", 1, 4, 4.0), 0.1, 0.1)
    cs = capi.Case(p)
    with pytest.raises(capi.MesoError) as e:
        capi.Device(cs, storage=7)
    assert e.value.code == capi.MESO_E_INVALID
