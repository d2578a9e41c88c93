"""CPU-side checks of the C ABI (include/pmx.h): the HIP library loads, exports
every declared entry point, and reports argument and device errors with
status codes instead of exiting.  No kernel is launched."""
import ctypes
import os
import re

import pytest

from patchmatch_cuda_amd import capi, configs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pmx_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    names = _declared("pmx.h")
    assert len(names) >= 35
    L = ctypes.CDLL(capi.LIB_PATH)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # and the Python front end binds exactly what the header declares
    assert sorted(capi.SIGNATURES) == names


def test_validate_params_messages():
    L = capi.lib()
    _, rp = configs.C1_SMALL
    p = capi.Params.from_run_params(rp)
    assert L.pmx_validate_params(ctypes.byref(p), 4) == capi.PMX_OK
    assert L.pmx_validate_params(None, 4) == 1
    assert L.pmx_validate_params(ctypes.byref(p), 0) == 1
    assert b"at least 2 images" in L.pmx_last_error(None)
    for bad in [dict(half_window=0), dict(half_window=23)]:
        q = capi.Params.from_run_params(rp.replace(**bad))
        assert L.pmx_validate_params(ctypes.byref(q), 4) == 1
        assert b"too large size of windows" in L.pmx_last_error(None)
    for bad in [dict(depth_near=0.0), dict(depth_far=1.0), dict(num_samples=0),
                dict(spm_alpha=0.0), dict(blur_mode=7), dict(init_mode=3), dict(block_x=0)]:
        q = capi.Params.from_run_params(rp.replace(**bad))
        assert L.pmx_validate_params(ctypes.byref(q), 4) == 1, bad


def test_null_arguments_are_errors():
    L = capi.lib()
    assert L.pmx_device_count(None) == 1
    assert L.pmx_ctx_create(0, None) == 1
    assert L.pmx_engine_destroy(None) in (0, 1)


@pytest.mark.skipif(capi.device_count() > 0, reason="a HIP device is present")
def test_no_device_reports_status():
    assert capi.device_count() == 0
    with pytest.raises(capi.PmxError) as ei:
        capi.Context(0)
    assert ei.value.code == 5
    assert "no HIP device" in str(ei.value)
