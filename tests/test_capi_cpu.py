"""C-ABI library checks that need no GPU: it loads, it exports every entry
point declared in include/*.h, and argument errors come back as status codes
without touching a device."""
import ctypes as C
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    inc = os.path.join(ROOT, "include")
    for f in os.listdir(inc):
        if f.endswith(".h"):
            txt = open(os.path.join(inc, f)).read()
            names |= set(re.findall(r"FCT_API\s+[\w\s\*]+?\b(fct_\w+)\s*\(", txt))
    return names


def test_library_exports_every_declared_symbol():
    from freect_amd import wfbp
    lib = wfbp.load_library()
    declared = _declared()
    assert len(declared) >= 15
    out = subprocess.run(["nm", "-D", "--defined-only", wfbp.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r"\bT (fct_\w+)", out))
    assert declared <= exported, declared - exported
    # ctypes signatures cover every declared entry point, and nothing else.
    assert set(wfbp.SIGNATURES) == declared
    for name in declared:
        getattr(lib, name)


def test_status_codes_without_device():
    from freect_amd import wfbp
    lib = wfbp.load_library()
    # Null arguments are rejected before any device call.
    assert lib.fct_wfbp_gpu_plan_create(None, None, None, None, None, None) == 1
    assert lib.fct_wfbp_gpu_rebin_filter(None, None, None, None) == 1
    assert lib.fct_wfbp_gpu_backproject(None, None, None, None) == 1
    assert lib.fct_wfbp_gpu_plan_info(None, None) == 1
    assert lib.fct_wfbp_gpu_last_error(None) == b"null context"


def test_create_fails_cleanly_without_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from freect_amd import wfbp
    h = C.c_void_p()
    st = wfbp.load_library().fct_wfbp_gpu_create(0, C.byref(h))
    assert st in (1, 4) and not h.value


def test_buffer_arguments_are_checked_before_the_abi():
    # Every tensor handed to a plan stage is checked against the plan's sizes
    # before its address crosses the C ABI: element type, contiguity, element
    # count (or byte count for the opaque filtered buffer) and device.
    import torch
    from freect_amd import wfbp
    chk = wfbp._dev_arg
    good = torch.zeros(12, dtype=torch.float32)
    with pytest.raises(wfbp.FctError, match="dtype must be float32"):
        chk(good.double(), "raw", 0, numel=12, dtype="f32")
    with pytest.raises(wfbp.FctError, match="dtype must be int32"):
        chk(good, "per_voxel", 0, numel=12, dtype="i32")
    with pytest.raises(wfbp.FctError, match="contiguous"):
        chk(torch.zeros(4, 6)[:, :3], "raw", 0, numel=12, dtype="f32")
    with pytest.raises(wfbp.FctError, match="11 elements, the plan needs 12"):
        chk(good[:11], "raw", 0, numel=12, dtype="f32")
    with pytest.raises(wfbp.FctError, match="48 bytes, the plan needs 40"):
        chk(good, "filtered", 0, nbytes=40)
    with pytest.raises(wfbp.FctError, match="must live on cuda:0"):
        chk(good, "raw", 0, numel=12, dtype="f32")
    with pytest.raises(wfbp.FctError, match="a device buffer is required"):
        chk(None, "native", 0, numel=12, dtype="f32")
    with pytest.raises(wfbp.FctError, match="expected a torch tensor"):
        chk(good.numpy(), "raw", 0, numel=12, dtype="f32")
    assert chk(None, "scale", 0, optional=True) is None
    assert chk(1234, "raw", 0, numel=12, dtype="f32") == 1234   # raw address: the C contract
