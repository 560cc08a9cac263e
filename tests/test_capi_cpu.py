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
