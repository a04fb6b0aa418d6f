"""The C-ABI library loads and exports every entry point include/*.h
declares (no device calls)."""
import os
import re
import subprocess

import gpu_cdt_refine_amd as g

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "gpu_cdt_refine.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return set(re.findall(r"\b(gcdt_[a-z0-9_]+)\s*\(", txt))


def test_library_exports_declared_symbols():
    out = subprocess.run(["nm", "-D", "--defined-only", g.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r"\s[TW]\s+(\S+)", out))
    missing = declared_symbols() - exported
    assert not missing, missing
    # the drop-in keeps the reference's C++ linkage and signature
    mangled = "_Z18GPU_Refine_QualityP13triangulateioS0_d12InsertPolicy12DeletePolicyiiPP7PStatusPP7TStatus"
    assert mangled in exported


def test_library_loads_without_device():
    L = g.lib()
    o = g.make_options(theta=30.0)
    assert o.theta_deg == 30.0 and o.cand_factor == 4 and o.claim_iters == 1
    assert L.gcdt_last_error() == b""


def test_options_check_refuses_claim_epoch_overflow():
    """One graph launch may use at most 2047 claim epochs (11 bits of the
    claim key); the options check refuses settings beyond that before any
    device work."""
    import ctypes as C
    L = g.lib()
    out = C.c_void_p()
    bad = g.make_options(graph=1, graph_rounds=64, graph_seg_passes=16, claim_iters=2)
    rc = L.gcdt_refine_host(None, 0, None, None, 0, None, 0, C.byref(bad), C.byref(out), None)
    assert rc == -1 and b"claim epochs" in L.gcdt_last_error()
    # 64 * (15 + 1) * (1 + 1) + 2 = 2050 epochs: just over the limit
    edge = g.make_options(graph=1, graph_rounds=64, graph_seg_passes=15, claim_iters=1)
    rc = L.gcdt_refine_host(None, 0, None, None, 0, None, 0, C.byref(edge), C.byref(out), None)
    assert rc == -1 and b"claim epochs" in L.gcdt_last_error()
