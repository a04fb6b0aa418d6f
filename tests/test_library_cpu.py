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
