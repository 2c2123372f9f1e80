"""CPU-side checks of the C ABI: the library loads and exports every function include/*.h declares.
No compute call is made (there is no GPU here)."""
import ctypes
import glob
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(ruv_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_header_declares_the_boundary():
    names = declared_functions()
    for n in ("ruv_available", "ruv_last_error", "ruv_cg_solve_precond", "ruv_cg_solve", "ruv_apply_ic0",
              "ruv_cg_solve_mixed", "ruv_plan_create", "ruv_plan_solve_dev", "ruv_plan_destroy"):
        assert n in names


def test_library_exports_every_declared_symbol():
    import researchuv_amd._lib as L

    lib = L.lib()
    missing = [n for n in sorted(declared_functions()) if not hasattr(lib, n)]
    assert not missing, f"declared in include/*.h but not exported: {missing}"


def test_library_carries_gfx950_code():
    import researchuv_amd._lib as L

    data = open(L.LIB_PATH, "rb").read()
    assert b"gfx950" in data


def test_status_codes_match_header():
    import researchuv_amd._lib as L

    text = open(os.path.join(ROOT, "include", "ruv_lscm.h")).read()
    for code, name in L.STATUS.items():
        if code == 0:
            continue
        m = re.search(r"RUV_ERR_%s\s*=\s*(-?\d+)" % name, text)
        assert m and int(m.group(1)) == code, name
