"""The C-ABI library loads and exports every entry point include/wtb_bse.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "wtb_bse.h"
LIB = ROOT / "wantibexos_amd" / "libwtb_bse.so"


def declared():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(wtb_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared()
    for n in ["wtb_create", "wtb_eigsys", "wtb_zheev", "wtb_bsesolver", "wtb_bse_spectrum",
              "wtb_dielbsev", "wtb_dielbsep", "wtb_monkhorst_pack"]:
        assert n in names


@pytest.mark.skipif(not LIB.exists(), reason="library not built (python -m wantibexos_amd.build)")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(LIB))
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.skipif(not LIB.exists(), reason="library not built")
def test_no_device_fails_loudly():
    """Without a usable gfx950 device the library refuses to create a context."""
    from wantibexos_amd import bse
    try:
        c = bse.Context(0)
    except bse.WTBError:
        return
    c.close()
    pytest.skip("a GPU is present")


def c_offsets(tmp_path, struct, fields):
    src = tmp_path / "off.c"
    body = "".join(f'  printf("%s %zu\\n", "{f}", offsetof({struct}, {f}));\n' for f in fields)
    src.write_text(f'#include <stdio.h>\n#include <stddef.h>\n#include "wtb_bse.h"\nint main(void) {{\n'
                   f'  printf("sizeof %zu\\n", sizeof({struct}));\n{body}  return 0;\n}}\n')
    exe = tmp_path / "off"
    import subprocess
    subprocess.run(["gcc", "-I", str(ROOT / "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout
    return {k: int(v) for k, v in (l.split() for l in out.splitlines())}


@pytest.mark.parametrize("struct,pyname", [("wtb_bse_params", "_Params"),
                                           ("wtb_davidson_params", "_DavidsonParams"),
                                           ("wtb_davidson_info", "_DavidsonInfo")])
def test_ctypes_structures_match_the_header(tmp_path, struct, pyname):
    """The ctypes mirrors of the parameter structures have the C layout (gcc's offsetof)."""
    import shutil
    if shutil.which("gcc") is None:
        pytest.skip("no gcc")
    from wantibexos_amd import bse
    cls = getattr(bse, pyname)
    names = [f for f, _ in cls._fields_]
    off = c_offsets(tmp_path, struct, names)
    assert off["sizeof"] == ctypes.sizeof(cls)
    for f in names:
        assert off[f] == getattr(cls, f).offset, f


@pytest.mark.skipif(not LIB.exists(), reason="library not built")
@pytest.mark.parametrize("field,value,text", [("nev", 0, "nev"), ("guard", -1, "guard"),
                                              ("max_iter", 0, "max_iter"), ("max_sub", -5, "max_sub"),
                                              ("coarse", -1, "coarse"), ("tol", 0.0, "tol"),
                                              ("tol", float("nan"), "tol"), ("lanczos_steps", -1, "lanczos")])
def test_lowest_rejects_bad_parameters(field, value, text):
    """wtb_bse_plan_lowest checks every value that sizes a buffer or a loop before it
    looks at the plan, so the rejection is visible without a device (null plan)."""
    from wantibexos_amd import bse
    dp = bse._DavidsonParams(nev=8, guard=4, max_sub=40, max_iter=10, tol=1e-6, lanczos_steps=10,
                             fallback=0, coarse=0)
    setattr(dp, field, value)
    lib = bse.lib()
    w = (ctypes.c_double * 8)()
    r = (ctypes.c_double * 8)()
    rc = lib.wtb_bse_plan_lowest(None, ctypes.byref(dp), None, w, r, None, None, None)
    assert rc == 1
    assert text in lib.wtb_last_error().decode()
    # good parameters and a null plan: refused for the plan, not the parameters
    dp2 = bse._DavidsonParams(nev=8, guard=4, max_sub=40, max_iter=10, tol=1e-6, lanczos_steps=10,
                              fallback=0, coarse=0)
    assert lib.wtb_bse_plan_lowest(None, ctypes.byref(dp2), None, w, r, None, None, None) == 1
    assert "null plan" in lib.wtb_last_error().decode()
