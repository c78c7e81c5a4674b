"""The C-ABI boundary: the library loads without a GPU and exports every
function include/midas_caking.h declares; no compute call is made here."""
import ctypes as C
import re
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "midas_caking.h"


def declared():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^\s*(?:int|void\s*\*|const char\s*\*)\s*(mc_\w+)\s*\(",
                                 text, re.M)))


@pytest.fixture(scope="module")
def lib():
    from midas_amd import _lib
    if not _lib.LIB_PATH.exists():
        subprocess.run([sys.executable, "-m", "midas_amd.build"], check=True, cwd=ROOT)
    return _lib.lib()


def test_header_declares_the_boundary():
    names = declared()
    for must in ("mc_map_build", "mc_map_load", "mc_map_write", "mc_integrate",
                 "mc_profile", "mc_fit_windows", "mc_server_submit", "mc_last_error"):
        assert must in names


def test_every_declared_symbol_is_exported(lib):
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, f"declared but not exported: {missing}"


def test_bindings_cover_the_header():
    from midas_amd import _lib
    assert set(declared()) <= set(_lib.SIGNATURES)


def test_struct_layouts(tmp_path):
    """ctypes mirrors of the header structs have the C compiler's layout."""
    from midas_amd import _lib
    src = tmp_path / "layout.c"
    src.write_text(
        '#include <stdio.h>\n#include <stddef.h>\n#include "midas_caking.h"\n'
        "int main(void){printf(\"%zu %zu %zu %zu %zu %zu %zu\\n\", sizeof(mc_geometry),"
        " offsetof(mc_geometry, TransOpt), offsetof(mc_geometry, PolarizationFraction),"
        " sizeof(mc_panel), sizeof(mc_map_info), offsetof(mc_map_info, hbm_bytes),"
        " sizeof(mc_server_stats));return 0;}\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", f"-I{ROOT / 'include'}", str(src), "-o", str(exe)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True,
                                          check=True).stdout.split()]
    want = [C.sizeof(_lib.Geometry), _lib.Geometry.TransOpt.offset,
            _lib.Geometry.PolarizationFraction.offset, C.sizeof(_lib.PanelC),
            C.sizeof(_lib.MapInfo), _lib.MapInfo.hbm_bytes.offset,
            C.sizeof(_lib.ServerStats)]
    assert got == want


def test_errors_without_gpu_are_status_codes(lib):
    from midas_amd import _lib
    n = C.c_int(-1)
    assert lib.mc_device_count(C.byref(n)) == 0
    if n.value == 0:
        h = C.c_void_p()
        rc = lib.mc_ctx_create(0, C.byref(h))
        assert rc != 0 and _lib.last_error()
    assert lib.mc_integrate(None, None, 1, 1, 0, 0, 0, None, 0, None, 0) == -1
    assert "bad arguments" in _lib.last_error()


def test_version(lib):
    a, b, c = C.c_int(), C.c_int(), C.c_int()
    assert lib.mc_version(C.byref(a), C.byref(b), C.byref(c)) == 0
    assert (a.value, b.value) >= (0, 1)
