"""CPU checks of the C ABI: libptgpu.so loads, exports every function include/ptgpu.h
declares, and the ctypes mirrors have the C layouts. No GPU call is made."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import ptref

pk = ptref.package()
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "ptgpu.h")


def declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(ptg_\w+)\s*\(", text, flags=re.M)))


def test_library_exports_every_declared_symbol():
    lib = pk.load_library()
    names = declared()
    assert len(names) >= 25
    for n in names:
        assert hasattr(lib, n), n
    assert sorted(pk.exported_symbols()) == names


def test_abi_version_and_null_context():
    lib = pk.load_library()
    assert lib.ptg_abi_version() == 1
    assert lib.ptg_last_error(None) == b"null context"
    assert lib.ptg_destroy(None) == 0
    assert lib.ptg_sync(None) == pk.PTG_ERR_ARG


def test_struct_layouts_match_c(tmp_path):
    src = tmp_path / "layout.c"
    src.write_text(f"""#include <stdio.h>
#include <stddef.h>
#include "{HEADER}"
#define F(T, m) printf(#T "." #m " %zu\\n", offsetof(T, m));
int main(void) {{
  printf("ptg_search_zone %zu\\nptg_grain %zu\\nptg_subpix_params %zu\\nptg_track_params %zu\\n",
         sizeof(ptg_search_zone), sizeof(ptg_grain), sizeof(ptg_subpix_params), sizeof(ptg_track_params));
  F(ptg_search_zone, inc_rot) F(ptg_grain, refrot) F(ptg_grain, NCC_subpix) F(ptg_grain, C0C0) F(ptg_grain, masked)
  F(ptg_grain, status) F(ptg_track_params, NCC_min) F(ptg_track_params, subpix) F(ptg_subpix_params, method)
  return 0;
}}
""")
    exe = tmp_path / "layout"
    subprocess.check_call(["gcc", "-std=c11", "-o", str(exe), str(src)])
    out = dict(line.rsplit(" ", 1) for line in subprocess.check_output([str(exe)]).decode().splitlines())
    assert int(out["ptg_search_zone"]) == ctypes.sizeof(pk.SearchZone)
    assert int(out["ptg_grain"]) == ctypes.sizeof(pk.Grain) == pk.GRAIN_DTYPE.itemsize
    assert int(out["ptg_subpix_params"]) == ctypes.sizeof(pk.SubpixParams)
    assert int(out["ptg_track_params"]) == ctypes.sizeof(pk.TrackParams)
    for key, val in out.items():
        if "." in key:
            t, m = key.split(".")
            cls = {"ptg_search_zone": pk.SearchZone, "ptg_grain": pk.Grain, "ptg_track_params": pk.TrackParams,
                   "ptg_subpix_params": pk.SubpixParams}[t]
            assert getattr(cls, m).offset == int(val), key


def test_grain_dtype_round_trips_through_ctypes():
    g = np.zeros(3, dtype=pk.GRAIN_DTYPE)
    g["refcoord_xpix"] = [1, 2, 3]
    g["NCC_subpix"] = [0.5, 0.25, 0.125]
    arr = (pk.Grain * 3).from_buffer(g)
    assert arr[1].refcoord_xpix == 2 and arr[2].NCC_subpix == 0.125


def test_search_zone_layout_is_search_zone_type():
    # same members, order and padding as search_zone_type (src/search_zone_type.hpp:20-27)
    assert [f[0] for f in pk.SearchZone._fields_] == ["left", "right", "up", "down", "num_rot", "inc_rot"]
    assert pk.SearchZone.inc_rot.offset == 24 and ctypes.sizeof(pk.SearchZone) == 32


def test_create_failure_reports_its_cause():
    # without a visible device ptg_create fails, releases what it made and leaves the
    # failing call's message for ptg_last_error(NULL); with a device it succeeds
    lib = pk.load_library()
    h = ctypes.c_void_p()
    rc = lib.ptg_create(ctypes.byref(h), 0)
    if rc == pk.PTG_OK:
        lib.ptg_destroy(h)
        pytest.skip("a device is visible: ptg_create succeeded")
    assert h.value is None
    msg = lib.ptg_last_error(None).decode()
    assert msg.startswith("ptg_create: ") and len(msg) > len("ptg_create: ")
    assert lib.ptg_create(None, 0) == pk.PTG_ERR_ARG
    assert lib.ptg_last_error(None).decode() == "ptg_create: null output pointer: invalid argument"
