"""CPU checks of the C ABI: the library loads, exports every function include/tosca_rhs.h declares,
and its host-side error paths return codes (never exit) without touching a GPU."""
import ctypes as C
import os
import re
import subprocess
import sys
from pathlib import Path

import pytest

from tosca_amd import abi

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "tosca_rhs.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(tosca_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_expected_api():
    names = declared_functions()
    assert set(names) == set(abi.SIGNATURES), names


def test_library_exports_every_declared_symbol():
    lib = abi.load()
    out = subprocess.run(["nm", "-D", "--defined-only", str(abi.LIB_PATH)], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r"\sT\s(tosca_\w+)", out))
    for n in declared_functions():
        assert n in exported, n
        assert getattr(lib, n) is not None


def test_struct_sizes_match_header(tmp_path):
    # tosca_dmda_info: 15 int64; tosca_mesh: 18 pointers; tosca_params: see header field order
    assert C.sizeof(abi.DmdaInfo) == 15 * 8
    assert C.sizeof(abi.Mesh) == 18 * 8
    assert C.sizeof(abi.Params) == 64 + 48 + 24 + 24 + 16 + 16 + 32
    src = r'''
#include <stdio.h>
#include <stddef.h>
#include "tosca_rhs.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu %zu %zu\n", offsetof(tosca_inlet, noise), offsetof(tosca_inlet, plane),
         offsetof(tosca_inlet, merge1), offsetof(tosca_inlet, avg_top_coords), offsetof(tosca_inlet, avg_top_u),
         offsetof(tosca_inlet, spline), offsetof(tosca_inlet, cells), offsetof(tosca_inlet, y_weights));
  printf("%zu %zu %zu %zu %zu %zu %zu %zu %zu %zu\n", sizeof(tosca_inlet), offsetof(tosca_inlet, udir), sizeof(tosca_dmda_info), sizeof(tosca_mesh), sizeof(tosca_params),
         offsetof(tosca_params, nut_value), offsetof(tosca_params, nu), offsetof(tosca_params, tau_wall_i),
         offsetof(tosca_params, wall_u), offsetof(tosca_params, wall_rough));
  return 0; }
'''
    d = tmp_path
    (d / "p.c").write_text(src)
    subprocess.run(["gcc", "-I", str(ROOT / "include"), str(d / "p.c"), "-o", str(d / "p")], check=True)
    vals = list(map(int, subprocess.run([str(d / "p")], capture_output=True, text=True, check=True).stdout.split()))
    inlet_fields = ("noise", "plane", "merge1", "avg_top_coords", "avg_top_u", "spline", "cells", "y_weights")
    assert vals[:8] == [getattr(abi.Inlet, f).offset for f in inlet_fields]
    vals = vals[8:]
    assert vals[0] == C.sizeof(abi.Inlet) and vals[1] == abi.Inlet.udir.offset
    vals = vals[2:]
    assert vals[0] == C.sizeof(abi.DmdaInfo)
    assert vals[1] == C.sizeof(abi.Mesh)
    assert vals[2] == C.sizeof(abi.Params)
    assert vals[3] == abi.Params.nut_value.offset
    assert vals[4] == abi.Params.nu.offset
    assert vals[5] == abi.Params.tau_wall_i.offset
    assert vals[6] == abi.Params.wall_u.offset
    assert vals[7] == abi.Params.wall_rough.offset


def _info(mx=10, my=8, mz=6, **kw):
    d = dict(xs=0, ys=0, zs=0, xm=mx, ym=my, zm=mz, gxs=-3, gys=-3, gzs=-3, gxm=mx + 6, gym=my + 6, gzm=mz + 6,
             mx=mx, my=my, mz=mz)
    d.update(kw)
    return abi.make_info(d)


def test_version_string():
    assert b"gfx950" in abi.load().tosca_version()


def test_null_arguments_are_errors():
    lib = abi.load()
    assert lib.tosca_ctx_create(None, None, 0) == abi.TOSCA_ERR_ARG
    h = C.c_void_p()
    assert lib.tosca_ctx_create(C.byref(h), None, 0) == abi.TOSCA_ERR_ARG
    assert lib.tosca_formu(None, None, None, None, None, None, None, 1.0, 0) == abi.TOSCA_ERR_ARG
    assert lib.tosca_update_cs(None, None, None, None, None) == abi.TOSCA_ERR_ARG
    assert lib.tosca_update_nut(None, None, None, None, None, None, None) == abi.TOSCA_ERR_ARG
    assert lib.tosca_buoyancy(None, None, None, None, 1.0, None, None) == abi.TOSCA_ERR_ARG
    assert lib.tosca_rhs_les_step(None, None, None, None, None, None, None, 1.0, None, None, None, 1.0) == abi.TOSCA_ERR_ARG
    assert lib.tosca_ctx_destroy(None) == abi.TOSCA_OK
    assert b"NULL" in lib.tosca_last_error(None)


@pytest.mark.parametrize("bad", [dict(xs=1), dict(gxs=-2), dict(gxm=15), dict(xm=9)])
def test_layout_errors(bad):
    lib = abi.load()
    h = C.c_void_p()
    info = _info(**bad)
    assert lib.tosca_ctx_create(C.byref(h), C.byref(info), 0) == abi.TOSCA_ERR_LAYOUT
    assert not h.value


def test_too_small_grid_is_layout_error():
    lib = abi.load()
    h = C.c_void_p()
    info = _info(mx=3, my=8, mz=8)
    assert lib.tosca_ctx_create(C.byref(h), C.byref(info), 0) == abi.TOSCA_ERR_LAYOUT


def test_forced_device_unavailable_is_device_error():
    code = (
        "import ctypes as C, sys; sys.path.insert(0, %r)\n"
        "from tosca_amd import abi\n"
        "lib = abi.load(); h = C.c_void_p()\n"
        "info = abi.make_info(dict(xs=0,ys=0,zs=0,xm=10,ym=8,zm=6,gxs=-3,gys=-3,gzs=-3,gxm=16,gym=14,gzm=12,mx=10,my=8,mz=6))\n"
        "rc = lib.tosca_ctx_create(C.byref(h), C.byref(info), 0)\n"
        "print(rc, lib.tosca_last_error(None).decode())\n" % str(ROOT))
    env = dict(os.environ, TOSCA_FORCE_NO_DEVICE="1")
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, env=env, check=True).stdout
    rc, msg = out.split(" ", 1)
    assert int(rc) == abi.TOSCA_ERR_DEVICE
    assert "TOSCA_FORCE_NO_DEVICE" in msg
