"""Error returns of the C ABI with device arrays: each case must come back as an error code from
the host-side checks, before any kernel is queued (no case here launches a kernel on bad input).

- tosca_mesh_check rejects planted nvert > 0, meshTag > 0, aj <= 0 and NaN metrics
  (TOSCA_ERR_MESH), the restatement of SetMeshMetrics' checks (src/mesh.c:1306-1346) plus the
  IBM/overset flags;
- every entry point that takes a mesh refuses one that has not passed that check on the context
  (TOSCA_ERR_MESH), and a failed check does not register the mesh;
- a misaligned pointer and a too-small allocation are TOSCA_ERR_ARG;
- NaN in the flow fields is not an error: the step returns TOSCA_OK and the NaN stays local."""
import ctypes as C

import numpy as np
import pytest
import torch

from tosca_amd import abi
from tosca_amd import case as Ca
from tests._util import random_case

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return abi.load()


def _setup(seed=3):
    lay, mesh, f, params = random_case(n=(7, 6, 5), periodic=(2, 0, 1), seed=seed, div_scheme=Ca.DIV_CENTRAL,
                                       les_model=Ca.LES_SMAGORINSKY,
                                       bcu_nonper=(Ca.BCU_GENERIC, Ca.BCU_SLIP, Ca.BCU_NOSLIP))
    dev = {k: v.cuda() for k, v in mesh.items() if k != "cent"}
    g = {k: v.cuda() for k, v in f.items()}
    return lay, dev, g, params


def _ctx(lib, lay):
    h = C.c_void_p()
    info = abi.make_info(lay.info())
    assert lib.tosca_ctx_create(C.byref(h), C.byref(info), 0) == abi.TOSCA_OK
    return h


def _mesh(dev):
    return abi.make_mesh({n: dev[n].data_ptr() for n in abi.MESH_FIELDS})


def _formu(lib, h, dev, params, g, rhs, mesh=None):
    m = mesh if mesh is not None else _mesh(dev)
    p = abi.make_params(params)
    return lib.tosca_formu(h, C.byref(m), C.byref(p), g["ucont"].data_ptr(), g["ucat"].data_ptr(),
                           g["nut"].data_ptr(), rhs.data_ptr(), 1.0, 0)


@pytest.mark.parametrize("plant", ["nvert", "meshtag", "aj_zero", "aj_negative", "aj_nan", "csi_nan", "zet_zero"])
def test_mesh_check_rejects_planted_values(lib, plant):
    lay, dev, g, params = _setup()
    k, j, i = 5, 5, 5                                   # local indices of the interior cell (2, 2, 2)
    if plant in ("nvert", "meshtag"):
        dev[plant][k, j, i] = 1.0
    elif plant == "aj_zero":
        dev["aj"][k, j, i] = 0.0
    elif plant == "aj_negative":
        dev["aj"][k, j, i] = -dev["aj"][k, j, i]
    elif plant == "aj_nan":
        dev["aj"][k, j, i] = float("nan")
    elif plant == "csi_nan":
        dev["csi"][k, j, i, 1] = float("nan")
    else:
        dev["zet"][k, j, i] = 0.0
    h = _ctx(lib, lay)
    m = _mesh(dev)
    rc = lib.tosca_mesh_check(h, C.byref(m))
    msg = lib.tosca_last_error(h).decode()
    print(f"\nplanted {plant}: rc {rc} ({msg})", flush=True)
    assert rc == abi.TOSCA_ERR_MESH, msg
    # the rejected mesh is not registered: FormU on it is refused as unchecked
    rhs = torch.zeros((lay.mz, lay.my, lay.mx, 3), dtype=torch.float64, device="cuda")
    rc = _formu(lib, h, dev, params, g, rhs, m)
    assert rc == abi.TOSCA_ERR_MESH
    assert "tosca_mesh_check" in lib.tosca_last_error(h).decode()
    lib.tosca_ctx_destroy(h)


def test_unchecked_mesh_is_refused_everywhere(lib):
    lay, dev, g, params = _setup(4)
    h = _ctx(lib, lay)
    m, p = _mesh(dev), abi.make_params(params)
    glob = torch.zeros((lay.mz, lay.my, lay.mx, 3), dtype=torch.float64, device="cuda")
    cs = torch.zeros(lay.gshape, dtype=torch.float64, device="cuda")
    nut = torch.zeros(lay.gshape, dtype=torch.float64, device="cuda")
    bt = torch.zeros_like(glob)
    calls = {
        "formu": lambda: _formu(lib, h, dev, params, g, glob),
        "update_nut": lambda: lib.tosca_update_nut(h, C.byref(m), C.byref(p), g["ucat"].data_ptr(), cs.data_ptr(),
                                                   nut.data_ptr(), None),
        "buoyancy": lambda: lib.tosca_buoyancy(h, C.byref(m), C.byref(p), g["theta"].data_ptr(), 1.0, bt.data_ptr(),
                                               None),
        "step": lambda: lib.tosca_rhs_les_step(h, C.byref(m), C.byref(p), g["ucont"].data_ptr(), g["ucat"].data_ptr(),
                                               g["theta"].data_ptr(), glob.data_ptr(), 1.0, cs.data_ptr(),
                                               nut.data_ptr(), bt.data_ptr(), 1.0),
        "grad_p": lambda: lib.tosca_grad_p(h, C.byref(m), C.byref(p), g["theta"].data_ptr(), glob.data_ptr()),
    }
    for name, call in calls.items():
        rc = call()
        assert rc == abi.TOSCA_ERR_MESH, (name, lib.tosca_last_error(h).decode())
    assert lib.tosca_mesh_check(h, C.byref(m)) == abi.TOSCA_OK
    assert _formu(lib, h, dev, params, g, glob) == abi.TOSCA_OK
    assert lib.tosca_ctx_synchronize(h) == abi.TOSCA_OK
    lib.tosca_ctx_destroy(h)


def test_misaligned_and_short_pointers(lib):
    lay, dev, g, params = _setup(5)
    h = _ctx(lib, lay)
    m, p = _mesh(dev), abi.make_params(params)
    assert lib.tosca_mesh_check(h, C.byref(m)) == abi.TOSCA_OK
    rhs = torch.zeros((lay.mz, lay.my, lay.mx, 3), dtype=torch.float64, device="cuda")
    rc = lib.tosca_formu(h, C.byref(m), C.byref(p), g["ucont"].data_ptr() + 4, g["ucat"].data_ptr(),
                         g["nut"].data_ptr(), rhs.data_ptr(), 1.0, 0)
    assert rc == abi.TOSCA_ERR_ARG and b"aligned" in lib.tosca_last_error(h)
    short = torch.zeros(lay.mz * lay.my * lay.mx * 3 - 8, dtype=torch.float64, device="cuda")
    rc = lib.tosca_formu(h, C.byref(m), C.byref(p), g["ucont"].data_ptr(), g["ucat"].data_ptr(),
                         g["nut"].data_ptr(), short.data_ptr(), 1.0, 0)
    # torch's caching allocator may round the block up; the check is against the allocation
    if rc != abi.TOSCA_OK:
        assert rc == abi.TOSCA_ERR_ARG and b"too small" in lib.tosca_last_error(h)
    host = np.zeros(lay.mz * lay.my * lay.mx * 3)
    rc = lib.tosca_formu(h, C.byref(m), C.byref(p), g["ucont"].data_ptr(), g["ucat"].data_ptr(),
                         g["nut"].data_ptr(), host.ctypes.data, 1.0, 0)
    assert rc == abi.TOSCA_ERR_ARG
    assert lib.tosca_ctx_synchronize(h) == abi.TOSCA_OK
    lib.tosca_ctx_destroy(h)


def test_nan_flow_field_stays_local(lib):
    """NaN in one ucat value is propagated, not rejected: FormU's stencil reaches at most two cells
    away, so rhs is NaN only on faces near that cell."""
    lay, dev, g, params = _setup(6)
    params = dict(params, les_active=0)
    h = _ctx(lib, lay)
    m = _mesh(dev)
    assert lib.tosca_mesh_check(h, C.byref(m)) == abi.TOSCA_OK
    k, j, i = 4, 3, 3                                  # a cell two or more cells from every seam
    g["ucat"][k + 3, j + 3, i + 3, 0] = float("nan")
    rhs = torch.zeros((lay.mz, lay.my, lay.mx, 3), dtype=torch.float64, device="cuda")
    assert _formu(lib, h, dev, params, g, rhs) == abi.TOSCA_OK
    assert lib.tosca_ctx_synchronize(h) == abi.TOSCA_OK
    r = rhs.cpu().numpy()
    bad = np.argwhere(np.isnan(r).any(axis=-1))
    assert len(bad) > 0
    assert np.abs(bad - np.array([k, j, i])).max() <= 2, bad
    lib.tosca_ctx_destroy(h)
