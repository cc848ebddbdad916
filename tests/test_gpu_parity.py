"""GPU parity: the HIP library (through the C ABI) against the CPU oracle on identical inputs.

Bar: normwise relative L2 error <= 1e-12 per output field for the FP64 path.
Output buffers start from the same random values on both sides, so entries a routine must not
touch are checked as well.
"""
import itertools

import numpy as np
import pytest
import torch

from tosca_amd import case as C
from oracle import oracle as O
from tests._util import np_mesh, to_np, to_dev, relerr, maxrel, random_case, out_arrays

pytestmark = pytest.mark.gpu

TOL = 1e-12


@pytest.fixture(scope="module", params=["fused", "ref"])
def rhs_mod(request):
    """The library module, with contexts created for the fused (default) or the
    reference-structured kernel path."""
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    import os
    from tosca_amd import rhs
    old = os.environ.get("TOSCA_PATH")
    os.environ["TOSCA_PATH"] = request.param
    yield rhs
    if old is None:
        os.environ.pop("TOSCA_PATH", None)
    else:
        os.environ["TOSCA_PATH"] = old


def _dev_mesh(mesh):
    return {k: to_dev(v) for k, v in np_mesh(mesh).items()}


def _report(name, a, b):
    e, m = relerr(a, b), maxrel(a, b)
    print(f"  {name:8s} relL2={e:.3e} maxrel={m:.3e}", flush=True)
    return e


PER_COMBOS = [(2, 0, 1), (0, 0, 0), (1, 1, 1), (2, 2, 2), (0, 2, 0), (1, 0, 2)]


@pytest.mark.parametrize("periodic", PER_COMBOS)
@pytest.mark.parametrize("scheme", [C.DIV_CENTRAL, C.DIV_QUICK])
@pytest.mark.parametrize("mode", [0, 1, 2])
def test_formu_random(rhs_mod, periodic, scheme, mode):
    seed = hash((periodic, scheme, mode)) % 1000
    lay, mesh, f, params = random_case(n=(7, 6, 5), periodic=periodic, seed=seed, div_scheme=scheme,
                                       bcu_nonper=(C.BCU_GENERIC, C.BCU_SLIP, C.BCU_WALLFUNCTION))
    print(f"\nFormU periodic={periodic} scheme={scheme} mode={mode}", flush=True)
    out = out_arrays(lay)
    nm = np_mesh(mesh)
    tau_i, tau_j = to_np(f["tau_i"]), to_np(f["tau_j"])
    po = dict(params, tau_wall_i=tau_i.ctypes.data, tau_wall_j=tau_j.ctypes.data)
    rhs_o = out["rhs"].copy()
    O.formu(lay.info(), nm, po, to_np(f["ucont"]), to_np(f["ucat"]), to_np(f["nut"]), rhs_o, 0.7, mode)
    ctx = rhs_mod.RhsContext(lay.info())
    dm = _dev_mesh(mesh)
    ti, tj = to_dev(tau_i), to_dev(tau_j)
    pg = dict(params, tau_wall_i=ti.data_ptr(), tau_wall_j=tj.data_ptr())
    rhs_g = to_dev(out["rhs"])
    ctx.FormU(dm, pg, to_dev(to_np(f["ucont"])), to_dev(to_np(f["ucat"])), to_dev(to_np(f["nut"])), rhs_g, 0.7, mode)
    torch.cuda.synchronize()
    assert _report("rhs", to_np(rhs_g), rhs_o) <= TOL


@pytest.mark.parametrize("model", [C.LES_SMAGORINSKY, C.LES_DSM, C.LES_DPASD])
@pytest.mark.parametrize("periodic", [(2, 0, 1), (0, 0, 0), (1, 1, 1), (2, 2, 2)])
@pytest.mark.parametrize("max_cs", [0.5, 1e6])
def test_update_cs_nut_random(rhs_mod, model, periodic, max_cs):
    seed = hash((model, periodic, max_cs)) % 1000
    lay, mesh, f, params = random_case(n=(7, 6, 8), periodic=periodic, seed=seed, les_model=model, max_cs=max_cs)
    # smooth-ish velocity so the dynamic coefficient is not all clipped
    print(f"\nUpdateCs/UpdateNut model={model} periodic={periodic} max_cs={max_cs}", flush=True)
    out = out_arrays(lay)
    nm = np_mesh(mesh)
    ucat = to_np(f["ucat"])
    cs_o, nut_o, ks_o = out["cs"].copy(), out["nut"].copy(), out["ksgs"].copy()
    O.update_cs(lay.info(), nm, params, ucat, cs_o)
    O.update_nut(lay.info(), nm, params, ucat, cs_o, nut_o, ks_o)
    ctx = rhs_mod.RhsContext(lay.info())
    dm = _dev_mesh(mesh)
    cs_g, nut_g, ks_g = to_dev(out["cs"]), to_dev(out["nut"]), to_dev(out["ksgs"])
    u = to_dev(ucat)
    ctx.UpdateCs(dm, params, u, cs_g)
    ctx.UpdateNut(dm, params, u, cs_g, nut_g, ks_g)
    torch.cuda.synchronize()
    e1 = _report("cs", to_np(cs_g), cs_o)
    e2 = _report("nut", to_np(nut_g), nut_o)
    e3 = _report("ksgs", to_np(ks_g), ks_o)
    assert max(e1, e2, e3) <= TOL


@pytest.mark.parametrize("periodic", [(2, 0, 1), (0, 0, 0), (1, 1, 1)])
def test_buoyancy_random(rhs_mod, periodic):
    lay, mesh, f, params = random_case(n=(6, 5, 7), periodic=periodic, seed=3)
    out = out_arrays(lay)
    nm = np_mesh(mesh)
    bt_o, gc_o = out["btheta"].copy(), out["gcont"].copy()
    O.buoyancy(lay.info(), nm, params, to_np(f["theta"]), 1.3, bt_o, gc_o)
    ctx = rhs_mod.RhsContext(lay.info())
    bt_g, gc_g = to_dev(out["btheta"]), to_dev(out["gcont"])
    ctx.Buoyancy(_dev_mesh(mesh), params, to_dev(to_np(f["theta"])), 1.3, bt_g, gc_g)
    torch.cuda.synchronize()
    print(f"\nBuoyancy periodic={periodic}", flush=True)
    assert max(_report("btheta", to_np(bt_g), bt_o), _report("gcont", to_np(gc_g), gc_o)) <= TOL


def _step_parity(rhs_mod, case, label):
    lay = case.lay
    print(f"\nrhs_les_step {label} {lay}", flush=True)
    out = out_arrays(lay)
    nm = np_mesh(case.mesh)
    ucont, ucat, theta = to_np(case.ucont), to_np(case.ucat), to_np(case.theta)
    o = {k: out[k].copy() for k in ("rhs", "cs", "nut", "btheta")}
    O.rhs_les_step(lay.info(), nm, case.params, ucont, ucat, theta, o["rhs"], 1.0, o["cs"], o["nut"], o["btheta"], 1.0)
    ctx = rhs_mod.RhsContext(lay.info())
    g = {k: to_dev(out[k]) for k in ("rhs", "cs", "nut", "btheta")}
    ctx.rhs_les_step(_dev_mesh(case.mesh), case.params, to_dev(ucont), to_dev(ucat), to_dev(theta), g["rhs"], 1.0,
                     g["cs"], g["nut"], g["btheta"], 1.0)
    torch.cuda.synchronize()
    errs = [_report(k, to_np(g[k]), o[k]) for k in ("rhs", "cs", "nut", "btheta")]
    assert max(errs) <= TOL


@pytest.mark.parametrize("ptype", [C.PERIODIC_COPY, C.PERIODIC_WRAP])
@pytest.mark.parametrize("model,scheme", [(C.LES_SMAGORINSKY, C.DIV_CENTRAL), (C.LES_DSM, C.DIV_QUICK)])
def test_step_tgv(rhs_mod, ptype, model, scheme):
    _step_parity(rhs_mod, C.tgv_case(n=16, ptype=ptype, les_model=model, div_scheme=scheme), f"tgv p{ptype} m{model}")


@pytest.mark.parametrize("model", [C.LES_DSM, C.LES_DPASD])
@pytest.mark.parametrize("iper", [C.PERIODIC_COPY, C.PERIODIC_WRAP])
def test_step_hill(rhs_mod, model, iper):
    _step_parity(rhs_mod, C.hill_case(nx=24, ny=12, nz=10, les_model=model, iperiod=iper), f"hill m{model} i{iper}")


def test_step_abl(rhs_mod):
    _step_parity(rhs_mod, C.abl_case(nx=20, ny=12, nz=10, n_modes=64), "abl smag central")


# multi-tile cases: several 62-wide i tiles, several j tiles and several k chunks, so the tile
# seams, the periodic seams and the k-chunk priming of the fused kernels are all exercised
@pytest.mark.parametrize("periodic", [(2, 0, 2), (1, 1, 0), (0, 0, 0)])
@pytest.mark.parametrize("scheme", [C.DIV_CENTRAL, C.DIV_QUICK])
def test_formu_random_multitile(rhs_mod, periodic, scheme):
    lay, mesh, f, params = random_case(n=(37, 130, 14), periodic=periodic, seed=11, div_scheme=scheme, amp=0.1)
    print(f"\nFormU multi-tile periodic={periodic} scheme={scheme} {lay}", flush=True)
    out = out_arrays(lay)
    nm = np_mesh(mesh)
    rhs_o = out["rhs"].copy()
    O.formu(lay.info(), nm, params, to_np(f["ucont"]), to_np(f["ucat"]), to_np(f["nut"]), rhs_o, 1.0, 0)
    ctx = rhs_mod.RhsContext(lay.info())
    rhs_g = to_dev(out["rhs"])
    ctx.FormU(_dev_mesh(mesh), params, to_dev(to_np(f["ucont"])), to_dev(to_np(f["ucat"])), to_dev(to_np(f["nut"])),
              rhs_g, 1.0, 0)
    torch.cuda.synchronize()
    assert _report("rhs", to_np(rhs_g), rhs_o) <= TOL


@pytest.mark.parametrize("builder,kw", [
    (C.tgv_case, dict(n=70, ptype=C.PERIODIC_WRAP, les_model=C.LES_DSM, div_scheme=C.DIV_QUICK)),
    (C.tgv_case, dict(n=66, ptype=C.PERIODIC_COPY, les_model=C.LES_SMAGORINSKY, div_scheme=C.DIV_CENTRAL)),
    (C.hill_case, dict(nx=40, ny=130, nz=20, les_model=C.LES_DSM)),
    (C.abl_case, dict(nx=36, ny=70, nz=24, les_model=C.LES_DSM, n_modes=64)),
    (C.hill_case, dict(nx=40, ny=130, nz=28, les_model=C.LES_DPASD)),
    (C.abl_case, dict(nx=40, ny=70, nz=30, les_model=C.LES_DPASD, n_modes=64)),
])
def test_step_multitile(rhs_mod, builder, kw):
    _step_parity(rhs_mod, builder(**kw), f"{builder.__name__} {kw}")


# several i tiles, j tiles and k chunks of the plane-averaged model's two filter kernels, and
# several partials per j plane
@pytest.mark.parametrize("periodic", [(2, 0, 1), (0, 0, 0), (1, 1, 2)])
def test_update_cs_dpasd_multitile(rhs_mod, periodic):
    lay, mesh, f, params = random_case(n=(40, 130, 28), periodic=periodic, seed=17, les_model=C.LES_DPASD, max_cs=1e6,
                                       amp=0.1)
    print(f"\nUpdateCs DPASD multi-tile periodic={periodic} {lay}", flush=True)
    out = out_arrays(lay)
    nm = np_mesh(mesh)
    ucat = to_np(f["ucat"])
    cs_o = out["cs"].copy()
    O.update_cs(lay.info(), nm, params, ucat, cs_o)
    ctx = rhs_mod.RhsContext(lay.info())
    cs_g = to_dev(out["cs"])
    ctx.UpdateCs(_dev_mesh(mesh), params, to_dev(ucat), cs_g)
    torch.cuda.synchronize()
    assert _report("cs", to_np(cs_g), cs_o) <= TOL


@pytest.mark.parametrize("model", [C.LES_DSM, C.LES_DPASD])
@pytest.mark.parametrize("periodic", [(2, 0, 1), (0, 0, 0)])
def test_update_cs_uniform_velocity(rhs_mod, model, periodic):
    """Uniform velocity: L = M = 0.  The dynamic coefficient is 0; the plane-averaged model divides
    0 by beta = 0 and its clip (two plain comparisons) turns the NaN into maxCs."""
    lay, mesh, f, params = random_case(n=(7, 6, 8), periodic=periodic, seed=2, les_model=model)
    u = np.broadcast_to(np.array([1.5, -0.5, 0.25]), lay.gshape + (3,)).copy()
    out = out_arrays(lay)
    nm = np_mesh(mesh)
    cs_o = out["cs"].copy()
    O.update_cs(lay.info(), nm, params, u, cs_o)
    ctx = rhs_mod.RhsContext(lay.info())
    cs_g = to_dev(out["cs"])
    ctx.UpdateCs(_dev_mesh(mesh), params, to_dev(u), cs_g)
    torch.cuda.synchronize()
    print(f"\nUpdateCs uniform velocity model={model} periodic={periodic}", flush=True)
    assert np.array_equal(to_np(cs_g), cs_o)


@pytest.mark.parametrize("t_ref", [300.0, 287.3])
def test_step_buoyancy_near_tref(rhs_mod, t_ref):
    """theta within 1e-3 K of t_ref: (t_ref - theta) cancels to a few digits, so any change in how
    t_ref enters (a reciprocal, a re-derived t_ref) shows at the 1e-12 level."""
    case = C.hill_case(nx=20, ny=12, nz=10, les_model=C.LES_DSM)
    rng = np.random.Generator(np.random.PCG64(17))
    case.params = dict(case.params, t_ref=t_ref)
    case.theta = torch.as_tensor(t_ref + 1e-3 * rng.standard_normal(tuple(case.theta.shape)))
    _step_parity(rhs_mod, case, f"buoyancy near t_ref={t_ref}")
