"""GPU parity of the wind-farm body force against the oracle (through the C ABI):
tosca_project_uadm (projectBladeForce, uniformADM branch), tosca_body_force_to_cont
(bodyForceCartesian2Contravariant) and the step with the farm source (tosca_rhs_les_step_farm,
and its mixed-precision form against the FP64 one).

The projection sums exp(-(r/eps)^2) terms whose exp comes from the device's math library on one
side and glibc on the other; both are faithfully rounded, so the bar is 1e-13 (normwise) there.
The face conversion is a copy of the oracle's expressions under FP contraction off: bit-identical.
"""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from tosca_amd import case as C
from tosca_amd import farm as F
from tests._util import np_mesh, out_arrays, random_case, relerr, to_dev, to_np

pytestmark = pytest.mark.gpu


def _ctx(lay):
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from tosca_amd import rhs
    return rhs, rhs.RhsContext(lay.info())


def _small_farm(lengths, eps, n=(2, 2), spacing=None):
    t = F.UadmTurbine(r_tip=2.5, r_hub=0.4, h_tower=4.0, over_hang=0.3, epsilon=eps, n_rad=5, n_azi=12,
                      ct=0.75, uref=3.0, up_tilt=4.0)
    sp = spacing or 0.4 * min(lengths)
    return F.grid_farm(t, n[0], n[1], sp, lengths)


@pytest.mark.parametrize("periodic", [(0, 0, 0), (2, 0, 2), (1, 0, 1)])
def test_project_uadm_matches_oracle(periodic):
    ax = np.arange(25) * 0.5
    nodes = C.cartesian_nodes(ax, ax[:21], ax[:17] * 0.9, "cpu")
    lay, mesh = C.mesh_metrics(nodes, periodic)
    rhs, ctx = _ctx(lay)
    farm = _small_farm((12.0, 10.0), 0.9)
    params = C.default_params(periodic, (C.BC_PERIODIC if periodic[0] else C.BC_SLIP,) * 6, (C.BC_ZEROGRAD,) * 6)
    cent = mesh["cent"].numpy().copy()
    ref = np.full(lay.gshape + (3,), np.nan)
    O.project_uadm(lay.info(), params, cent, farm, ref)
    got = torch.full(lay.gshape + (3,), float("nan"), dtype=torch.float64, device="cuda")
    print(f"\nproject_uadm periodic={periodic} rotors={farm.n_rotors} points={farm.n_points}", flush=True)
    ctx.ProjectUADM(params, to_dev(cent), rhs.DeviceFarm(farm), got)
    torch.cuda.synchronize()
    g = to_np(got)
    assert np.abs(ref).max() > 0
    assert np.array_equal(np.isnan(g), np.isnan(ref))
    mask = ~np.isnan(ref)
    e = relerr(g[mask], ref[mask])
    print(f"  scat relL2 vs oracle {e:.2e}", flush=True)
    assert e <= 1e-13
    ctx.close()


def test_body_force_to_cont_bit_identical():
    lay, mesh, f, params = random_case(n=(7, 6, 8), periodic=(2, 0, 1), seed=11)
    rhs, ctx = _ctx(lay)
    m = np_mesh(mesh)
    rng = np.random.Generator(np.random.PCG64(5))
    scat = rng.standard_normal(lay.gshape + (3,))
    ref = out_arrays(lay)["rhs"]
    O.body_force_to_cont(lay.info(), m, params, scat, ref)
    got = to_dev(out_arrays(lay)["rhs"])
    dmesh = {k: to_dev(v) for k, v in m.items()}
    ctx.BodyForceToCont(dmesh, params, to_dev(scat), got)
    torch.cuda.synchronize()
    assert np.array_equal(to_np(got), ref)
    ctx.close()


@pytest.mark.parametrize("seed,periodic,model", [(1, (2, 0, 1), C.LES_DSM), (2, (0, 1, 2), C.LES_SMAGORINSKY),
                                                 (3, (1, 0, 0), C.LES_DSM)])
def test_rhs_les_step_farm_matches_oracle(seed, periodic, model):
    lay, mesh, f, params = random_case(n=(9, 8, 10), periodic=periodic, seed=seed, les_model=model)
    rhs, ctx = _ctx(lay)
    m = np_mesh(mesh)
    rng = np.random.Generator(np.random.PCG64(40 + seed))
    farm = rng.standard_normal((lay.mz, lay.my, lay.mx, 3))
    o = out_arrays(lay)
    O.rhs_les_step_farm(lay.info(), m, params, to_np(f["ucont"]), to_np(f["ucat"]), to_np(f["theta"]), farm, o["rhs"],
                        0.7, o["cs"], o["nut"], o["btheta"], 1.3)
    g = {k: to_dev(v) for k, v in out_arrays(lay).items()}
    dmesh = {k: to_dev(v) for k, v in m.items()}
    print(f"\nrhs_les_step_farm seed={seed} periodic={periodic} model={model}", flush=True)
    ctx.rhs_les_step_farm(dmesh, params, to_dev(f["ucont"]), to_dev(f["ucat"]), to_dev(f["theta"]), to_dev(farm),
                          g["rhs"], 0.7, g["cs"], g["nut"], g["btheta"], 1.3)
    torch.cuda.synchronize()
    for k in ("rhs", "cs", "nut", "btheta"):
        e = relerr(to_np(g[k]), o[k])
        print(f"  {k:7s} relL2 {e:.2e}", flush=True)
        assert e <= 1e-12, k
    ctx.close()


def test_mixed_farm_matches_fp64_farm():
    case = C.hill_case(nx=24, ny=20, nz=16, les_model=C.LES_DSM, div_scheme=C.DIV_QUICK, device="cpu")
    lay = case.lay
    rhs, ctx = _ctx(lay)
    mesh = {k: v.cuda() for k, v in case.mesh.items() if k != "cent"}
    t = F.UadmTurbine(r_tip=63.0, r_hub=1.5, epsilon=40.0, n_rad=6, n_azi=16, ct=0.75, uref=8.0)
    farm_t = F.grid_farm(t, 2, 1, 900.0, (3000.0, 1500.0))
    scat = torch.zeros(lay.gshape + (3,), dtype=torch.float64, device="cuda")
    scont = torch.zeros((lay.mz, lay.my, lay.mx, 3), dtype=torch.float64, device="cuda")
    ctx.ProjectUADM(case.params, case.mesh["cent"].cuda(), rhs.DeviceFarm(farm_t), scat)
    ctx.BodyForceToCont(mesh, case.params, scat, scont)
    torch.cuda.synchronize()
    assert float(scont.abs().max()) > 0
    gl = (lay.mz, lay.my, lay.mx, 3)
    o64 = dict(rhs=torch.zeros(gl, dtype=torch.float64, device="cuda"), btheta=torch.zeros(gl, dtype=torch.float64, device="cuda"),
               cs=torch.zeros(lay.gshape, dtype=torch.float64, device="cuda"),
               nut=torch.zeros(lay.gshape, dtype=torch.float64, device="cuda"))
    o32 = dict(o64, rhs=o64["rhs"].clone(), btheta=o64["btheta"].clone(), cs=o64["cs"].clone(),
               nut=torch.zeros(lay.gshape, dtype=torch.float32, device="cuda"))
    # FP32-rounded flow fields on both sides: the comparison sees the mixed arithmetic only
    uc, ua, th = (case.ucont.float().double().cuda(), case.ucat.float().double().cuda(),
                  case.theta.float().double().cuda())
    ctx.rhs_les_step_farm(mesh, case.params, uc, ua, th, scont, o64["rhs"], 1.0, o64["cs"], o64["nut"], o64["btheta"])
    ctx.rhs_les_step_mixed_farm(mesh, case.params, uc.float(), ua.float(), th.float(), scont, o32["rhs"], 1.0,
                                o32["cs"], o32["nut"], o32["btheta"])
    torch.cuda.synchronize()
    for k in ("rhs", "cs", "nut", "btheta"):
        a, b = o32[k].double(), o64[k]
        e = float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))
        print(f"  mixed farm {k:7s} relL2 vs FP64 {e:.2e}", flush=True)
        assert e <= 1e-5, k
    # the farm term itself: rhs with farm minus rhs without farm is the farm source, to round-off
    r0 = torch.zeros(gl, dtype=torch.float64, device="cuda")
    ctx.rhs_les_step(mesh, case.params, uc, ua, th, r0, 1.0, o64["cs"].clone(), o64["nut"].clone(),
                     o64["btheta"].clone())
    torch.cuda.synchronize()
    d = o64["rhs"] - r0
    e = float(torch.linalg.vector_norm(d - scont) / torch.linalg.vector_norm(scont))
    print(f"  farm term recovered relL2 {e:.2e}", flush=True)
    assert e <= 1e-9
    ctx.close()
