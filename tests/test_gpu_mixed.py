"""Mixed-precision step (tosca_rhs_les_step_mixed) against the FP64 CPU oracle.

The mixed path takes FP32 flow fields (ucont, ucat, theta; nut out) and keeps the metrics and
the outputs rhs, cs, btheta in FP64.  Two errors are reported per output field:
  - arithmetic: against the oracle run on the same FP32-rounded inputs (promoted to FP64), which
    isolates the FP32 arithmetic of the kernels.  Bar: relative L2 <= 1e-5;
  - total: against the oracle on the original FP64 inputs, which adds the rounding of the inputs
    to FP32 (printed, and held to 1e-3 as a sanity bound).
"""
import numpy as np
import pytest
import torch

from tosca_amd import case as C
from oracle import oracle as O
from tests._util import np_mesh, to_np, to_dev, relerr, maxrel

pytestmark = pytest.mark.gpu

ARITH_TOL = 1e-5
TOTAL_TOL = 1e-3


@pytest.fixture(scope="module")
def rhs_mod():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from tosca_amd import rhs
    return rhs


def _rms_rel(a, b):
    return float(np.sqrt(np.mean((a - b) ** 2)) / max(np.sqrt(np.mean(b ** 2)), 1e-300))


def _oracle(case, ucont, ucat, theta):
    lay = case.lay
    out = dict(rhs=np.zeros((lay.mz, lay.my, lay.mx, 3)), cs=np.zeros(lay.gshape), nut=np.zeros(lay.gshape),
               btheta=np.zeros((lay.mz, lay.my, lay.mx, 3)))
    O.rhs_les_step(lay.info(), np_mesh(case.mesh), case.params, ucont, ucat, theta, out["rhs"], 1.0, out["cs"],
                   out["nut"], out["btheta"], 1.0)
    return out


CASES = [
    ("tgv smagorinsky", lambda: C.tgv_case(n=16, ptype=C.PERIODIC_WRAP)),
    ("tgv dsm quick copy", lambda: C.tgv_case(n=16, ptype=C.PERIODIC_COPY, les_model=C.LES_DSM, div_scheme=C.DIV_QUICK)),
    ("hill dsm quick", lambda: C.hill_case(nx=24, ny=12, nz=10, les_model=C.LES_DSM)),
    ("hill multitile", lambda: C.hill_case(nx=40, ny=130, nz=20, les_model=C.LES_DSM)),
    ("abl dsm", lambda: C.abl_case(nx=36, ny=70, nz=24, les_model=C.LES_DSM, n_modes=64)),
    ("shard dsm", lambda: C.shard_case(nx=64, ny=64, nz=32)),
    ("hill dpasd", lambda: C.hill_case(nx=24, ny=12, nz=10, les_model=C.LES_DPASD)),
    ("hill dpasd multitile", lambda: C.hill_case(nx=40, ny=130, nz=28, les_model=C.LES_DPASD)),
    # 25 k planes in chunks of 9, 9, 7: the FP32 k pass's last step of four has one or three outputs
    ("hill dpasd odd chunks", lambda: C.hill_case(nx=25, ny=12, nz=10, les_model=C.LES_DPASD)),
]


@pytest.mark.parametrize("name,builder", CASES, ids=[c[0] for c in CASES])
def test_mixed_step(rhs_mod, name, builder):
    case = builder()
    lay = case.lay
    print(f"\nmixed step: {name} {lay}", flush=True)
    u64, V64, t64 = to_np(case.ucat), to_np(case.ucont), to_np(case.theta)
    u32, V32, t32 = (a.astype(np.float32) for a in (u64, V64, t64))
    ref_arith = _oracle(case, V32.astype(np.float64), u32.astype(np.float64), t32.astype(np.float64))
    ref_total = _oracle(case, V64, u64, t64)

    ctx = rhs_mod.RhsContext(lay.info())
    mesh = {k: to_dev(v) for k, v in np_mesh(case.mesh).items()}
    dev32 = lambda a: torch.as_tensor(a).to("cuda")
    rhs = torch.zeros((lay.mz, lay.my, lay.mx, 3), dtype=torch.float64, device="cuda")
    bt = torch.zeros_like(rhs)
    cs = torch.zeros(lay.gshape, dtype=torch.float64, device="cuda")
    nut = torch.zeros(lay.gshape, dtype=torch.float32, device="cuda")
    ctx.rhs_les_step_mixed(mesh, case.params, dev32(V32), dev32(u32), dev32(t32), rhs, 1.0, cs, nut, bt, 1.0)
    torch.cuda.synchronize()
    got = dict(rhs=to_np(rhs), cs=to_np(cs), nut=to_np(nut).astype(np.float64), btheta=to_np(bt))
    worst = 0.0
    for k in ("rhs", "cs", "nut", "btheta"):
        ea, et = relerr(got[k], ref_arith[k]), relerr(got[k], ref_total[k])
        print(f"  {k:7s} arithmetic relL2={ea:.2e} max={maxrel(got[k], ref_arith[k]):.2e} "
              f"rms={_rms_rel(got[k], ref_arith[k]):.2e} | total relL2={et:.2e} "
              f"max={maxrel(got[k], ref_total[k]):.2e}", flush=True)
        worst = max(worst, ea)
        assert et <= TOTAL_TOL, (k, et)
    assert worst <= ARITH_TOL


def test_mixed_rejects_unsupported_model(rhs_mod):
    case = C.hill_case(nx=12, ny=8, nz=8, les_model=C.LES_AMD)
    lay = case.lay
    ctx = rhs_mod.RhsContext(lay.info())
    mesh = {k: to_dev(v) for k, v in np_mesh(case.mesh).items()}
    f = lambda t: t.to(torch.float32).to("cuda")
    rhs = torch.zeros((lay.mz, lay.my, lay.mx, 3), dtype=torch.float64, device="cuda")
    cs = torch.zeros(lay.gshape, dtype=torch.float64, device="cuda")
    nut = torch.zeros(lay.gshape, dtype=torch.float32, device="cuda")
    from tosca_amd.abi import ToscaError
    with pytest.raises(ToscaError):
        ctx.rhs_les_step_mixed(mesh, case.params, f(case.ucont), f(case.ucat), f(case.theta), rhs, 1.0, cs, nut,
                               rhs.clone(), 1.0)


F32_FIELDS = ("csi", "eta", "zet", "icsi", "ieta", "izet", "jcsi", "jeta", "jzet", "kcsi", "keta", "kzet",
              "iaj", "jaj", "kaj")


@pytest.mark.parametrize("name,builder", CASES, ids=[c[0] for c in CASES])
def test_mixed_step_fp32_metrics(rhs_mod, name, builder):
    """tosca_rhs_les_step_mixed_m32: FormU (and the fused Buoyancy) read FP32 metric copies.
    Arithmetic reference: the oracle's UpdateCs / UpdateNut on the FP64 mesh and FP32-rounded
    velocity, its Buoyancy and FormU on the FP32-rounded metrics; bar 1e-5 as for the mixed step."""
    case = builder()
    lay = case.lay
    print(f"\nmixed step, FP32 metrics: {name} {lay}", flush=True)
    nm = np_mesh(case.mesh)
    nm_r = dict(nm)
    for k in F32_FIELDS:
        nm_r[k] = nm[k].astype(np.float32).astype(np.float64)
    u64, V64, t64 = to_np(case.ucat), to_np(case.ucont), to_np(case.theta)
    u, V, t = (a.astype(np.float32).astype(np.float64) for a in (u64, V64, t64))
    info, prm = lay.info(), case.params
    cs = np.zeros(lay.gshape); nut = np.zeros(lay.gshape)
    O.update_cs(info, nm, prm, u, cs)
    O.update_nut(info, nm, prm, u, cs, nut)
    bt = np.zeros((lay.mz, lay.my, lay.mx, 3))
    O.buoyancy(info, nm_r, prm, t, 1.0, bt)
    rhs = np.zeros((lay.mz, lay.my, lay.mx, 3))
    O.formu(info, nm_r, prm, V, u, nut, rhs, 1.0, 0)
    ref_arith = dict(rhs=rhs, cs=cs, nut=nut, btheta=bt)
    ref_total = _oracle(case, V64, u64, t64)

    ctx = rhs_mod.RhsContext(info)
    mesh = {k: to_dev(v) for k, v in nm.items()}
    mesh32 = {k: torch.as_tensor(nm[k].astype(np.float32)).to("cuda") for k in F32_FIELDS}
    dev32 = lambda a: torch.as_tensor(a.astype(np.float32)).to("cuda")
    rhs_g = torch.zeros((lay.mz, lay.my, lay.mx, 3), dtype=torch.float64, device="cuda")
    bt_g = torch.zeros_like(rhs_g)
    cs_g = torch.zeros(lay.gshape, dtype=torch.float64, device="cuda")
    nut_g = torch.zeros(lay.gshape, dtype=torch.float32, device="cuda")
    ctx.rhs_les_step_mixed_m32(mesh, mesh32, prm, dev32(V64), dev32(u64), dev32(t64), rhs_g, 1.0, cs_g, nut_g, bt_g,
                               1.0)
    torch.cuda.synchronize()
    got = dict(rhs=to_np(rhs_g), cs=to_np(cs_g), nut=to_np(nut_g).astype(np.float64), btheta=to_np(bt_g))
    worst = 0.0
    for k in ("rhs", "cs", "nut", "btheta"):
        ea, et = relerr(got[k], ref_arith[k]), relerr(got[k], ref_total[k])
        print(f"  {k:7s} arithmetic relL2={ea:.2e} max={maxrel(got[k], ref_arith[k]):.2e} "
              f"rms={_rms_rel(got[k], ref_arith[k]):.2e} | total relL2={et:.2e} "
              f"max={maxrel(got[k], ref_total[k]):.2e}", flush=True)
        worst = max(worst, ea)
        assert et <= TOTAL_TOL, (k, et)
    assert worst <= ARITH_TOL
