"""The 384x256x128 stretched ABL box (central, constant Smagorinsky, random-Fourier turbulence) and
the 192x192x128 curvilinear hill (QUICK, dynamic Smagorinsky, random-Fourier turbulence) on the GPU.

At sizes the oracle finishes in seconds:
  - the ABL physics (Smagorinsky, central, stretched z, random-Fourier field) through the mixed step,
    against the oracle on the FP32-rounded inputs (relative L2 <= 1e-5 per field);
  - the hill physics through the adjoint: the dot-product identity <rbar, J d> = <J^T rbar, d>
    against the FP64 tangent oracle, rbar ~ N(0, 1), to 1e-10.
At the full sizes, through size-independent properties:
  - ABL: fused step against the reference-structured kernels (<= 1e-12 per field), both mixed steps
    against the FP64 step (<= 1e-5), free-stream preservation of FormU on the stretched mesh;
  - hill: the adjoint against central finite differences of the forward (eps = 1e-6) for
    cotangents rbar ~ N(0, 1) and random directions in ucont, ucat, theta and Cs.
"""
import os

import numpy as np
import pytest
import torch

from tosca_amd import case as C
from oracle import oracle as O
from tests._util import np_mesh, to_np, to_dev, relerr, maxrel

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rhs_mod():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from tosca_amd import rhs
    return rhs


def _ctx(rhs, lay, path):
    old = os.environ.get("TOSCA_PATH")
    os.environ["TOSCA_PATH"] = path
    try:
        return rhs.RhsContext(lay.info())
    finally:
        if old is None:
            os.environ.pop("TOSCA_PATH", None)
        else:
            os.environ["TOSCA_PATH"] = old


def _rel(a, b):
    d = float(torch.linalg.vector_norm((a - b).double()))
    n = float(torch.linalg.vector_norm(b.double()))
    return d / n if n > 0 else d


def _outs(lay, nut_dtype=torch.float64):
    gl = (lay.mz, lay.my, lay.mx, 3)
    return dict(rhs=torch.zeros(gl, dtype=torch.float64, device="cuda"),
                btheta=torch.zeros(gl, dtype=torch.float64, device="cuda"),
                cs=torch.zeros(lay.gshape, dtype=torch.float64, device="cuda"),
                nut=torch.zeros(lay.gshape, dtype=nut_dtype, device="cuda"))


F32_FIELDS = ("csi", "eta", "zet", "icsi", "ieta", "izet", "jcsi", "jeta", "jzet", "kcsi", "keta", "kzet",
              "iaj", "jaj", "kaj")


# ---- ABL physics at oracle size -------------------------------------------------------------------

def test_abl_smagorinsky_mixed_vs_oracle(rhs_mod):
    case = C.abl_case(nx=36, ny=70, nz=24, les_model=C.LES_SMAGORINSKY, n_modes=64, seed=3)
    lay = case.lay
    print(f"\nABL Smagorinsky mixed step {lay}", flush=True)
    nm = np_mesh(case.mesh)
    V32, u32, t32 = (to_np(t).astype(np.float32) for t in (case.ucont, case.ucat, case.theta))
    ref = dict(rhs=np.zeros((lay.mz, lay.my, lay.mx, 3)), cs=np.zeros(lay.gshape), nut=np.zeros(lay.gshape),
               btheta=np.zeros((lay.mz, lay.my, lay.mx, 3)))
    O.rhs_les_step(lay.info(), nm, case.params, V32.astype(np.float64), u32.astype(np.float64),
                   t32.astype(np.float64), ref["rhs"], 1.0, ref["cs"], ref["nut"], ref["btheta"], 1.0)
    ctx = rhs_mod.RhsContext(lay.info())
    mesh = {k: to_dev(v) for k, v in nm.items()}
    mesh32 = {k: mesh[k].float().contiguous() for k in F32_FIELDS}
    d32 = lambda a: torch.as_tensor(a).to("cuda")
    for variant in ("fp64 metrics", "fp32 metric copies"):
        o = _outs(lay, torch.float32)
        if variant == "fp64 metrics":
            ctx.rhs_les_step_mixed(mesh, case.params, d32(V32), d32(u32), d32(t32), o["rhs"], 1.0, o["cs"], o["nut"],
                                   o["btheta"], 1.0)
            want = ref
        else:
            ctx.rhs_les_step_mixed_m32(mesh, mesh32, case.params, d32(V32), d32(u32), d32(t32), o["rhs"], 1.0,
                                       o["cs"], o["nut"], o["btheta"], 1.0)
            # FormU and Buoyancy on the FP32-rounded metrics, UpdateCs / UpdateNut on the FP64 mesh
            nm32 = dict(nm)
            for k in F32_FIELDS:
                nm32[k] = nm[k].astype(np.float32).astype(np.float64)
            want = dict(cs=np.zeros(lay.gshape), nut=np.zeros(lay.gshape), rhs=np.zeros_like(ref["rhs"]),
                        btheta=np.zeros_like(ref["btheta"]))
            O.update_cs(lay.info(), nm, case.params, u32.astype(np.float64), want["cs"])
            O.update_nut(lay.info(), nm, case.params, u32.astype(np.float64), want["cs"], want["nut"])
            O.buoyancy(lay.info(), nm32, case.params, t32.astype(np.float64), 1.0, want["btheta"])
            O.formu(lay.info(), nm32, case.params, V32.astype(np.float64), u32.astype(np.float64),
                    want["nut"], want["rhs"], 1.0, 0)
        torch.cuda.synchronize()
        for k in ("rhs", "cs", "nut", "btheta"):
            got = to_np(o[k]).astype(np.float64)
            e = relerr(got, want[k])
            print(f"  {variant:18s} {k:7s} relL2={e:.2e} max={maxrel(got, want[k]):.2e}", flush=True)
            assert e <= 1e-5, (variant, k, e)
    ctx.close()


# ---- hill physics at oracle size: adjoint against the tangent oracle ------------------------------

@pytest.mark.parametrize("iperiod", [C.PERIODIC_WRAP, C.PERIODIC_COPY])
def test_hill_vjp_vs_tangent_oracle(rhs_mod, iperiod):
    case = C.hill_case(nx=16, ny=12, nz=10, turbulence="fourier", seed=4, les_model=C.LES_DSM,
                       div_scheme=C.DIV_QUICK, iperiod=iperiod)
    lay = case.lay
    info = lay.info()
    nm = np_mesh(case.mesh)
    x = dict(ucont=to_np(case.ucont), ucat=to_np(case.ucat), theta=to_np(case.theta), cs=np.zeros(lay.gshape))
    O.update_cs(info, nm, case.params, x["ucat"], x["cs"])
    rng = np.random.Generator(np.random.PCG64(4))
    rbar = rng.standard_normal((lay.mz, lay.my, lay.mx, 3))
    ctx = rhs_mod.RhsContext(info)
    dm = {k: to_dev(v) for k, v in nm.items()}
    xd = {k: to_dev(v) for k, v in x.items()}
    bars = {k: torch.zeros_like(v) for k, v in xd.items()}
    ctx.rhs_vjp(dm, case.params, xd["ucont"], xd["ucat"], xd["theta"], xd["cs"], to_dev(rbar), 1.0, 1.0,
                bars["ucont"], bars["ucat"], bars["theta"], bars["cs"])
    torch.cuda.synchronize()
    bars = {k: to_np(v) for k, v in bars.items()}
    print(f"\nhill VJP {lay} iperiod={iperiod}", flush=True)
    for name in ("ucont", "ucat", "theta", "cs"):
        d = rng.standard_normal(x[name].shape)
        _, dR = O.rhs_jvp(info, nm, case.params, x["ucont"], x["ucat"], x["theta"], x["cs"],
                          **{"d_" + name: d})
        lhs, rhs_ = float(np.sum(rbar * dR)), float(np.sum(bars[name] * d))
        err = abs(lhs - rhs_) / (float(np.sum(np.abs(rbar * dR))) + 1e-300)
        print(f"  {name:6s} <rbar,Jd>={lhs:+.12e} <J^T rbar,d>={rhs_:+.12e} rel={err:.2e}", flush=True)
        assert err <= 1e-10, (name, err)
    ctx.close()


# ---- full 384x256x128 ABL box ---------------------------------------------------------------------

@pytest.fixture(scope="module")
def abl_full(rhs_mod):
    print("\nbuilding the 384x256x128 ABL box (Smagorinsky, central, 2000 Fourier modes) on the GPU", flush=True)
    case = C.abl_case(device="cuda")
    mesh = {k: v.contiguous() for k, v in case.mesh.items() if k != "cent"}
    yield case, mesh
    del case, mesh
    torch.cuda.empty_cache()


def test_abl_full_fused_vs_reference_structured(rhs_mod, abl_full):
    case, mesh = abl_full
    lay = case.lay
    outs = {}
    for path in ("fused", "ref"):
        ctx = _ctx(rhs_mod, lay, path)
        o = _outs(lay)
        ctx.rhs_les_step(mesh, case.params, case.ucont, case.ucat, case.theta, o["rhs"], 1.0, o["cs"], o["nut"],
                         o["btheta"], 1.0)
        torch.cuda.synchronize()
        outs[path] = o
        ctx.close()
    for k in ("rhs", "cs", "nut", "btheta"):
        e = _rel(outs["fused"][k], outs["ref"][k])
        print(f"  ABL 384x256x128 fused vs reference-structured {k:7s} relL2={e:.2e}", flush=True)
        assert e <= 1e-12, (k, e)


def test_abl_full_mixed_vs_fp64(rhs_mod, abl_full):
    case, mesh = abl_full
    lay = case.lay
    ctx = _ctx(rhs_mod, lay, "fused")
    ref = _outs(lay)
    ctx.rhs_les_step(mesh, case.params, case.ucont, case.ucat, case.theta, ref["rhs"], 1.0, ref["cs"], ref["nut"],
                     ref["btheta"], 1.0)
    V32, u32, t32 = case.ucont.float().contiguous(), case.ucat.float().contiguous(), case.theta.float().contiguous()
    mesh32 = {k: mesh[k].float().contiguous() for k in F32_FIELDS}
    for variant in ("fp64 metrics", "fp32 metric copies"):
        o = _outs(lay, torch.float32)
        if variant == "fp64 metrics":
            ctx.rhs_les_step_mixed(mesh, case.params, V32, u32, t32, o["rhs"], 1.0, o["cs"], o["nut"], o["btheta"], 1.0)
        else:
            ctx.rhs_les_step_mixed_m32(mesh, mesh32, case.params, V32, u32, t32, o["rhs"], 1.0, o["cs"], o["nut"],
                                       o["btheta"], 1.0)
        torch.cuda.synchronize()
        for k in ("rhs", "cs", "nut", "btheta"):
            e = _rel(o[k].double(), ref[k])
            print(f"  ABL 384x256x128 mixed ({variant}) {k:7s} relL2={e:.2e}", flush=True)
            assert e <= 1e-5, (variant, k, e)
    del mesh32
    ctx.close()


def test_abl_full_free_stream(rhs_mod, abl_full):
    """A uniform velocity with the fluxes S.U it implies: FormU's convective part cancels to
    round-off and its viscous part is exactly 0 on the stretched mesh."""
    case, mesh = abl_full
    lay = case.lay
    ctx = _ctx(rhs_mod, lay, "fused")
    u0 = torch.tensor([7.0, -2.0, 0.5], dtype=torch.float64, device="cuda")
    ucat = u0.expand(lay.gshape + (3,)).contiguous()
    ucont = torch.stack([(mesh[f] * u0).sum(-1) for f in ("icsi", "jeta", "kzet")], -1).contiguous()
    nut = torch.full(lay.gshape, 0.3, dtype=torch.float64, device="cuda")
    gl = (lay.mz, lay.my, lay.mx, 3)
    conv = torch.zeros(gl, dtype=torch.float64, device="cuda")
    visc = torch.zeros(gl, dtype=torch.float64, device="cuda")
    ctx.FormU(mesh, case.params, ucont, ucat, nut, conv, 1.0, 1)
    ctx.FormU(mesh, case.params, ucont, ucat, nut, visc, 1.0, 2)
    torch.cuda.synchronize()
    flux_scale = float(ucont.abs().max()) * float(u0.abs().max())
    c = float(conv[1:-1, 1:-1, 1:-1].abs().max()) / flux_scale
    v = float(visc[1:-1, 1:-1, 1:-1].abs().max())
    print(f"  ABL 384x256x128 free stream: convective max/scale={c:.2e}, viscous max={v:.2e}", flush=True)
    assert c <= 1e-12
    assert v == 0.0
    ctx.close()


# ---- full 192x192x128 hill ------------------------------------------------------------------------

def test_hill_full_vjp_finite_differences(rhs_mod):
    """<rbar, (R(x + eps d) - R(x - eps d)) / (2 eps)> = <J^T rbar, d>, rbar ~ N(0, 1), eps = 1e-6,
    one random direction per input block (ucont, ucat, theta, Cs); Cs from the dynamic model."""
    print("\nbuilding the 192x192x128 hill (QUICK, DSM, Fourier turbulence) on the GPU", flush=True)
    case = C.hill_case(nx=192, ny=192, nz=128, lx=3000.0, ly=3000.0, turbulence="fourier", seed=4,
                       les_model=C.LES_DSM, div_scheme=C.DIV_QUICK, device="cuda")
    mesh = {k: v.contiguous() for k, v in case.mesh.items() if k != "cent"}
    lay = case.lay
    ctx = _ctx(rhs_mod, lay, "fused")
    gl = (lay.mz, lay.my, lay.mx, 3)
    cs = torch.zeros(lay.gshape, dtype=torch.float64, device="cuda")
    ctx.UpdateCs(mesh, case.params, case.ucat, cs)
    g = torch.Generator(device="cuda").manual_seed(4)
    rbar = torch.randn(gl, dtype=torch.float64, device="cuda", generator=g)
    x = dict(ucont=case.ucont, ucat=case.ucat, theta=case.theta, cs=cs)
    bars = {k: torch.empty_like(v) for k, v in x.items()}
    ctx.rhs_vjp(mesh, case.params, x["ucont"], x["ucat"], x["theta"], x["cs"], rbar, 1.0, 1.0,
                bars["ucont"], bars["ucat"], bars["theta"], bars["cs"])
    nut = torch.zeros(lay.gshape, dtype=torch.float64, device="cuda")
    eps = 1e-6

    def forward(xx):
        r = torch.zeros(gl, dtype=torch.float64, device="cuda")
        bt = torch.zeros(gl, dtype=torch.float64, device="cuda")
        ctx.UpdateNut(mesh, case.params, xx["ucat"], xx["cs"], nut)
        ctx.FormU(mesh, case.params, xx["ucont"], xx["ucat"], nut, r, 1.0, 0)
        ctx.Buoyancy(mesh, case.params, xx["theta"], 1.0, bt)
        return r + bt

    amp = dict(ucont=1.0, ucat=1.0, theta=1.0, cs=1e-3)
    for name in ("ucont", "ucat", "theta", "cs"):
        d = torch.randn(x[name].shape, dtype=torch.float64, device="cuda", generator=g) * amp[name]
        if name == "ucont":
            # QUICK picks its upwind cells by the sign of the face flux; the adjoint holds that
            # choice, so the direction leaves the fluxes that eps could flip (the wall faces carry
            # exactly 0) unperturbed
            d = d * (x["ucont"].abs() > 1e3 * eps).double()
        xp, xm = dict(x), dict(x)
        xp[name] = (x[name] + eps * d).contiguous()
        xm[name] = (x[name] - eps * d).contiguous()
        fd = float((rbar * (forward(xp) - forward(xm))).sum()) / (2 * eps)
        adj = float((bars[name] * d).sum())
        e = abs(adj - fd) / abs(fd)
        print(f"  hill 192x192x128 {name:6s} <J^T rbar, d> = {adj:+.10e}, FD {fd:+.10e}, rel {e:.2e}", flush=True)
        assert e <= 1e-6, (name, e)
    ctx.close()
