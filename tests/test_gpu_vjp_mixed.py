"""Mixed-precision adjoint (tosca_rhs_vjp_mixed) on the GPU.

The mixed adjoint differentiates the function the mixed step evaluates, with FP32 flow fields in and
FP32 velocity and temperature cotangents out.  It is checked against the FP64 adjoint
(tosca_rhs_vjp, itself pinned to the FP64 tangent oracle by the dot-product test) evaluated at the
same FP32-rounded inputs promoted to FP64, so the difference isolates the FP32 forward recomputation
and the FP32 cotangent packets.  Bar: normwise relative L2 <= 1e-5 per output, the bar of the mixed
step.  Cases: random curvilinear lattices with slip and wall-function walls and both periodic types,
central and QUICK; the hill of the adjoint benchmark at reduced size.
"""
import numpy as np
import pytest
import torch

from tosca_amd import case as C
from tests._util import np_mesh, to_np, to_dev, random_case, relerr

pytestmark = pytest.mark.gpu

TOL = 1e-5


@pytest.fixture(scope="module")
def rhs_mod():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from tosca_amd import rhs
    return rhs


def _compare(rhs_mod, lay, nm, params, ucont, ucat, theta, cs, rbar, label, scale=0.7, bscale=1.3):
    ctx = rhs_mod.RhsContext(lay.info())
    dm = {k: to_dev(v) for k, v in nm.items()}
    u32 = {k: torch.as_tensor(v.astype(np.float32)).to("cuda") for k, v in
           (("ucont", ucont), ("ucat", ucat), ("theta", theta))}
    u64 = {k: v.double() for k, v in u32.items()}
    cs_d, rb = to_dev(cs), to_dev(rbar)
    b64 = {k: torch.zeros_like(v) for k, v in u64.items()}
    b64["cs"] = torch.zeros_like(cs_d)
    b32 = {k: torch.zeros_like(v) for k, v in u32.items()}
    b32["cs"] = torch.zeros_like(cs_d)
    ctx.rhs_vjp(dm, params, u64["ucont"], u64["ucat"], u64["theta"], cs_d, rb, scale, bscale,
                b64["ucont"], b64["ucat"], b64["theta"], b64["cs"])
    ctx.rhs_vjp_mixed(dm, params, u32["ucont"], u32["ucat"], u32["theta"], cs_d, rb, scale, bscale,
                      b32["ucont"], b32["ucat"], b32["theta"], b32["cs"])
    torch.cuda.synchronize()
    print(f"\nmixed VJP {label} {lay}", flush=True)
    worst = 0.0
    for k in ("ucont", "ucat", "theta", "cs"):
        a, b = to_np(b32[k]).astype(np.float64), to_np(b64[k])
        e = relerr(a, b)
        print(f"  {k:6s} relL2={e:.2e} |ref|={np.linalg.norm(b):.3e}", flush=True)
        worst = max(worst, e)
    ctx.close()
    assert worst <= TOL


PERS = [(0, 0, 0), (2, 0, 1), (1, 1, 1), (1, 2, 0)]


@pytest.mark.parametrize("periodic", PERS)
@pytest.mark.parametrize("scheme", [C.DIV_CENTRAL, C.DIV_QUICK])
def test_vjp_mixed_random(rhs_mod, periodic, scheme):
    lay, mesh, f, params = random_case(n=(9, 8, 10), periodic=periodic, seed=23, div_scheme=scheme,
                                       bcu_nonper=(C.BCU_GENERIC, C.BCU_SLIP, C.BCU_WALLFUNCTION))
    rng = np.random.Generator(np.random.PCG64(123))
    ti, tj = to_dev(to_np(f["tau_i"])), to_dev(to_np(f["tau_j"]))
    pg = dict(params, tau_wall_i=ti.data_ptr(), tau_wall_j=tj.data_ptr())
    cs = rng.uniform(0.01, 0.2, lay.gshape)
    rbar = rng.standard_normal((lay.mz, lay.my, lay.mx, 3))
    _compare(rhs_mod, lay, np_mesh(mesh), pg, to_np(f["ucont"]), to_np(f["ucat"]), to_np(f["theta"]), cs, rbar,
             f"periodic={periodic} scheme={scheme}")


@pytest.mark.parametrize("model", [C.LES_DSM, C.LES_SMAGORINSKY])
def test_vjp_mixed_hill(rhs_mod, model):
    case = C.hill_case(nx=40, ny=70, nz=24, turbulence="fourier", seed=4, les_model=model)
    lay = case.lay
    rng = np.random.Generator(np.random.PCG64(44))
    cs = rng.uniform(0.0, 0.05, lay.gshape)
    rbar = rng.standard_normal((lay.mz, lay.my, lay.mx, 3))
    _compare(rhs_mod, lay, np_mesh(case.mesh), case.params, to_np(case.ucont), to_np(case.ucat), to_np(case.theta),
             cs, rbar, f"hill model={model}", scale=1.0, bscale=1.0)
