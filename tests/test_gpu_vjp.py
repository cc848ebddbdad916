"""Hand-written adjoint of the right-hand side (tosca_rhs_vjp) on the GPU.

R(ucont, ucat, theta, Cs) = FormU(ucont, ucat, UpdateNut(ucat, Cs)) + Buoyancy(theta).  Checks:
  - dot-product test against the FP64 tangent oracle (oracle/tosca_oracle_tangent.cpp, itself pinned
    to finite differences on the CPU): <rbar, J d> = <J^T rbar, d> for random directions in each
    input block, to 1e-10 relative;
  - central finite differences of the GPU forward (eps = 1e-6), to 1e-6 relative;
  - torch.autograd.gradcheck of the autograd wrapper on a small grid.
"""
import numpy as np
import pytest
import torch

from tosca_amd import case as C
from oracle import oracle as O
from tests._util import np_mesh, to_np, to_dev, random_case

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mods():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from tosca_amd import rhs, autograd
    return rhs, autograd


def _setup(periodic, seed, scheme, n=(7, 6, 8)):
    lay, mesh, f, params = random_case(n=n, periodic=periodic, seed=seed, div_scheme=scheme,
                                       bcu_nonper=(C.BCU_GENERIC, C.BCU_SLIP, C.BCU_WALLFUNCTION))
    rng = np.random.Generator(np.random.PCG64(70 + seed))
    x = dict(ucont=to_np(f["ucont"]), ucat=to_np(f["ucat"]), theta=to_np(f["theta"]),
             cs=rng.uniform(0.01, 0.2, lay.gshape))
    tau = (to_np(f["tau_i"]), to_np(f["tau_j"]))
    return lay, np_mesh(mesh), params, x, tau, rng


def _dot(a, b):
    return float(np.sum(a * b))


PERS = [(0, 0, 0), (2, 0, 1), (1, 1, 1), (2, 2, 2), (1, 2, 0)]


@pytest.mark.parametrize("periodic", PERS)
@pytest.mark.parametrize("scheme", [C.DIV_CENTRAL, C.DIV_QUICK])
def test_vjp_dot_product_vs_tangent_oracle(mods, periodic, scheme):
    rhs, _ = mods
    lay, nm, params, x, tau, rng = _setup(periodic, 5, scheme)
    info = lay.info()
    po = dict(params, tau_wall_i=tau[0].ctypes.data, tau_wall_j=tau[1].ctypes.data)
    rbar = rng.standard_normal((lay.mz, lay.my, lay.mx, 3))
    # GPU adjoint
    ctx = rhs.RhsContext(info)
    ti, tj = to_dev(tau[0]), to_dev(tau[1])
    pg = dict(params, tau_wall_i=ti.data_ptr(), tau_wall_j=tj.data_ptr())
    dm = {k: to_dev(v) for k, v in nm.items()}
    xd = {k: to_dev(v) for k, v in x.items()}
    bars = {k: torch.zeros_like(v) for k, v in xd.items()}
    ctx.rhs_vjp(dm, pg, xd["ucont"], xd["ucat"], xd["theta"], xd["cs"], to_dev(rbar), 0.7, 1.3,
                bars["ucont"], bars["ucat"], bars["theta"], bars["cs"])
    torch.cuda.synchronize()
    bars = {k: to_np(v) for k, v in bars.items()}
    print(f"\nVJP periodic={periodic} scheme={scheme}", flush=True)
    for name in ("ucont", "ucat", "theta", "cs"):
        d = rng.standard_normal(x[name].shape)
        _, dR = O.rhs_jvp(info, nm, po, x["ucont"], x["ucat"], x["theta"], x["cs"], scale=0.7, buoy_scale=1.3,
                          **{"d_" + name: d})
        lhs, rhs_ = _dot(rbar, dR), _dot(bars[name], d)
        scale_ = np.sum(np.abs(rbar * dR)) + 1e-300
        err = abs(lhs - rhs_) / scale_
        print(f"  {name:6s} <rbar,Jd>={lhs:+.12e} <J^T rbar,d>={rhs_:+.12e} rel={err:.2e}", flush=True)
        assert err <= 1e-10, (name, err)


def _gpu_forward(ctx, dm, pg, xd, scale, bscale):
    info = ctx.info
    nut = torch.zeros_like(xd["cs"])
    ctx.UpdateNut(dm, pg, xd["ucat"], xd["cs"], nut)
    r = torch.zeros((info.mz, info.my, info.mx, 3), dtype=torch.float64, device="cuda")
    ctx.FormU(dm, pg, xd["ucont"], xd["ucat"], nut, r, scale, 0)
    bt = torch.zeros_like(r)
    ctx.Buoyancy(dm, pg, xd["theta"], bscale, bt)
    return r + bt


@pytest.mark.parametrize("periodic", [(2, 0, 1), (0, 0, 0)])
def test_vjp_finite_differences(mods, periodic):
    rhs, _ = mods
    lay, nm, params, x, tau, rng = _setup(periodic, 6, C.DIV_CENTRAL)
    ctx = rhs.RhsContext(lay.info())
    ti, tj = to_dev(tau[0]), to_dev(tau[1])
    pg = dict(params, tau_wall_i=ti.data_ptr(), tau_wall_j=tj.data_ptr())
    dm = {k: to_dev(v) for k, v in nm.items()}
    xd = {k: to_dev(v) for k, v in x.items()}
    rbar = to_dev(rng.standard_normal((lay.mz, lay.my, lay.mx, 3)))
    bars = {k: torch.zeros_like(v) for k, v in xd.items()}
    ctx.rhs_vjp(dm, pg, xd["ucont"], xd["ucat"], xd["theta"], xd["cs"], rbar, 1.0, 1.0, bars["ucont"], bars["ucat"],
                bars["theta"], bars["cs"])
    for name in ("ucont", "ucat", "theta", "cs"):
        d = to_dev(rng.standard_normal(x[name].shape))
        eps = 1e-6 * max(1.0, float(xd[name].abs().max()))
        xp, xm = dict(xd), dict(xd)
        xp[name] = xd[name] + eps * d
        xm[name] = xd[name] - eps * d
        fd = (_gpu_forward(ctx, dm, pg, xp, 1.0, 1.0) - _gpu_forward(ctx, dm, pg, xm, 1.0, 1.0)) / (2 * eps)
        lhs, rhs_ = float((rbar * fd).sum()), float((bars[name] * d).sum())
        err = abs(lhs - rhs_) / (float((rbar * fd).abs().sum()) + 1e-300)
        print(f"  FD {name:6s} rel={err:.2e}", flush=True)
        assert err <= 1e-6, (name, err)


def test_autograd_gradcheck(mods):
    rhs, autograd = mods
    lay, nm, params, x, tau, rng = _setup((2, 0, 1), 7, C.DIV_CENTRAL, n=(4, 4, 3))
    ctx = rhs.RhsContext(lay.info())
    ti, tj = to_dev(tau[0]), to_dev(tau[1])
    pg = dict(params, tau_wall_i=ti.data_ptr(), tau_wall_j=tj.data_ptr())
    dm = {k: to_dev(v) for k, v in nm.items()}
    ins = tuple(to_dev(x[k]).requires_grad_(True) for k in ("ucont", "ucat", "theta", "cs"))
    fn = lambda *a: autograd.rhs_function(ctx, dm, pg, *a, scale=0.9, buoy_scale=1.1)
    assert torch.autograd.gradcheck(fn, ins, eps=1e-6, atol=1e-7, rtol=1e-5, nondet_tol=1e-12)
