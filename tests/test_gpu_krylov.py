"""GPU parity of the IMEX solve (tosca_imex_solve: restarted GMRES on the device with IMEXMatVec as
the operator) and of the IMEX predictors (tosca_ueqn_rk3cn: RK3WCN and RK3SOCN; tosca_ueqn_abcn)
against the CPU oracle's restatements, on both kernel paths.

The GPU sums its dot products in another (fixed) association than the CPU, so the Krylov iterates
agree to round-off, not bit for bit: the checks are the iteration count, the residual estimate and,
where the operator is linear, the true residual of the GPU solution (evaluated with the oracle's
IMEX operator) against atol, and the solutions to 1e-9 relative.  The predictors are compared at
the reference's settings (restart 30, atol 1e-5) and at a tight atol."""
import numpy as np
import pytest
import torch

from tosca_amd import case as C
from oracle import oracle as O
from tests._util import np_mesh, to_np, to_dev, relerr, random_case

pytestmark = pytest.mark.gpu

KINDS = (C.BCU_GENERIC, C.BCU_SLIP, C.BCU_WALLFUNCTION, C.BCU_NOSLIP, C.BCU_ZEROGRADIENT)


@pytest.fixture(scope="module", params=["fused", "ref"])
def rhs_mod(request):
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


PERS = [(0, 0, 0), (2, 0, 1), (1, 1, 1), (2, 2, 2)]


def _setup(periodic, seed, kinds=KINDS):
    lay, mesh, f, params = random_case(n=(7, 6, 5), periodic=periodic, seed=seed, div_scheme=C.DIV_CENTRAL,
                                       bcu_nonper=kinds)
    rng = np.random.Generator(np.random.PCG64(950 + seed))
    glob = (lay.mz, lay.my, lay.mx, 3)
    tau_i, tau_j = to_np(f["tau_i"]), to_np(f["tau_j"])
    return lay, np_mesh(mesh), f, params, rng, glob, tau_i, tau_j


@pytest.mark.parametrize("ucat0", ["zero", "random"])
@pytest.mark.parametrize("periodic", PERS)
def test_imex_solve(rhs_mod, periodic, ucat0):
    """ucat0 "random": the reference's operator as it stands, affine or piecewise linear: the
    boundary cells of the ucat work array are not rewritten by contravariantToCartesian, wall-
    function faces add the wall stress, zeroGradient patches clip inflow by sign; only the
    comparison with the oracle is meaningful.  ucat0 "zero" with generic, slip and noSlip patches
    only: the operator is linear and the true residual of the solution is checked against atol."""
    seed = 20 + PERS.index(periodic)
    kinds = KINDS if ucat0 == "random" else (C.BCU_GENERIC, C.BCU_SLIP, C.BCU_NOSLIP)
    lay, nm, f, params, rng, glob, tau_i, tau_j = _setup(periodic, seed, kinds)
    if ucat0 == "zero":
        f = dict(f, ucat=torch.zeros_like(f["ucat"]))
    print(f"\nIMEX solve periodic={periodic} ucat0={ucat0} bc_u={params['bc_u']}", flush=True)
    b = rng.standard_normal(glob)
    x0 = b + 0.1 * rng.standard_normal(glob)
    uo = C.owned(f["ucont"], lay).numpy().copy()
    nut = to_np(f["nut"])
    coef, atol = 0.9, 1e-9
    po = dict(params, tau_wall_i=tau_i.ctypes.data, tau_wall_j=tau_j.ctypes.data)
    xo = x0.copy()
    its_o, res_o, r_o = O.imex_solve(lay.info(), nm, po, b, xo, to_np(f["ucont"]), to_np(f["ucat"]), nut, uo, coef,
                                     atol=atol)
    ctx = rhs_mod.RhsContext(lay.info())
    dm = {k: to_dev(v) for k, v in nm.items()}
    ti, tj = to_dev(tau_i), to_dev(tau_j)
    pg = dict(params, tau_wall_i=ti.data_ptr(), tau_wall_j=tj.data_ptr())
    xg = to_dev(x0)
    its_g, res_g, r_g = ctx.IMEXSolve(dm, pg, to_dev(b), xg, to_dev(to_np(f["ucont"])), to_dev(to_np(f["ucat"])),
                                      to_dev(nut), to_dev(uo), coef, atol=atol)
    torch.cuda.synchronize()
    x = to_np(xg)
    ax = np.zeros(glob)
    O.imex_matvec(lay.info(), nm, po, x, to_np(f["ucont"]), to_np(f["ucat"]), nut, uo, coef, ax)
    true = float(np.linalg.norm((ax - b).ravel()))
    e = relerr(x, xo)
    print(f"  its gpu {its_g} oracle {its_o}; residual gpu {res_g:.3e} true {true:.3e}; x relL2 {e:.2e}", flush=True)
    assert r_g > 0 and r_o > 0
    assert abs(its_g - its_o) <= 1
    assert res_g <= atol
    if ucat0 == "zero":
        assert true <= atol * 1.01
    assert e <= 1e-9
    ctx.close()


@pytest.mark.parametrize("periodic", [(0, 0, 0), (2, 0, 1)])
def test_imex_solve_rtol_governs(rhs_mod, periodic):
    """A relative tolerance with atol = 0 and a guess far from the solution: PETSc's rule measures
    rtol against ||b|| (nonzero initial guess), so the GPU and the oracle must stop at the same
    iteration with reason KSP_CONVERGED_RTOL (2); the residual there is <= rtol ||b||."""
    seed = 60 + PERS.index(periodic)
    lay, nm, f, params, rng, glob, tau_i, tau_j = _setup(periodic, seed, (C.BCU_GENERIC, C.BCU_SLIP, C.BCU_NOSLIP))
    f = dict(f, ucat=torch.zeros_like(f["ucat"]))
    print(f"\nIMEX solve, rtol governs, periodic={periodic}", flush=True)
    b = rng.standard_normal(glob)
    x0 = 50.0 * rng.standard_normal(glob)
    uo = C.owned(f["ucont"], lay).numpy().copy()
    nut = to_np(f["nut"])
    coef, rtol = 0.9, 1e-3
    po = dict(params, tau_wall_i=tau_i.ctypes.data, tau_wall_j=tau_j.ctypes.data)
    xo = x0.copy()
    its_o, res_o, r_o = O.imex_solve(lay.info(), nm, po, b, xo, to_np(f["ucont"]), to_np(f["ucat"]), nut, uo, coef,
                                     rtol=rtol, atol=0.0)
    ctx = rhs_mod.RhsContext(lay.info())
    dm = {k: to_dev(v) for k, v in nm.items()}
    ti, tj = to_dev(tau_i), to_dev(tau_j)
    pg = dict(params, tau_wall_i=ti.data_ptr(), tau_wall_j=tj.data_ptr())
    xg = to_dev(x0)
    its_g, res_g, r_g = ctx.IMEXSolve(dm, pg, to_dev(b), xg, to_dev(to_np(f["ucont"])), to_dev(to_np(f["ucat"])),
                                      to_dev(nut), to_dev(uo), coef, rtol=rtol, atol=0.0)
    torch.cuda.synchronize()
    bn = float(np.linalg.norm(b.ravel()))
    print(f"  its gpu {its_g} oracle {its_o}; reason gpu {r_g} oracle {r_o}; res gpu {res_g:.3e} "
          f"rtol*||b|| {rtol * bn:.3e}", flush=True)
    assert r_o == 2 and r_g == 2
    assert its_g == its_o
    assert res_g <= rtol * bn and res_o <= rtol * bn
    ctx.close()


@pytest.mark.parametrize("scheme", [0, 1])
@pytest.mark.parametrize("atol", [1e-5, 1e-11])
@pytest.mark.parametrize("periodic", [(0, 0, 0), (2, 0, 1), (1, 1, 1)])
def test_rk3cn(rhs_mod, periodic, atol, scheme):
    seed = 40 + PERS.index(periodic)
    lay, nm, f, params, rng, glob, tau_i, tau_j = _setup(periodic, seed)
    print(f"\nRK3CN scheme={scheme} atol={atol} periodic={periodic} bc_u={params['bc_u']}", flush=True)
    src = {k: 0.1 * rng.standard_normal(glob) for k in ("dp", "bt", "bto", "farm")}
    uo = C.owned(f["ucont"], lay).numpy().copy()
    nut = to_np(f["nut"])
    dt = 0.05
    po = dict(params, tau_wall_i=tau_i.ctypes.data, tau_wall_j=tau_j.ctypes.data)
    uc_o, ua_o, rhs_o = to_np(f["ucont"]), to_np(f["ucat"]), np.zeros(glob)
    its_o, res_o, r_o = O.ueqn_rk3cn(lay.info(), nm, po, scheme, uc_o, ua_o, nut, src["dp"], src["bt"], src["bto"], 1,
                                     src["farm"], uo, dt, rhs_o, atol=atol)
    ctx = rhs_mod.RhsContext(lay.info())
    dm = {k: to_dev(v) for k, v in nm.items()}
    ti, tj = to_dev(tau_i), to_dev(tau_j)
    pg = dict(params, tau_wall_i=ti.data_ptr(), tau_wall_j=tj.data_ptr())
    uc_g, ua_g, rhs_g = to_dev(to_np(f["ucont"])), to_dev(to_np(f["ucat"])), to_dev(np.zeros(glob))
    its_g, res_g, r_g = ctx.UeqnRK3CN(dm, pg, scheme, uc_g, ua_g, to_dev(nut), to_dev(src["dp"]), to_dev(src["bt"]),
                                      to_dev(src["bto"]), 1, to_dev(src["farm"]), to_dev(uo), dt, rhs_g, atol=atol)
    torch.cuda.synchronize()
    e = relerr(to_np(uc_g), uc_o)
    er = relerr(to_np(rhs_g), rhs_o)
    print(f"  last stage its gpu {its_g} oracle {its_o}; ucont relL2 {e:.2e}; rhs relL2 {er:.2e}", flush=True)
    assert r_g > 0 and r_o > 0
    assert its_g == its_o
    assert e <= 1e-9 and er <= 1e-9
    ctx.close()


@pytest.mark.parametrize("periodic", [(0, 0, 0), (2, 0, 1), (1, 1, 1)])
def test_abcn_two_steps(rhs_mod, periodic):
    """UeqnABCN on a first step (ab2 = 0) and a second one (ab2 = 1) that uses the first's
    RhsConv_o"""
    seed = 70 + PERS.index(periodic)
    lay, nm, f, params, rng, glob, tau_i, tau_j = _setup(periodic, seed)
    print(f"\nABCN periodic={periodic} bc_u={params['bc_u']}", flush=True)
    src = {k: 0.1 * rng.standard_normal(glob) for k in ("dp", "bt", "bto", "farm")}
    uo = C.owned(f["ucont"], lay).numpy().copy()
    nut = to_np(f["nut"])
    dt = 0.05
    po = dict(params, tau_wall_i=tau_i.ctypes.data, tau_wall_j=tau_j.ctypes.data)
    uc_o, ua_o, rhs_o, cv_o = to_np(f["ucont"]), to_np(f["ucat"]), np.zeros(glob), np.zeros(glob)
    ctx = rhs_mod.RhsContext(lay.info())
    dm = {k: to_dev(v) for k, v in nm.items()}
    ti, tj = to_dev(tau_i), to_dev(tau_j)
    pg = dict(params, tau_wall_i=ti.data_ptr(), tau_wall_j=tj.data_ptr())
    uc_g, ua_g, rhs_g, cv_g = (to_dev(to_np(f["ucont"])), to_dev(to_np(f["ucat"])), to_dev(np.zeros(glob)),
                               to_dev(np.zeros(glob)))
    for ab2 in (0, 1):
        its_o, _, r_o = O.ueqn_abcn(lay.info(), nm, po, uc_o, ua_o, nut, src["dp"], src["bt"], src["bto"], ab2,
                                    src["farm"], uo, dt, rhs_o, cv_o)
        its_g, _, r_g = ctx.UeqnABCN(dm, pg, uc_g, ua_g, to_dev(nut), to_dev(src["dp"]), to_dev(src["bt"]),
                                     to_dev(src["bto"]), ab2, to_dev(src["farm"]), to_dev(uo), dt, rhs_g, cv_g)
        torch.cuda.synchronize()
        e, ec = relerr(to_np(uc_g), uc_o), relerr(to_np(cv_g), cv_o)
        print(f"  ab2={ab2}: its gpu {its_g} oracle {its_o}; ucont relL2 {e:.2e}; RhsConv relL2 {ec:.2e}", flush=True)
        assert r_g > 0 and r_o > 0 and its_g == its_o
        assert e <= 1e-9 and ec <= 1e-9
    ctx.close()


def test_krylov_argument_checks(rhs_mod):
    lay, nm, f, params, rng, glob, tau_i, tau_j = _setup((1, 1, 1), 60)
    from tosca_amd.abi import ToscaError
    ctx = rhs_mod.RhsContext(lay.info())
    dm = {k: to_dev(v) for k, v in nm.items()}
    b = to_dev(rng.standard_normal(glob))
    args = (to_dev(to_np(f["ucont"])), to_dev(to_np(f["ucat"])), to_dev(to_np(f["nut"])),
            to_dev(C.owned(f["ucont"], lay).numpy().copy()), 0.5)
    with pytest.raises(ToscaError):
        ctx.IMEXSolve(dm, params, b, b, *args)                 # x aliases b
    with pytest.raises(ToscaError):
        ctx.IMEXSolve(dm, params, b, b.clone(), *args, restart=0)
    ctx.close()
