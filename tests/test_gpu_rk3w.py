"""The explicit Wray Runge-Kutta predictor (tosca_ueqn_rk3w) and its checkpointed adjoint
(tosca_ueqn_rk3w_adjoint) on the GPU, against the oracle.

  - forward: every output (ucont with its ghosts, ucat, rhs, rhs_o, ucont_o) against
    oracle_ueqn_rk3w over 3 and 4 stages (one time step, and the first stage of a second), on both
    kernel paths; bar: normwise relative L2 <= 1e-12 per array;
  - adjoint: the dot-product identity <xbar, J d> = <J^T xbar, d> against the FP64 tangent oracle
    (oracle_rk3w_jvp, pinned to finite differences on the CPU) for random cotangents of the final
    Ucont and random directions in the initial Ucont and in nu_t, over 4 and 10 stages, with walls
    (noSlip, slip, wall function, zeroGradient) and periodic axes of both types, central and QUICK;
    bar 1e-10.  The adjoint call also leaves the forward's outputs, bit for bit.
"""
import numpy as np
import pytest
import torch

from tosca_amd import case as C
from oracle import oracle as O
from tests._util import np_mesh, to_np, to_dev, relerr, random_case

pytestmark = pytest.mark.gpu

TOL = 1e-12
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


CASES = [
    # periodic, scheme, with bTheta (AB2), with farm
    ((0, 0, 0), C.DIV_CENTRAL, True, True),
    ((2, 0, 1), C.DIV_QUICK, True, False),
    ((1, 1, 1), C.DIV_CENTRAL, False, True),
    ((2, 2, 2), C.DIV_QUICK, False, False),
    ((1, 2, 0), C.DIV_CENTRAL, True, False),
]


def _setup(case_idx):
    periodic, scheme, with_bt, with_farm = CASES[case_idx]
    seed = 91 + case_idx
    lay, mesh, f, params = random_case(n=(7, 6, 5), periodic=periodic, seed=seed, div_scheme=scheme, bcu_nonper=KINDS)
    rng = np.random.Generator(np.random.PCG64(900 + seed))
    glob = (lay.mz, lay.my, lay.mx, 3)
    src = {k: 0.1 * rng.standard_normal(glob) for k in ("dp", "bt", "bto", "farm")}
    if not with_bt:
        src["bt"] = src["bto"] = None
    if not with_farm:
        src["farm"] = None
    nm = np_mesh(mesh)
    tau = (to_np(f["tau_i"]), to_np(f["tau_j"]))
    return lay, nm, f, params, src, tau, rng


def _report(name, a, b):
    e = relerr(a, b)
    print(f"  {name:8s} relL2={e:.3e} identical={np.array_equal(a, b)}", flush=True)
    return e


def _dev_args(params, src, tau):
    ti, tj = to_dev(tau[0]), to_dev(tau[1])
    pg = dict(params, tau_wall_i=ti.data_ptr(), tau_wall_j=tj.data_ptr())
    dev = lambda a: None if a is None else to_dev(a)
    return pg, (ti, tj), {k: dev(v) for k, v in src.items()}


@pytest.mark.parametrize("n_stages", [3, 4])
@pytest.mark.parametrize("case_idx", range(len(CASES)))
def test_rk3w_forward(rhs_mod, case_idx, n_stages):
    lay, nm, f, params, src, tau, rng = _setup(case_idx)
    print(f"\nUeqnRK3W {CASES[case_idx]} stages={n_stages} bc_u={params['bc_u']}", flush=True)
    glob = (lay.mz, lay.my, lay.mx, 3)
    dt = 0.05
    po = dict(params, tau_wall_i=tau[0].ctypes.data, tau_wall_j=tau[1].ctypes.data)
    uc_o, ua_o, nut = to_np(f["ucont"]), to_np(f["ucat"]), to_np(f["nut"])
    uo_o, r_o, ro_o = np.full(glob, 1.0), np.full(glob, 2.0), np.full(glob, 3.0)
    O.ueqn_rk3w(lay.info(), nm, po, uc_o, ua_o, nut, src["dp"], src["bt"], src["bto"], 1, src["farm"], dt, n_stages,
                uo_o, r_o, ro_o)
    ctx = rhs_mod.RhsContext(lay.info())
    dm = {k: to_dev(v) for k, v in nm.items()}
    pg, keep, sd = _dev_args(params, src, tau)
    uc, ua = to_dev(to_np(f["ucont"])), to_dev(to_np(f["ucat"]))
    uo, r, ro = to_dev(np.full(glob, 1.0)), to_dev(np.full(glob, 2.0)), to_dev(np.full(glob, 3.0))
    ctx.UeqnRK3W(dm, pg, uc, ua, to_dev(nut), sd["dp"], sd["bt"], sd["bto"], 1, sd["farm"], dt, n_stages, uo, r, ro)
    torch.cuda.synchronize()
    for name, a, b in (("ucont", uc, uc_o), ("ucat", ua, ua_o), ("rhs", r, r_o), ("rhs_o", ro, ro_o),
                       ("ucont_o", uo, uo_o)):
        assert _report(name, to_np(a), b) <= TOL, name
    ctx.close()


@pytest.mark.parametrize("n_stages", [4, 10])
@pytest.mark.parametrize("case_idx", range(len(CASES)))
def test_rk3w_adjoint_vs_tangent_oracle(rhs_mod, case_idx, n_stages):
    lay, nm, f, params, src, tau, rng = _setup(case_idx)
    info = lay.info()
    print(f"\nUeqnRK3W adjoint {CASES[case_idx]} stages={n_stages} bc_u={params['bc_u']}", flush=True)
    glob = (lay.mz, lay.my, lay.mx, 3)
    dt = 0.05
    po = dict(params, tau_wall_i=tau[0].ctypes.data, tau_wall_j=tau[1].ctypes.data)
    uc0, ua0, nut = to_np(f["ucont"]), to_np(f["ucat"]), to_np(f["nut"])
    xbar = rng.standard_normal(glob)
    ctx = rhs_mod.RhsContext(info)
    dm = {k: to_dev(v) for k, v in nm.items()}
    pg, keep, sd = _dev_args(params, src, tau)
    # the forward alone, for the outputs the adjoint call must also leave
    fw = dict(uc=to_dev(uc0), ua=to_dev(ua0), uo=to_dev(np.zeros(glob)), r=to_dev(np.zeros(glob)),
              ro=to_dev(np.zeros(glob)))
    ctx.UeqnRK3W(dm, pg, fw["uc"], fw["ua"], to_dev(nut), sd["dp"], sd["bt"], sd["bto"], 1, sd["farm"], dt, n_stages,
                 fw["uo"], fw["r"], fw["ro"])
    ad = dict(uc=to_dev(uc0), ua=to_dev(ua0), uo=to_dev(np.zeros(glob)), r=to_dev(np.zeros(glob)),
              ro=to_dev(np.zeros(glob)))
    ubar = to_dev(xbar)
    nbar = to_dev(np.full(lay.gshape, 9.0))
    ctx.UeqnRK3W_adjoint(dm, pg, ad["uc"], ad["ua"], to_dev(nut), sd["dp"], sd["bt"], sd["bto"], 1, sd["farm"], dt,
                         n_stages, ad["uo"], ad["r"], ad["ro"], ubar, nbar)
    torch.cuda.synchronize()
    for k in fw:
        assert torch.equal(fw[k], ad[k]), k
    ubar, nbar = to_np(ubar), to_np(nbar)
    du = rng.standard_normal(glob)
    dn = rng.standard_normal(lay.gshape)
    args = (nut, src["dp"], src["bt"], src["bto"], 1, src["farm"], dt, n_stages)
    for name, kw, adj in (("ucont", dict(d_ucont=du), float(np.sum(ubar * du))),
                          ("nut", dict(d_nut=dn), float(np.sum(nbar * dn)))):
        _, dF = O.rk3w_jvp(info, nm, po, uc0, ua0, *args, **kw)
        lhs = float(np.sum(xbar * dF))
        err = abs(lhs - adj) / (float(np.sum(np.abs(xbar * dF))) + 1e-300)
        print(f"  {name:6s} <xbar,J d>={lhs:+.12e} <J^T xbar,d>={adj:+.12e} rel={err:.2e}", flush=True)
        assert err <= 1e-10, (name, err)
    ctx.close()
