"""GPU parity of the explicit momentum predictors and the IMEX operator (tosca_ueqn_fe,
tosca_ueqn_rk4, tosca_imex_matvec) against the CPU oracle's restatements, on both kernel paths.
Every output is compared: the new ucont (local, ghosts included), ucat as the last stage leaves
it, rhs, the RK4 accumulator and the IMEX product.  Bar: normwise relative L2 <= 1e-12 per array;
ucont after four RK4 stages carries the FormU round-off of each stage."""
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


def _report(name, a, b):
    e = relerr(a, b)
    print(f"  {name:7s} relL2={e:.3e} identical={np.array_equal(a, b)}", flush=True)
    return e


CASES = [
    # periodic, scheme, form_mode, with bTheta (AB2), with farm
    ((0, 0, 0), C.DIV_CENTRAL, 0, True, True),
    ((2, 0, 1), C.DIV_QUICK, 0, True, False),
    ((1, 1, 1), C.DIV_CENTRAL, 0, False, True),
    ((2, 2, 2), C.DIV_QUICK, 2, True, False),
    ((1, 2, 0), C.DIV_CENTRAL, 1, False, False),
]


def _setup(case_idx):
    periodic, scheme, mode, with_bt, with_farm = CASES[case_idx]
    seed = 61 + case_idx
    lay, mesh, f, params = random_case(n=(7, 6, 5), periodic=periodic, seed=seed, div_scheme=scheme,
                                       bcu_nonper=KINDS)
    rng = np.random.Generator(np.random.PCG64(800 + seed))
    glob = (lay.mz, lay.my, lay.mx, 3)
    src = {k: 0.1 * rng.standard_normal(glob) for k in ("dp", "bt", "bto", "farm")}
    src["ucont_o"] = C.owned(f["ucont"], lay).numpy().copy()
    nm = np_mesh(mesh)
    tau_i, tau_j = to_np(f["tau_i"]), to_np(f["tau_j"])
    return periodic, mode, with_bt, with_farm, lay, nm, f, params, src, tau_i, tau_j


@pytest.mark.parametrize("driver", ["fe", "rk4"])
@pytest.mark.parametrize("case_idx", range(len(CASES)))
def test_explicit_predictors(rhs_mod, case_idx, driver):
    periodic, mode, with_bt, with_farm, lay, nm, f, params, src, tau_i, tau_j = _setup(case_idx)
    print(f"\nUeqn{driver.upper()} periodic={periodic} mode={mode} bTheta={with_bt} farm={with_farm} "
          f"bc_u={params['bc_u']}", flush=True)
    glob = (lay.mz, lay.my, lay.mx, 3)
    bt, bto = (src["bt"], src["bto"]) if with_bt else (None, None)
    farm = src["farm"] if with_farm else None
    dt = 0.05
    uc_o, ua_o = to_np(f["ucont"]), to_np(f["ucat"])
    rhs_o, ut_o = np.full(glob, 7.0), np.full(glob, 5.0)
    po = dict(params, tau_wall_i=tau_i.ctypes.data, tau_wall_j=tau_j.ctypes.data)
    nut = to_np(f["nut"])
    if driver == "fe":
        O.ueqn_fe(lay.info(), nm, po, uc_o, ua_o, nut, src["dp"], bt, bto, 1, farm, src["ucont_o"], dt, rhs_o, mode)
    else:
        O.ueqn_rk4(lay.info(), nm, po, uc_o, ua_o, nut, src["dp"], bt, bto, 1, farm, src["ucont_o"], dt, rhs_o, ut_o,
                   mode)
    ctx = rhs_mod.RhsContext(lay.info())
    dm = {k: to_dev(v) for k, v in nm.items()}
    ti, tj = to_dev(tau_i), to_dev(tau_j)
    pg = dict(params, tau_wall_i=ti.data_ptr(), tau_wall_j=tj.data_ptr())
    dev = lambda a: None if a is None else to_dev(a)
    uc_g, ua_g = to_dev(to_np(f["ucont"])), to_dev(to_np(f["ucat"]))
    rhs_g, ut_g = to_dev(np.full(glob, 7.0)), to_dev(np.full(glob, 5.0))
    if driver == "fe":
        ctx.UeqnFE(dm, pg, uc_g, ua_g, to_dev(nut), to_dev(src["dp"]), dev(bt), dev(bto), 1, dev(farm),
                   to_dev(src["ucont_o"]), dt, rhs_g, mode)
    else:
        ctx.UeqnRK4(dm, pg, uc_g, ua_g, to_dev(nut), to_dev(src["dp"]), dev(bt), dev(bto), 1, dev(farm),
                    to_dev(src["ucont_o"]), dt, rhs_g, ut_g, mode)
    torch.cuda.synchronize()
    assert _report("ucont", to_np(uc_g), uc_o) <= TOL
    assert _report("ucat", to_np(ua_g), ua_o) <= TOL
    assert _report("rhs", to_np(rhs_g), rhs_o) <= TOL
    if driver == "rk4":
        assert _report("utmp", to_np(ut_g), ut_o) <= TOL
    ctx.close()


@pytest.mark.parametrize("case_idx", range(len(CASES)))
def test_imex_matvec(rhs_mod, case_idx):
    periodic, mode, with_bt, with_farm, lay, nm, f, params, src, tau_i, tau_j = _setup(case_idx)
    print(f"\nIMEXMatVec periodic={periodic} bc_u={params['bc_u']}", flush=True)
    glob = (lay.mz, lay.my, lay.mx, 3)
    rng = np.random.Generator(np.random.PCG64(77 + case_idx))
    v = rng.standard_normal(glob)
    coef = 0.37
    uc_o, ua_o = to_np(f["ucont"]), to_np(f["ucat"])
    av_o = np.full(glob, 3.0)
    po = dict(params, tau_wall_i=tau_i.ctypes.data, tau_wall_j=tau_j.ctypes.data)
    nut = to_np(f["nut"])
    O.imex_matvec(lay.info(), nm, po, v, uc_o, ua_o, nut, src["ucont_o"], coef, av_o)
    ctx = rhs_mod.RhsContext(lay.info())
    dm = {k: to_dev(v_) for k, v_ in nm.items()}
    ti, tj = to_dev(tau_i), to_dev(tau_j)
    pg = dict(params, tau_wall_i=ti.data_ptr(), tau_wall_j=tj.data_ptr())
    uc_g, ua_g = to_dev(to_np(f["ucont"])), to_dev(to_np(f["ucat"]))
    av_g = to_dev(np.full(glob, 3.0))
    ctx.IMEXMatVec(dm, pg, to_dev(v), uc_g, ua_g, to_dev(nut), to_dev(src["ucont_o"]), coef, av_g)
    torch.cuda.synchronize()
    assert np.array_equal(to_np(uc_g), uc_o), "ucont"
    assert _report("ucat", to_np(ua_g), ua_o) <= TOL
    assert _report("av", to_np(av_g), av_o) <= TOL
    ctx.close()


def test_rk4_rejects_aliased_work(rhs_mod):
    periodic, mode, with_bt, with_farm, lay, nm, f, params, src, tau_i, tau_j = _setup(2)
    from tosca_amd.abi import ToscaError
    glob = (lay.mz, lay.my, lay.mx, 3)
    ctx = rhs_mod.RhsContext(lay.info())
    dm = {k: to_dev(v) for k, v in nm.items()}
    rhs = to_dev(np.zeros(glob))
    with pytest.raises(ToscaError):
        ctx.UeqnRK4(dm, params, to_dev(to_np(f["ucont"])), to_dev(to_np(f["ucat"])), to_dev(to_np(f["nut"])),
                    to_dev(src["dp"]), None, None, 0, None, to_dev(src["ucont_o"]), 0.1, rhs, rhs, 0)
    ctx.close()
