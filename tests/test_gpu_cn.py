"""GPU parity of the Crank-Nicolson momentum residual (tosca_cn_rhs, tosca_snes_func_eval) against
the CPU oracle's restatement, on both kernel paths, first step and later steps, with the wall,
slip, zeroGradient and periodic patch kinds.  The constant right-hand side is a sequence of AXPYs
evaluated in the same order on both sides (bit-identical); the residual carries FormU's round-off:
bar 1e-12 relative L2.  The work arrays (lUcont, lUcat) are compared too."""
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


@pytest.mark.parametrize("ab2", [0, 1])
@pytest.mark.parametrize("periodic", [(0, 0, 0), (2, 0, 1), (1, 1, 1), (2, 2, 2)])
def test_snes_residual(rhs_mod, periodic, ab2):
    seed = 80 + 2 * [(0, 0, 0), (2, 0, 1), (1, 1, 1), (2, 2, 2)].index(periodic) + ab2
    lay, mesh, f, params = random_case(n=(7, 6, 5), periodic=periodic, seed=seed, div_scheme=C.DIV_QUICK,
                                       bcu_nonper=KINDS)
    print(f"\nSNESFuncEval periodic={periodic} ab2={ab2} bc_u={params['bc_u']}", flush=True)
    rng = np.random.Generator(np.random.PCG64(990 + seed))
    glob = (lay.mz, lay.my, lay.mx, 3)
    dp, bt, bto, ro, farm, u = (rng.standard_normal(glob) for _ in range(6))
    uo = C.owned(f["ucont"], lay).numpy().copy()
    nut = to_np(f["nut"])
    tau_i, tau_j = to_np(f["tau_i"]), to_np(f["tau_j"])
    po = dict(params, tau_wall_i=tau_i.ctypes.data, tau_wall_j=tau_j.ctypes.data)
    dt = 0.07
    bU_o = np.zeros(glob)
    O.cn_rhs(lay.info(), po, dp, bt, bto, ab2, ro, farm, bU_o)
    uc_o, ua_o, F_o = to_np(f["ucont"]), to_np(f["ucat"]), np.full(glob, 3.0)
    O.snes_func_eval(lay.info(), np_mesh(mesh), po, u, uc_o, ua_o, nut, bU_o, uo, dt, ab2, F_o)
    ctx = rhs_mod.RhsContext(lay.info())
    dm = {k: to_dev(v) for k, v in np_mesh(mesh).items()}
    ti, tj = to_dev(tau_i), to_dev(tau_j)
    pg = dict(params, tau_wall_i=ti.data_ptr(), tau_wall_j=tj.data_ptr())
    bU_g = to_dev(np.zeros(glob))
    ctx.CNRhs(pg, to_dev(dp), to_dev(bt), to_dev(bto), ab2, to_dev(ro), to_dev(farm), bU_g)
    uc_g, ua_g, F_g = to_dev(to_np(f["ucont"])), to_dev(to_np(f["ucat"])), to_dev(np.full(glob, 3.0))
    ctx.SNESFuncEval(dm, pg, to_dev(u), uc_g, ua_g, to_dev(nut), bU_g, to_dev(uo), dt, ab2, F_g)
    torch.cuda.synchronize()
    assert np.array_equal(to_np(bU_g), bU_o)
    assert np.array_equal(to_np(uc_g), uc_o)
    e_u, e_f = relerr(to_np(ua_g), ua_o), relerr(to_np(F_g), F_o)
    print(f"  ucat relL2={e_u:.2e}  F relL2={e_f:.2e}", flush=True)
    assert e_u <= 1e-12 and e_f <= 1e-12
    ctx.close()
