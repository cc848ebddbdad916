"""GPU parity of ExplicitRhsU (tosca_explicit_rhs_u) against the CPU oracle's restatement.

Every output is compared: ucont after the no-penetration and periodic resets, ucat from the
contravariant fluxes (interior cells, periodic cells and ghosts; the other ghost cells must keep
their input values), and rhs = -dP + bTheta terms + farm + FormU with the non-resolved faces reset.
Bar: normwise relative L2 <= 1e-12 per array (the resets are exact selections and copies).
"""
import numpy as np
import pytest
import torch

from tosca_amd import case as C
from oracle import oracle as O
from tests._util import np_mesh, to_np, to_dev, relerr, maxrel, random_case

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
    e, m = relerr(a, b), maxrel(a, b)
    print(f"  {name:7s} relL2={e:.3e} maxrel={m:.3e}", flush=True)
    return e


def _sources(lay, seed):
    rng = np.random.Generator(np.random.PCG64(500 + seed))
    glob = (lay.mz, lay.my, lay.mx, 3)
    return {k: rng.standard_normal(glob) for k in ("dp", "bt", "bto", "farm", "ucont_o")}


CASES = [
    # periodic, scheme, form_mode, ab2, with bTheta, with farm
    ((0, 0, 0), C.DIV_CENTRAL, 0, 1, True, True),
    ((0, 0, 0), C.DIV_QUICK, 0, 0, True, False),
    ((2, 0, 1), C.DIV_QUICK, 0, 1, True, True),
    ((1, 1, 1), C.DIV_CENTRAL, 1, 1, False, True),
    ((2, 2, 2), C.DIV_QUICK, 2, 0, True, False),
    ((0, 2, 0), C.DIV_CENTRAL, 0, 1, False, False),
    ((1, 0, 2), C.DIV_QUICK, 0, 1, True, True),
]


@pytest.mark.parametrize("case_idx", range(len(CASES)))
def test_explicit_rhs_u_random(rhs_mod, case_idx):
    periodic, scheme, mode, ab2, with_bt, with_farm = CASES[case_idx]
    seed = 31 + case_idx
    lay, mesh, f, params = random_case(n=(7, 6, 5), periodic=periodic, seed=seed, div_scheme=scheme,
                                       bcu_nonper=KINDS)
    print(f"\nExplicitRhsU periodic={periodic} scheme={scheme} mode={mode} ab2={ab2} bTheta={with_bt} "
          f"farm={with_farm} bc_u={params['bc_u']}", flush=True)
    src = _sources(lay, seed)
    nm = np_mesh(mesh)
    tau_i, tau_j = to_np(f["tau_i"]), to_np(f["tau_j"])
    bt = src["bt"] if with_bt else None
    bto = src["bto"] if with_bt else None
    farm = src["farm"] if with_farm else None
    # oracle
    uc_o, ua_o = to_np(f["ucont"]), to_np(f["ucat"])
    rhs_o = np.full((lay.mz, lay.my, lay.mx, 3), 7.0)
    po = dict(params, tau_wall_i=tau_i.ctypes.data, tau_wall_j=tau_j.ctypes.data)
    O.explicit_rhs_u(lay.info(), nm, po, uc_o, ua_o, to_np(f["nut"]), src["dp"], bt, bto, ab2, farm, src["ucont_o"],
                     rhs_o, mode)
    # library
    ctx = rhs_mod.RhsContext(lay.info())
    dm = {k: to_dev(v) for k, v in nm.items()}
    ti, tj = to_dev(tau_i), to_dev(tau_j)
    pg = dict(params, tau_wall_i=ti.data_ptr(), tau_wall_j=tj.data_ptr())
    uc_g, ua_g = to_dev(to_np(f["ucont"])), to_dev(to_np(f["ucat"]))
    rhs_g = to_dev(np.full((lay.mz, lay.my, lay.mx, 3), 7.0))
    dev = lambda a: None if a is None else to_dev(a)
    ctx.ExplicitRhsU(dm, pg, uc_g, ua_g, to_dev(to_np(f["nut"])), to_dev(src["dp"]), dev(bt), dev(bto), ab2, dev(farm),
                     to_dev(src["ucont_o"]), rhs_g, mode)
    torch.cuda.synchronize()
    assert np.array_equal(to_np(uc_g), uc_o), "ucont resets"
    assert _report("ucat", to_np(ua_g), ua_o) <= TOL
    assert _report("rhs", to_np(rhs_g), rhs_o) <= TOL
    ctx.close()


def test_explicit_rhs_u_tgv(rhs_mod):
    """Taylor-Green box (type-2 periodic on every axis): the assembled rhs with the AB2 buoyancy
    terms, on the builder's consistent fields."""
    case = C.tgv_case(n=16, ptype=C.PERIODIC_WRAP)
    lay = case.lay
    print(f"\nExplicitRhsU tgv {lay}", flush=True)
    src = _sources(lay, 99)
    nm = np_mesh(case.mesh)
    nut = np.zeros(lay.gshape)
    cs = np.zeros(lay.gshape)
    O.update_cs(lay.info(), nm, case.params, to_np(case.ucat), cs)
    O.update_nut(lay.info(), nm, case.params, to_np(case.ucat), cs, nut)
    uc_o, ua_o = to_np(case.ucont), to_np(case.ucat)
    rhs_o = np.zeros((lay.mz, lay.my, lay.mx, 3))
    O.explicit_rhs_u(lay.info(), nm, case.params, uc_o, ua_o, nut, src["dp"], src["bt"], src["bto"], 1, None,
                     src["ucont_o"], rhs_o, 0)
    ctx = rhs_mod.RhsContext(lay.info())
    dm = {k: to_dev(v) for k, v in nm.items()}
    uc_g, ua_g = to_dev(to_np(case.ucont)), to_dev(to_np(case.ucat))
    rhs_g = torch.zeros((lay.mz, lay.my, lay.mx, 3), dtype=torch.float64, device="cuda")
    ctx.ExplicitRhsU(dm, case.params, uc_g, ua_g, to_dev(nut), to_dev(src["dp"]), to_dev(src["bt"]),
                     to_dev(src["bto"]), 1, None, to_dev(src["ucont_o"]), rhs_g, 0)
    torch.cuda.synchronize()
    assert np.array_equal(to_np(uc_g), uc_o)
    assert _report("ucat", to_np(ua_g), ua_o) <= TOL
    assert _report("rhs", to_np(rhs_g), rhs_o) <= TOL
    ctx.close()


def test_explicit_rhs_u_rejects_missing_arrays(rhs_mod):
    lay, mesh, f, params = random_case(n=(5, 5, 5), periodic=(0, 0, 0), seed=3, bcu_nonper=(C.BCU_GENERIC,))
    ctx = rhs_mod.RhsContext(lay.info())
    dm = {k: to_dev(v) for k, v in np_mesh(mesh).items()}
    z = torch.zeros((lay.mz, lay.my, lay.mx, 3), dtype=torch.float64, device="cuda")
    from tosca_amd.abi import ToscaError
    with pytest.raises(ToscaError):       # ab2 without bTheta_o
        ctx.ExplicitRhsU(dm, params, to_dev(to_np(f["ucont"])), to_dev(to_np(f["ucat"])), to_dev(to_np(f["nut"])),
                         z, z.clone(), None, 1, None, z.clone(), z.clone(), 0)
    with pytest.raises(ToscaError):       # no dP
        ctx.ExplicitRhsU(dm, params, to_dev(to_np(f["ucont"])), to_dev(to_np(f["ucat"])), to_dev(to_np(f["nut"])),
                         None, None, None, 0, None, z.clone(), z.clone(), 0)
    ctx.close()
