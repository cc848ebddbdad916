"""GPU parity of FormT (tosca_formt, src/solvers/teqn_solvers.c:6-592) against the CPU oracle's
restatement, on seeded random curvilinear meshes with every periodic type per axis, without LES and
with the closures FormT distinguishes (subgrid k_t of DSM / DPASD / AMD, nu_t over the stability-
corrected turbulent Prandtl number of Smagorinsky), thetaWallFunction patches and all form modes.
The output array is pre-filled with random values on both sides, so the boundary cells FormT must
leave alone are compared too.  Both sides evaluate without FP contraction; the only differences are
the last-ulp ones of pow(): bar 1e-12 relative L2."""
import numpy as np
import pytest
import torch

from tosca_amd import case as C
from oracle import oracle as O
from tests._util import np_mesh, to_np, to_dev, relerr, random_nodes

pytestmark = pytest.mark.gpu

TOL = 1e-12


@pytest.fixture(scope="module")
def rhs_mod():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from tosca_amd import rhs
    return rhs


def _case(n, periodic, seed, les, model, walls):
    rng = np.random.Generator(np.random.PCG64(900 + seed))
    lay, mesh = C.mesh_metrics(random_nodes(n, seed), periodic)
    gs = lay.gshape
    bcs = [C.BC_PERIODIC if periodic[a // 2] else C.BC_FIXED for a in range(6)]
    bcn = [C.BC_PERIODIC if periodic[a // 2] else C.BC_ZEROGRAD for a in range(6)]
    params = C.default_params(periodic, bcs, bcn, les_model=model, nu=0.013, t_ref=300.0, les_active=les)
    # nu_t spans both branches of the stability length (small nu_t: l < delta in stable faces)
    nut = rng.uniform(0.0, 0.1, gs) * np.where(rng.uniform(size=gs) < 0.5, 1e-3, 1.0)
    f = dict(ucont=rng.standard_normal(gs + (3,)), theta=300.0 + rng.standard_normal(gs), nut=nut,
             kt=rng.uniform(0.0, 0.05, gs), rhs=rng.standard_normal((lay.mz, lay.my, lay.mx)))
    wall = tuple(int(walls and not periodic[q // 2] and rng.uniform() < 0.7) for q in range(4))
    tp = dict(pr=0.71, wall_t=wall, q_wall_i=rng.standard_normal((3, lay.mz, lay.my)),
              q_wall_j=rng.standard_normal((3, lay.mz, lay.mx)))
    return lay, np_mesh(mesh), params, f, tp


def _compare(rhs_mod, lay, mesh, params, f, tp, scale, mode, label):
    ro = f["rhs"].copy()
    O.formt(lay.info(), mesh, params, tp, f["ucont"], f["theta"], f["nut"], f["kt"], ro, scale, mode)
    ctx = rhs_mod.RhsContext(lay.info())
    rg = to_dev(f["rhs"])
    dm = {k: to_dev(v) for k, v in mesh.items()}
    tpg = dict(tp, q_wall_i=to_dev(tp["q_wall_i"]), q_wall_j=to_dev(tp["q_wall_j"]))
    ctx.FormT(dm, params, tpg, to_dev(f["ucont"]), to_dev(f["theta"]), to_dev(f["nut"]), to_dev(f["kt"]), rg, scale,
              mode)
    torch.cuda.synchronize()
    g = to_np(rg)
    e = relerr(g, ro)
    print(f"  {label}: relL2={e:.2e} identical={np.array_equal(g, ro)}", flush=True)
    ctx.close()
    return e


PERS = [(0, 0, 0), (1, 1, 1), (2, 2, 2), (2, 0, 1), (0, 1, 2), (1, 2, 0)]
CLOSURES = [(0, C.LES_DSM, "no LES"), (1, C.LES_SMAGORINSKY, "smagorinsky Prt"), (1, C.LES_DSM, "dsm k_t"),
            (1, C.LES_DPASD, "dpasd k_t"), (1, C.LES_AMD, "amd k_t")]


@pytest.mark.parametrize("periodic", PERS)
@pytest.mark.parametrize("closure", CLOSURES, ids=[c[2] for c in CLOSURES])
def test_formt_parity(rhs_mod, periodic, closure):
    les, model, name = closure
    seed = PERS.index(periodic) * 10 + CLOSURES.index(closure)
    lay, mesh, params, f, tp = _case((7, 6, 5), periodic, seed, les, model, walls=True)
    print(f"\nFormT periodic={periodic} {name} wall_t={tp['wall_t']}", flush=True)
    assert _compare(rhs_mod, lay, mesh, params, f, tp, 1.0, 0, "mode 0") <= TOL


@pytest.mark.parametrize("mode", [1, 2])
def test_formt_form_modes(rhs_mod, mode):
    lay, mesh, params, f, tp = _case((7, 6, 5), (2, 0, 1), 77 + mode, 1, C.LES_SMAGORINSKY, walls=True)
    print(f"\nFormT formMode={mode}", flush=True)
    assert _compare(rhs_mod, lay, mesh, params, f, tp, -0.75, mode, f"mode {mode}, scale -0.75") <= TOL


@pytest.mark.parametrize("periodic", [(2, 0, 2), (0, 0, 0), (1, 2, 1)])
def test_formt_larger_mesh(rhs_mod, periodic):
    """several workgroups along every axis"""
    lay, mesh, params, f, tp = _case((70, 37, 23), periodic, 5, 1, C.LES_SMAGORINSKY, walls=True)
    print(f"\nFormT 70x37x23 periodic={periodic}", flush=True)
    assert _compare(rhs_mod, lay, mesh, params, f, tp, 1.0, 0, "mode 0") <= TOL


def test_formt_rejects_bad_arguments(rhs_mod):
    from tosca_amd.abi import ToscaError
    lay, mesh, params, f, tp = _case((7, 6, 5), (0, 0, 0), 3, 1, C.LES_DSM, walls=False)
    ctx = rhs_mod.RhsContext(lay.info())
    dm = {k: to_dev(v) for k, v in mesh.items()}
    args = [to_dev(f["ucont"]), to_dev(f["theta"]), to_dev(f["nut"]), to_dev(f["kt"]), to_dev(f["rhs"])]
    for model in (C.LES_STABILITY, C.LES_BV, C.LES_BAMD):
        with pytest.raises(ToscaError):
            ctx.FormT(dm, dict(params, les_model=model), dict(tp, q_wall_i=None, q_wall_j=None), *args)
    with pytest.raises(ToscaError):       # a wall patch without its heat-flux array
        ctx.FormT(dm, params, dict(tp, wall_t=(0, 0, 1, 0), q_wall_i=None, q_wall_j=None), *args)
    with pytest.raises((ToscaError, ValueError)):       # host memory
        ctx.FormT(dm, params, dict(tp, q_wall_i=None, q_wall_j=None), args[0], torch.as_tensor(f["theta"]), *args[2:])
    with pytest.raises(ToscaError):
        ctx.FormT(dm, params, dict(tp, q_wall_i=None, q_wall_j=None), *args, 1.0, 3)
    ctx.close()
