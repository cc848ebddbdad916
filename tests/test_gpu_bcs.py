"""GPU parity of the boundary-condition ghost fills (tosca_update_contravariant_bcs,
tosca_update_cartesian_bcs, tosca_update_temperature_bcs) against the CPU oracle's restatement.
The whole local array is compared, so points a fill must leave alone are checked too.  The fills
are copies, reflections and short expressions compiled without contraction: bar 1e-14 relative
(they agree bit for bit in practice)."""
import numpy as np
import pytest
import torch

from tosca_amd import case as C
from oracle import oracle as O
from tests._util import np_mesh, to_np, to_dev, relerr, random_nodes, inflow_inlet

pytestmark = pytest.mark.gpu

TOL = 1e-14


@pytest.fixture(scope="module")
def rhs_mod():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from tosca_amd import rhs
    return rhs


def _case(periodic, seed, kinds):
    rng = np.random.Generator(np.random.PCG64(700 + seed))
    lay, mesh = C.mesh_metrics(random_nodes((7, 6, 5), seed), periodic)
    bc_u = []
    for q in range(6):
        bc_u.append(C.BCU_PERIODIC if periodic[q // 2] else int(rng.choice(kinds)))
    bcn = [C.BC_PERIODIC if periodic[a // 2] else C.BC_ZEROGRAD for a in range(6)]
    params = C.default_params(periodic, [C.BC_PERIODIC if periodic[a // 2] else C.BC_FIXED for a in range(6)], bcn,
                              div_scheme=C.DIV_CENTRAL, les_model=C.LES_DSM, nu=0.01)
    return lay, np_mesh(mesh), dict(params, bc_u=tuple(bc_u)), rng


PERS = [(0, 0, 0), (2, 0, 1), (1, 1, 1), (2, 2, 2), (0, 2, 0), (1, 0, 2)]


@pytest.mark.parametrize("periodic", PERS)
def test_contravariant_bcs(rhs_mod, periodic):
    seed = PERS.index(periodic)
    kinds = (C.BCU_NOSLIP, C.BCU_SLIP, C.BCU_WALLFUNCTION, C.BCU_ZEROGRADIENT, C.BCU_FIXEDVALUE, C.BCU_GENERIC)
    lay, mesh, params, rng = _case(periodic, seed, kinds)
    print(f"\nUpdateContravariantBCs periodic={periodic} bc_u={params['bc_u']}", flush=True)
    V0, U = rng.standard_normal(lay.gshape + (3,)), rng.standard_normal(lay.gshape + (3,))
    Vo = V0.copy()
    O.contravariant_bcs(lay.info(), mesh, params, Vo, U)
    ctx = rhs_mod.RhsContext(lay.info())
    Vg = to_dev(V0)
    ctx.UpdateContravariantBCs({k: to_dev(v) for k, v in mesh.items()}, params, Vg, to_dev(U))
    torch.cuda.synchronize()
    e = relerr(to_np(Vg), Vo)
    print(f"  ucont relL2={e:.2e} identical={np.array_equal(to_np(Vg), Vo)}", flush=True)
    assert e <= TOL
    ctx.close()


@pytest.mark.parametrize("periodic", PERS)
def test_cartesian_bcs_shumann(rhs_mod, periodic):
    """velocityWallFunction patches with the Shumann model on every non-periodic i and j patch
    (the i patches' ghosts stay as they are), the k patches from the other kinds"""
    seed = 30 + PERS.index(periodic)
    lay, mesh, params, rng = _case(periodic, seed, (C.BCU_NOSLIP, C.BCU_SLIP, C.BCU_ZEROGRADIENT, C.BCU_FIXEDVALUE))
    bc_u = tuple(C.BCU_WALLFUNCTION if q < 4 and not periodic[q // 2] else params["bc_u"][q] for q in range(6))
    params = dict(params, bc_u=bc_u, wall_u=(-3, -3, -3, -3))
    print(f"\nUpdateCartesianBCs Shumann periodic={periodic} bc_u={params['bc_u']}", flush=True)
    U0, ufix = rng.standard_normal(lay.gshape + (3,)), rng.standard_normal((6, 3))
    Uo = U0.copy()
    O.cartesian_bcs(lay.info(), mesh, params, Uo, ufix)
    ctx = rhs_mod.RhsContext(lay.info())
    Ug = to_dev(U0)
    ctx.UpdateCartesianBCs({k: to_dev(v) for k, v in mesh.items()}, params, Ug, ufix)
    torch.cuda.synchronize()
    e = relerr(to_np(Ug), Uo)
    print(f"  ucat relL2={e:.2e} identical={np.array_equal(to_np(Ug), Uo)}", flush=True)
    assert e <= TOL
    from tosca_amd.abi import ToscaError
    if not periodic[1]:
        with pytest.raises(ToscaError):
            ctx.UpdateCartesianBCs({k: to_dev(v) for k, v in mesh.items()}, dict(params, wall_u=(-3, -3, 0, -3)), Ug,
                                   ufix)
    ctx.close()


@pytest.mark.parametrize("rough", ["smooth", "rough", "mixed"])
@pytest.mark.parametrize("periodic", PERS)
def test_cartesian_bcs_cabot(rhs_mod, periodic, rough):
    """velocityWallFunction patches with the Cabot model on every non-periodic i and j patch
    (smooth walls through the integration table and the Newton solve, rough walls through the log
    law, or both), the k patches from the other kinds; ucat and ustar compared.  The smooth law's
    Newton iteration stops at steps below 1e-10 and the table's transcendental functions may round
    differently on the two sides: bar 1e-12."""
    seed = 50 + PERS.index(periodic)
    lay, mesh, params, rng = _case(periodic, seed, (C.BCU_NOSLIP, C.BCU_SLIP, C.BCU_ZEROGRADIENT, C.BCU_FIXEDVALUE))
    bc_u = tuple(C.BCU_WALLFUNCTION if q < 4 and not periodic[q // 2] else params["bc_u"][q] for q in range(6))
    ks = {"smooth": (0.0,) * 4, "rough": (0.05, 0.2, 0.01, 1e-12), "mixed": (0.0, 0.3, 0.05, 0.0)}[rough]
    params = dict(params, bc_u=bc_u, wall_u=(-1, -1, -1, -1), wall_rough=ks, nu=2e-3)
    print(f"\nUpdateCartesianBCs Cabot periodic={periodic} roughness={ks} bc_u={params['bc_u']}", flush=True)
    U0, ufix = rng.standard_normal(lay.gshape + (3,)), rng.standard_normal((6, 3))
    S0 = np.full(lay.gshape, -3.0)
    Uo, So = U0.copy(), S0.copy()
    O.cartesian_bcs(lay.info(), mesh, params, Uo, ufix, So)
    ctx = rhs_mod.RhsContext(lay.info())
    Ug, Sg = to_dev(U0), to_dev(S0)
    dm = {k: to_dev(v) for k, v in mesh.items()}
    for call in range(2):       # the second call reuses the context's table
        ctx.UpdateCartesianBCs(dm, params, Ug, ufix, Sg)
        torch.cuda.synchronize()
        eu, es = relerr(to_np(Ug), Uo), relerr(to_np(Sg), So)
        print(f"  call {call}: ucat relL2={eu:.2e} identical={np.array_equal(to_np(Ug), Uo)}; "
              f"ustar relL2={es:.2e} identical={np.array_equal(to_np(Sg), So)}", flush=True)
        assert eu <= 1e-12 and es <= 1e-12
        Ug, Sg = to_dev(U0), to_dev(S0)
    from tosca_amd.abi import ToscaError
    if any(not periodic[a] for a in range(2)):
        with pytest.raises(ToscaError):
            ctx.UpdateCartesianBCs(dm, params, Ug, ufix)                # Cabot patches need ustar
    ctx.close()


@pytest.mark.parametrize("periodic", PERS)
def test_cartesian_bcs(rhs_mod, periodic):
    seed = 10 + PERS.index(periodic)
    kinds = (C.BCU_NOSLIP, C.BCU_SLIP, C.BCU_ZEROGRADIENT, C.BCU_FIXEDVALUE)
    lay, mesh, params, rng = _case(periodic, seed, kinds)
    print(f"\nUpdateCartesianBCs periodic={periodic} bc_u={params['bc_u']}", flush=True)
    U0, ufix = rng.standard_normal(lay.gshape + (3,)), rng.standard_normal((6, 3))
    Uo = U0.copy()
    O.cartesian_bcs(lay.info(), mesh, params, Uo, ufix)
    ctx = rhs_mod.RhsContext(lay.info())
    Ug = to_dev(U0)
    ctx.UpdateCartesianBCs({k: to_dev(v) for k, v in mesh.items()}, params, Ug, ufix)
    torch.cuda.synchronize()
    e = relerr(to_np(Ug), Uo)
    print(f"  ucat relL2={e:.2e} identical={np.array_equal(to_np(Ug), Uo)}", flush=True)
    assert e <= TOL
    ctx.close()


@pytest.mark.parametrize("periodic", PERS)
def test_temperature_bcs(rhs_mod, periodic):
    seed = 20 + PERS.index(periodic)
    lay, mesh, params, rng = _case(periodic, seed, (C.BCU_ZEROGRADIENT,))
    bct = [C.BCT_PERIODIC if periodic[q // 2] else
           int(rng.choice((C.BCT_ZEROGRADIENT, C.BCT_FIXEDVALUE, C.BCT_FIXEDGRADIENT))) for q in range(6)]
    if not periodic[1] and seed % 2 == 0:
        bct[2] = C.BCT_SURFACETEMP           # jLeft with the surface-temperature model
    vals = rng.standard_normal(6)
    print(f"\nUpdateTemperatureBCs periodic={periodic} bc_t={bct}", flush=True)
    t0 = 300.0 + rng.standard_normal(lay.gshape)
    to = t0.copy()
    O.temperature_bcs(lay.info(), mesh, params, bct, vals, to)
    ctx = rhs_mod.RhsContext(lay.info())
    tg = to_dev(t0)
    ctx.UpdateTemperatureBCs({k: to_dev(v) for k, v in mesh.items()}, params, bct, vals, tg)
    torch.cuda.synchronize()
    e = relerr(to_np(tg), to)
    print(f"  theta relL2={e:.2e} identical={np.array_equal(to_np(tg), to)}", flush=True)
    assert e <= TOL
    ctx.close()


def test_cartesian_bcs_rejects_wall_function(rhs_mod):
    lay, mesh, params, rng = _case((0, 0, 0), 40, (C.BCU_ZEROGRADIENT,))
    params = dict(params, bc_u=(C.BCU_WALLFUNCTION,) + tuple(params["bc_u"][1:]))
    ctx = rhs_mod.RhsContext(lay.info())
    from tosca_amd.abi import ToscaError
    with pytest.raises(ToscaError):
        ctx.UpdateCartesianBCs({k: to_dev(v) for k, v in mesh.items()}, params, to_dev(rng.standard_normal(lay.gshape + (3,))))
    ctx.close()


@pytest.mark.parametrize("kind", ["loglaw", "sinusoidal", "nieuwstadt"])
def test_cartesian_bcs_inlet(rhs_mod, kind):
    """kLeft inletFunction profiles (log law, sinusoidal, Nieuwstadt) on a random curvilinear mesh
    with the other patches of the other kinds; the cell centres come from the mesh builder.  log
    and cos may round differently on the two sides: bar 1e-14; the Nieuwstadt profile's complex
    log, exp, sqrt, quotients and its series of up to thousands of terms: bar 1e-12."""
    rng = np.random.Generator(np.random.PCG64(61))
    lay, mesh_t = C.mesh_metrics(random_nodes((7, 6, 5), 61), (0, 0, 0))
    cent = to_np(mesh_t["cent"])
    mesh = np_mesh(mesh_t)
    bcn = [C.BC_ZEROGRAD] * 6
    params = C.default_params((0, 0, 0), [C.BC_FIXED] * 6, bcn, div_scheme=C.DIV_CENTRAL, les_model=C.LES_DSM, nu=0.01)
    params = dict(params, bc_u=(C.BCU_ZEROGRADIENT, C.BCU_SLIP, C.BCU_NOSLIP, C.BCU_SLIP, C.BCU_GENERIC,
                                C.BCU_ZEROGRADIENT))
    own = cent[4:-4, 4:-4, 4:-4]           # interior cells (the builder leaves boundary-cell centres at 0)
    zc = own[..., 2]
    if kind == "loglaw":
        inlet = dict(type=C.INLET_LOGLAW, grnd_level=float(zc.min()) - 0.05, u_tau=0.4, roughness=0.03,
                     h_inv=float(np.median(zc)), udir=(0.6, 0.8, 0.0))
    elif kind == "nieuwstadt":
        # cells below and above the capping height, the lowest within 0.05 of the ground (a series of
        # thousands of terms); Coriolis parameter for C = fc h_inv / (kappa u_tau) = 0.6
        h_inv = float(np.median(zc) - zc.min())
        inlet = dict(type=C.INLET_NIEUWSTADT, grnd_level=float(zc.min()) - 0.05, u_tau=0.4, roughness=0.03,
                     h_inv=h_inv, h_ref=0.3 * h_inv, fc=0.6 * 0.4 * 0.4 / h_inv, udir=(np.cos(0.3), np.sin(0.3), 0.0))
    else:
        yc = own[..., 1]
        inlet = dict(type=C.INLET_SINUSOIDAL, ymin=float(yc.min()), ly=float(np.ptp(yc)), amplitude=0.25, periods=1.5,
                     udir=(0.0, 0.6, 0.8), uref=(1.0, 2.0, 2.0))
    print(f"\nUpdateCartesianBCs inlet {kind} bc_u={params['bc_u']}", flush=True)
    U0 = rng.standard_normal(lay.gshape + (3,))
    Uo = U0.copy()
    O.cartesian_bcs(lay.info(), mesh, params, Uo, None, None, dict(inlet, cent=cent))
    ctx = rhs_mod.RhsContext(lay.info())
    Ug = to_dev(U0)
    ctx.UpdateCartesianBCs({k: to_dev(v) for k, v in mesh.items()}, params, Ug, None, None,
                           dict(inlet, cent=to_dev(cent)))
    torch.cuda.synchronize()
    e = relerr(to_np(Ug), Uo)
    print(f"  ucat relL2={e:.2e} identical={np.array_equal(to_np(Ug), Uo)}", flush=True)
    assert e <= (1e-12 if kind == "nieuwstadt" else TOL)
    if kind == "nieuwstadt":
        ghost = to_np(Ug)[3]                        # the k = 0 plane (ghost width 3), where the profile goes
        print(f"  inlet ghost plane max|relerr|={np.abs(ghost - Uo[3]).max() / np.abs(Uo[3]).max():.2e}", flush=True)
        assert np.abs(ghost - Uo[3]).max() <= 1e-12 * np.abs(Uo[3]).max()
    from tosca_amd.abi import ToscaError
    with pytest.raises(ToscaError):
        ctx.UpdateCartesianBCs({k: to_dev(v) for k, v in mesh.items()}, params, Ug, None, None,
                               dict(inlet, type=7, cent=to_dev(cent)))
    ctx.close()


def _dev_inlet(inlet):
    """The inlet dict with its arrays on the device (int32 tables stay int32)."""
    out = {}
    for k, v in inlet.items():
        if isinstance(v, np.ndarray):
            out[k] = torch.as_tensor(np.ascontiguousarray(v)).cuda()
        else:
            out[k] = v
    return out


@pytest.mark.parametrize("kind", ["powerlaw", "mapped", "mapped_merge", "interp", "interp_spline_shift_merge"])
def test_cartesian_bcs_inflow(rhs_mod, kind):
    """kLeft power-law (caller's noise draws), mapped (periodized plane, with and past its periods)
    and interpolated (4-point stencils; 6-point spline stencils for v and w, the shifted left/right
    mix) inflow ghosts, with and without the blend into the averaged top profile, on a random
    curvilinear mesh with random planes and tables.  pow and tanh may round differently on the two
    sides and the stencil sums may be contracted on the device: bar 1e-14 on the ghost plane."""
    lay, mesh_t = C.mesh_metrics(random_nodes((5, 9, 8), 63), (0, 0, 0))
    cent = to_np(mesh_t["cent"])
    mesh = np_mesh(mesh_t)
    params = C.default_params((0, 0, 0), [C.BC_FIXED] * 6, [C.BC_ZEROGRAD] * 6, div_scheme=C.DIV_CENTRAL,
                              les_model=C.LES_DSM, nu=0.01)
    params = dict(params, bc_u=(C.BCU_ZEROGRADIENT, C.BCU_SLIP, C.BCU_NOSLIP, C.BCU_ZEROGRADIENT, C.BCU_GENERIC,
                                C.BCU_ZEROGRADIENT))
    inlet = inflow_inlet(lay, kind, 5, cent)
    print(f"\nUpdateCartesianBCs inflow {kind} mx={lay.mx} my={lay.my} mz={lay.mz}", flush=True)
    rng = np.random.Generator(np.random.PCG64(64))
    U0 = rng.standard_normal(lay.gshape + (3,))
    Uo = U0.copy()
    O.cartesian_bcs(lay.info(), mesh, params, Uo, None, None, inlet)
    ctx = rhs_mod.RhsContext(lay.info())
    Ug = to_dev(U0)
    ctx.UpdateCartesianBCs({k: to_dev(v) for k, v in mesh.items()}, params, Ug, None, None, _dev_inlet(inlet))
    torch.cuda.synchronize()
    got = to_np(Ug)
    ghost, ref = got[3, 4:lay.my + 2, 4:lay.mx + 2], Uo[3, 4:lay.my + 2, 4:lay.mx + 2]
    e = np.abs(ghost - ref).max() / np.abs(ref).max()
    print(f"  ucat relL2={relerr(got, Uo):.2e} ghost plane max|relerr|={e:.2e}", flush=True)
    assert np.isfinite(ref).all() and np.abs(ref - U0[3, 4:lay.my + 2, 4:lay.mx + 2]).min() > 0
    assert e <= TOL
    assert relerr(got, Uo) <= TOL
    from tosca_amd.abi import ToscaError
    dev = {k: to_dev(v) for k, v in mesh.items()}
    missing = {"powerlaw": dict(noise=None), "mapped": dict(n1=0), "mapped_merge": dict(plane=None),
               "interp": dict(weights=None), "interp_spline_shift_merge": dict(y_ids=None)}[kind]
    with pytest.raises(ToscaError):
        ctx.UpdateCartesianBCs(dev, params, Ug, None, None, _dev_inlet(dict(inlet, **missing)))
    if kind == "powerlaw":
        # an int table needs 4-byte alignment only: the draws one element into a larger array
        di = _dev_inlet(inlet)
        off = torch.zeros(di["noise"].numel() + 1, dtype=torch.int32, device="cuda")[1:]
        off.copy_(di["noise"].reshape(-1))
        assert off.data_ptr() % 8 == 4
        U2 = to_dev(U0)
        ctx.UpdateCartesianBCs(dev, params, U2, None, None, dict(di, noise=off.view(di["noise"].shape)))
        torch.cuda.synchronize()
        assert np.array_equal(to_np(U2), got)
    ctx.close()
