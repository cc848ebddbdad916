"""Pins of the oracle's boundary-condition fills (UpdateContravariantBCs, UpdateCartesianBCs,
UpdateTemperatureBCs subsets) from facts that do not depend on its code:

  - each patch rule, written as a numpy expression on the patch's ghost plane (interior transverse
    range for the cell fills, the loop range k < mz-1, j < my-1, i < mx-1 for the face fill), and
    every other point unchanged;
  - slip: the ghost velocity is the mirror image of the boundary cell's velocity in the patch, and on
    an affine lattice the patch normal is the unit lattice vector of the axis;
  - fixedGradient on a box lattice: (t_ghost - t_cell) / h = g, h the cell width across the patch;
  - interpolated boundary fluxes of a uniform velocity on an affine lattice equal S . U.
"""
import numpy as np
import pytest

from tosca_amd import case as C
from oracle import oracle as O
from tests._util import np_mesh, random_nodes
from tests.test_oracle_pins import AFFINE, affine_nodes, affine_geometry

NOSLIP, ZG, SLIP, WALL, GEN, PER, FIX = (C.BCU_NOSLIP, C.BCU_ZEROGRADIENT, C.BCU_SLIP, C.BCU_WALLFUNCTION,
                                         C.BCU_GENERIC, C.BCU_PERIODIC, C.BCU_FIXEDVALUE)
AXIS = {0: 2, 1: 1, 2: 0}          # computational direction -> array axis of [k][j][i]


def _setup(nodes, periodic, bc_u, seed):
    lay, mesh = C.mesh_metrics(nodes, periodic)
    bcn = [C.BC_PERIODIC if periodic[a // 2] else C.BC_ZEROGRAD for a in range(6)]
    params = C.default_params(periodic, [C.BC_PERIODIC if periodic[a // 2] else C.BC_FIXED for a in range(6)], bcn,
                              div_scheme=C.DIV_CENTRAL, les_model=C.LES_DSM, nu=0.01)
    params = dict(params, bc_u=tuple(bc_u))
    rng = np.random.Generator(np.random.PCG64(seed))
    return lay, np_mesh(mesh), params, rng


def _plane(lay, a, idx):
    """index tuple of the plane idx along direction a (local array), interior transverse range"""
    m = (lay.mx, lay.my, lay.mz)
    sl = [None, None, None]
    for d in range(3):
        sl[AXIS[d]] = slice(4, 3 + m[d] - 1)
    sl[AXIS[a]] = 3 + idx
    return tuple(sl)


# ------------------------------------------------------------------------------------------------
# UpdateCartesianBCs
# ------------------------------------------------------------------------------------------------
def test_cartesian_bcs_patch_rules():
    bc = [NOSLIP, ZG, SLIP, FIX, ZG, NOSLIP]
    lay, mesh, params, rng = _setup(random_nodes((6, 7, 5), 11), (0, 0, 0), bc, 11)
    u0 = rng.standard_normal(lay.gshape + (3,))
    ufix = rng.standard_normal((6, 3))
    u = u0.copy()
    O.cartesian_bcs(lay.info(), mesh, params, u, ufix)
    m = (lay.mx, lay.my, lay.mz)
    want = u0.copy()
    for q in range(6):
        a, side = q // 2, q % 2
        g_idx, c1, c2 = (0, 1, 2) if side == 0 else (m[a] - 1, m[a] - 2, m[a] - 3)
        sg = _plane(lay, a, g_idx)
        sc, sc2 = _plane(lay, a, c1), _plane(lay, a, c2)
        if bc[q] == NOSLIP:
            want[sg] = -u0[sc]
        elif bc[q] == ZG:
            want[sg] = 2 * u0[sc] - u0[sc2]
        elif bc[q] == FIX:
            want[sg] = ufix[q]
        elif bc[q] == SLIP:
            M = np.stack([mesh[n][sc] for n in ("csi", "eta", "zet")], axis=-2)      # rows csi, eta, zet
            n = np.linalg.inv(M)[..., :, a]
            n = n / np.linalg.norm(n, axis=-1, keepdims=True)
            uc = u0[sc]
            mir = uc - 2.0 * (uc * n).sum(-1, keepdims=True) * n
            got = u[sg]
            assert np.abs(got - mir).max() <= 1e-13 * np.abs(uc).max()
            # a mirror image: same magnitude, normal component reversed, tangential kept
            assert np.allclose(np.linalg.norm(got, axis=-1), np.linalg.norm(uc, axis=-1), rtol=1e-13, atol=0)
            assert np.allclose((got * n).sum(-1), -(uc * n).sum(-1), rtol=0, atol=1e-13)
            want[sg] = got
    assert np.array_equal(u, want)


@pytest.mark.parametrize("geom", ["box", "sheared"])
def test_slip_normal_is_the_unit_lattice_vector(geom):
    vecs = AFFINE[geom]
    bc = [SLIP] * 6
    lay, mesh, params, rng = _setup(affine_nodes((5, 6, 4), vecs), (0, 0, 0), bc, 12)
    u0 = rng.standard_normal(lay.gshape + (3,))
    u = u0.copy()
    O.cartesian_bcs(lay.info(), mesh, params, u)
    m = (lay.mx, lay.my, lay.mz)
    for a in range(3):
        e = vecs[a] / np.linalg.norm(vecs[a])
        sg, sc = _plane(lay, a, 0), _plane(lay, a, 1)
        uc = u0[sc]
        want = uc - 2.0 * (uc @ e)[..., None] * e
        assert np.abs(u[sg] - want).max() <= 1e-13


@pytest.mark.parametrize("periodic", [(1, 0, 2), (2, 1, 0), (2, 2, 1)])
def test_cartesian_bcs_periodic(periodic):
    bc = []
    for a in range(3):
        bc += [PER, PER] if periodic[a] else [NOSLIP, ZG]
    lay, mesh, params, rng = _setup(random_nodes((6, 7, 5), 13), periodic, bc, 13)
    u0 = rng.standard_normal(lay.gshape + (3,))
    # type-2 ghosts consistent with the owned values, as a scattered local array holds them
    from tosca_amd.case import local_to_local
    import torch
    u0 = local_to_local(torch.as_tensor(u0), lay, periodic).numpy().copy()
    u = u0.copy()
    O.cartesian_bcs(lay.info(), mesh, params, u)
    m = (lay.mx, lay.my, lay.mz)
    for a in range(3):
        if not periodic[a]:
            continue
        assert np.array_equal(u[_plane(lay, a, 0)], u0[_plane(lay, a, m[a] - 2)]), a
        assert np.array_equal(u[_plane(lay, a, m[a] - 1)], u0[_plane(lay, a, 1)]), a


def test_cartesian_bcs_rejects_wall_and_inlet_patches():
    lay, mesh, params, rng = _setup(random_nodes((5, 5, 5), 14), (0, 0, 0), [WALL, ZG, ZG, ZG, ZG, ZG], 14)
    u = rng.standard_normal(lay.gshape + (3,))
    from tosca_amd.abi import ToscaError
    with pytest.raises(ToscaError):
        O.cartesian_bcs(lay.info(), mesh, params, u)
    with pytest.raises(ToscaError):
        O.cartesian_bcs(lay.info(), mesh, dict(params, bc_u=(GEN, ZG, ZG, ZG, ZG, ZG)), u)


def test_cartesian_bcs_rejects_k_wall_functions_and_missing_ustar():
    lay, mesh, params, rng = _setup(random_nodes((5, 5, 5), 15), (0, 0, 0), [ZG, ZG, WALL, ZG, ZG, ZG], 15)
    u = rng.standard_normal(lay.gshape + (3,))
    from tosca_amd.abi import ToscaError
    with pytest.raises(ToscaError):      # Cabot writes ustar
        O.cartesian_bcs(lay.info(), mesh, dict(params, wall_u=(0, 0, -1, 0)), u)
    with pytest.raises(ToscaError):      # a wall function without a wall model
        O.cartesian_bcs(lay.info(), mesh, dict(params, wall_u=(0, 0, 0, 0)), u)
    with pytest.raises(ToscaError):
        O.cartesian_bcs(lay.info(), mesh, dict(params, bc_u=(ZG, ZG, ZG, ZG, WALL, ZG), wall_u=(-3,) * 4), u)


# ------------------------------------------------------------------------------------------------
# kLeft inlet functions
# ------------------------------------------------------------------------------------------------
def _inlet_box(seed):
    """stretched box: x along k (the inflow patch is kLeft), y along i, z along j"""
    x = np.linspace(0.0, 60.0, 6)
    y = np.linspace(-20.0, 20.0, 9)                      # 8 cells across y
    z = np.cumsum(np.r_[0.0, 0.3 * 1.6 ** np.arange(7)])
    lay, mesh_t = C.mesh_metrics(C.cartesian_nodes(x, y, z, "cpu"), (0, 0, 0))
    cent = mesh_t["cent"].numpy().copy()
    bcn = [C.BC_ZEROGRAD] * 6
    params = C.default_params((0, 0, 0), [C.BC_FIXED] * 6, bcn, div_scheme=C.DIV_CENTRAL, les_model=C.LES_DSM, nu=0.01)
    params = dict(params, bc_u=(ZG, ZG, NOSLIP, SLIP, GEN, ZG))
    rng = np.random.Generator(np.random.PCG64(seed))
    return lay, np_mesh(mesh_t), params, cent, rng


def test_inlet_log_law():
    """log law max(uTau/0.4 ln(h/z0), 1e-5) along Udir below hInv, its value at hInv above, with
    h the height of the boundary cell's centre above the ground; nothing but ghost planes changes"""
    lay, mesh, params, cent, rng = _inlet_box(41)
    udir = np.array([0.8, 0.6, 0.0])
    inlet = dict(type=C.INLET_LOGLAW, cent=cent, grnd_level=0.05, u_tau=0.45, roughness=0.2, h_inv=8.0, udir=udir)
    u0 = rng.standard_normal(lay.gshape + (3,))
    u = u0.copy()
    O.cartesian_bcs(lay.info(), mesh, params, u, None, None, inlet)
    sg, sc = _plane(lay, 2, 0), _plane(lay, 2, 1)
    h = cent[sc][..., 2] - 0.05
    mag = np.where(h <= 8.0, np.maximum(0.45 / 0.4 * np.log(h / 0.2), 1e-5), 0.45 / 0.4 * np.log(8.0 / 0.2))
    assert (h < 0.2).any() and (h > 8.0).any()            # the clip and the cap both act
    assert np.abs(u[sg] - mag[..., None] * udir).max() <= 1e-14 * mag.max()
    above = u[sg][h > 8.0]
    assert np.array_equal(above, np.broadcast_to(above[0], above.shape))
    changed = u != u0                  # only ghost planes change (the other patches fill theirs)
    m = (lay.mx, lay.my, lay.mz)
    for q in range(6):
        changed[_plane(lay, q // 2, 0 if q % 2 == 0 else m[q // 2] - 1)] = False
    assert not changed.any()


def test_inlet_sinusoidal_mean():
    """sinusoidal profile: over whole periods across Ly on the uniform y cells the mean speed is
    |Uref|, the extremes are (1 +- amplitude cos(pi/4)) |Uref| at the cell-centre phases, along Udir"""
    lay, mesh, params, cent, rng = _inlet_box(42)
    uref = np.array([3.0, 0.0, 4.0])
    udir = uref / np.linalg.norm(uref)
    inlet = dict(type=C.INLET_SINUSOIDAL, cent=cent, ymin=-20.0, ly=40.0, amplitude=0.3, periods=2.0, udir=udir,
                 uref=uref)
    u = rng.standard_normal(lay.gshape + (3,))
    O.cartesian_bcs(lay.info(), mesh, params, u, None, None, inlet)
    g = u[_plane(lay, 2, 0)]                               # [j][i]
    mag = g @ udir
    assert np.abs(g - mag[..., None] * udir).max() <= 1e-14
    assert np.abs(mag.mean(axis=1) - 5.0).max() <= 1e-13
    # 4 cells per period: their centres sit at the phases pi/4 + k pi/2
    c45 = np.sqrt(0.5)
    assert abs(mag.max() - 5.0 * (1 + 0.3 * c45)) <= 1e-12 and abs(mag.min() - 5.0 * (1 - 0.3 * c45)) <= 1e-12
    from tosca_amd.abi import ToscaError
    with pytest.raises(ToscaError):                        # the power law needs the caller's noise draws
        O.cartesian_bcs(lay.info(), mesh, params, u, None, None, dict(inlet, type=1))
    with pytest.raises(ToscaError):                        # no typeU 7
        O.cartesian_bcs(lay.info(), mesh, params, u, None, None, dict(inlet, type=7))
    with pytest.raises(ToscaError):                        # inlet functions on kLeft only
        O.cartesian_bcs(lay.info(), mesh, dict(params, bc_u=(GEN, ZG, NOSLIP, SLIP, GEN, ZG)), u, None, None, inlet)


# ------------------------------------------------------------------------------------------------
# Cabot wall law
# ------------------------------------------------------------------------------------------------
def _inv_nut(y):
    """1 / (1 + nu_t+) with the near-wall eddy viscosity kappa y+ (1 - exp(-y+/19))^2, kappa 0.41"""
    return 1.0 / (1.0 + 0.41 * y * (1.0 - np.exp(-y / 19.0)) ** 2)


@pytest.mark.filterwarnings("ignore::scipy.integrate.IntegrationWarning")
@pytest.mark.parametrize("i", [1, 7, 60, 2500, 40000])
def test_cabot_table_is_the_integral(i):
    """table entry i = the integral of 1/(1 + nu_t+) over [0, 2 i] (scipy quad, per unit interval),
    to the 2.5e-11 relative error of 24 Simpson 3/8 panels per interval"""
    from scipy.integrate import quad
    want = sum(quad(_inv_nut, 2.0 * e, 2.0 * e + 2.0, epsabs=1e-16, epsrel=1e-15)[0] for e in range(min(i, 50)))
    if i > 50:
        want += quad(_inv_nut, 100.0, 2.0 * i, limit=400, epsabs=1e-13, epsrel=1e-15)[0]
    assert abs(O.cabot_table(i) - want) <= 3e-11 * want


@pytest.mark.filterwarnings("ignore::scipy.integrate.IntegrationWarning")
@pytest.mark.parametrize("nu,u,y", [(1.5e-5, 8.0, 5.0), (1e-3, 0.7, 0.05), (1e-2, 2.0, 0.3), (1.5e-5, 0.05, 1e-4),
                                    (1e-4, 1.0, 2e-3)])
def test_cabot_utau_solves_the_interpolated_law(nu, u, y):
    """u* from uTauCabot solves u = u* I(y u*/nu), I the table (entries from scipy quad) interpolated
    linearly at the even y+ points; the root is found independently with brentq"""
    from scipy.integrate import quad
    from scipy.optimize import brentq

    def I(yp):
        ib = int(yp / 2.0)
        lo = quad(_inv_nut, 0.0, 2.0 * ib, limit=400, epsabs=1e-13, epsrel=1e-13)[0] if ib else 0.0
        hi = lo + quad(_inv_nut, 2.0 * ib, 2.0 * ib + 2, epsabs=1e-15, epsrel=1e-15)[0]
        return (hi - lo) / 2.0 * (yp - 2.0 * ib) + lo

    us = O.cabot_utau(nu, u, y)
    ref = brentq(lambda t: t * I(y * t / nu) - u, 1e-8, 50.0, xtol=1e-15, rtol=1e-14)
    assert abs(us - ref) <= 1e-9 * ref
    yp = y * us / nu
    if yp < 0.5:              # viscous sublayer: u+ = y+
        assert abs(us - np.sqrt(nu * u / y)) <= 1e-3 * us


def _cabot_case(seed, wall_u, rough, nu=1e-3):
    bc = [WALL if wall_u[q] else ZG for q in range(4)] + [ZG, NOSLIP]
    lay, mesh, params, rng = _setup(affine_nodes((6, 7, 5), AFFINE["sheared"]), (0, 0, 0), bc, seed)
    params = dict(params, wall_u=tuple(wall_u), wall_rough=tuple(rough), nu=nu)
    u0 = rng.standard_normal(lay.gshape + (3,))
    return lay, mesh, params, u0


@pytest.mark.parametrize("rough", [0.0, 0.02])
def test_cabot_ghost_carries_the_wall_stress(rough):
    """on every Cabot patch: u* of the law for the boundary cell's parallel speed at
    sb = 0.5 / aj / |area| is stored in ustar (on the cells shared by an i and a j patch, the j
    patch's, which the reference computes last); the ghost's velocity difference to the boundary
    cell, over the distance 2 sb, times nu, is the wall stress u*^2 along the parallel direction
    where the clip at -|Ub_par| does not act, and the ghost's parallel speed is -|Ub_par| where it
    does; the normal part is reflected.  The smooth law always clips: its u+ = U/u* never exceeds
    y+ = sb u*/nu (the integrand 1/(1 + nu_t+) is at most 1), so the unclipped ghost speed
    u* (u+ - 2 y+) would be below -|Ub_par|.  The rough case runs at a viscosity where it does not."""
    nu = 1e-3 if rough == 0.0 else 2.0
    lay, mesh, params, u0 = _cabot_case(31, (-1, -1, -1, -1), (rough,) * 4, nu)
    u = u0.copy()
    ustar = np.full(lay.gshape, -7.0)
    O.cartesian_bcs(lay.info(), mesh, params, u, None, ustar)
    m = (lay.mx, lay.my, lay.mz)
    nu, nchecked = params["nu"], 0
    vecs = AFFINE["sheared"]
    for q in range(4):
        a, side = q // 2, q % 2
        sg, sc = _plane(lay, a, 0 if side == 0 else m[a] - 1), _plane(lay, a, 1 if side == 0 else m[a] - 2)
        n = vecs[a] / np.linalg.norm(vecs[a])              # calculateNormal's vector on an affine lattice
        area = np.linalg.norm(mesh[("csi", "eta")[a]][sc], axis=-1)
        sb = 0.5 / mesh["aj"][sc] / area
        ub = u0[sc]
        unor = (ub @ n)[..., None] * n
        upar = ub - unor
        mag = np.linalg.norm(upar, axis=-1)
        if rough == 0.0:
            us = np.vectorize(lambda mg, s: O.cabot_utau(nu, mg, s))(mag, sb)
        else:
            us = 0.41 * mag / np.log(sb / (0.033 * rough))
        stored = ustar[sc]
        if a == 0:            # rows j = 1 and my-2 hold the j patches' u*
            assert np.abs(stored[:, 1:-1] - us[:, 1:-1]).max() <= 1e-12 * us.max()
            assert np.abs(stored[:, [0, -1]] - us[:, [0, -1]]).max() > 1e-6 * us.max()
        else:
            assert np.abs(stored - us).max() <= 1e-12 * us.max()
        ug = u[sg]
        assert np.abs((ug + ub) @ n).max() <= 1e-12
        d = upar / mag[..., None]
        dot = lambda x, y: (x * y).sum(-1)
        stress = nu * dot(ub - ug, d) / (2 * sb)
        free = mag - 2 * sb * us ** 2 / nu > -mag
        clipped = -mag * mag / (mag + 1e-12)      # -|Ub_par| along nUnit's direction (its eps 1e-12)
        if rough == 0.0:
            assert not free.any() and (sb * us / nu >= mag / us).all()
            assert np.abs(dot(ug, d) - clipped).max() <= 1e-14
        else:
            assert np.abs(stress - us ** 2)[free].max() <= 1e-9 * (us ** 2).max()
            assert np.abs(dot(ug, d) - clipped)[~free].max(initial=0.0) <= 1e-14
            nchecked += free.sum()
        assert np.abs((ug - dot(ug, d)[..., None] * d + unor)).max() <= 1e-12
    assert rough == 0.0 or nchecked > 20
    written = ustar != -7.0             # only the boundary cells of the i and j patches
    for q in range(4):
        a, side = q // 2, q % 2
        written[_plane(lay, a, 1 if side == 0 else m[a] - 2)] = False
    assert not written.any()


def test_cabot_rough_wall_laws_and_shared_roughness():
    """rough walls: u* = kappa U / ln(sb / z0), z0 = 0.033 ks (log law; U the parallel speed), and
    U / u* = 8.3 (sb u* / nu)^(1/7) for 0.033 ks <= 1e-13 (power law); the right patch of an axis
    uses the left patch's roughness when both are Cabot"""
    lay, mesh, params, u0 = _cabot_case(32, (-1, -1, -1, -3), (0.1, 0.5, 1e-12, 0.0))
    u = u0.copy()
    ustar = np.zeros(lay.gshape)
    O.cartesian_bcs(lay.info(), mesh, params, u, None, ustar)
    m = (lay.mx, lay.my, lay.mz)
    vecs = AFFINE["sheared"]
    nu = params["nu"]
    for q, ks in ((0, 0.1), (1, 0.1), (2, 1e-12)):
        a, side = q // 2, q % 2
        sc = _plane(lay, a, 1 if side == 0 else m[a] - 2)
        n = vecs[a] / np.linalg.norm(vecs[a])
        area = np.linalg.norm(mesh[("csi", "eta")[a]][sc], axis=-1)
        sb = 0.5 / mesh["aj"][sc] / area
        ub = u0[sc]
        mag = np.linalg.norm(ub - (ub @ n)[..., None] * n, axis=-1)
        us = ustar[sc]
        if a == 0:            # row j = 1 holds the jLeft Cabot patch's u*
            us, mag, sb = us[:, 1:], mag[:, 1:], sb[:, 1:]
        if 0.033 * ks > 1e-13:
            assert np.abs(us / 0.41 * np.log(sb / (0.033 * ks)) - mag).max() <= 1e-13 * mag.max()
        else:
            assert np.abs(8.3 * us * (sb * us / nu) ** (1 / 7) - mag).max() <= 1e-12 * mag.max()


def test_shumann_patch_rule():
    """Shumann wall functions (wall_u = -3) on all four i and j patches: on the j patches the ghost
    is 2 Ub_par - Uc_par - Ub_nor, with the patch normal taken as the normalised j column of the
    inverse of the (csi, eta, zet) matrix (numpy.linalg.inv); the i patches' ghosts are untouched"""
    bc = [WALL, WALL, WALL, WALL, ZG, NOSLIP]
    lay, mesh, params, rng = _setup(random_nodes((6, 7, 5), 16), (0, 0, 0), bc, 16)
    params = dict(params, wall_u=(-3, -3, -3, -3))
    u0 = rng.standard_normal(lay.gshape + (3,))
    u = u0.copy()
    O.cartesian_bcs(lay.info(), mesh, params, u)
    m = (lay.mx, lay.my, lay.mz)
    want = u0.copy()
    for q in range(2, 6):
        a, side = q // 2, q % 2
        g_idx, c1, c2 = (0, 1, 2) if side == 0 else (m[a] - 1, m[a] - 2, m[a] - 3)
        sg, sc, sc2 = _plane(lay, a, g_idx), _plane(lay, a, c1), _plane(lay, a, c2)
        if bc[q] == NOSLIP:
            want[sg] = -u0[sc]
        elif bc[q] == ZG:
            want[sg] = 2 * u0[sc] - u0[sc2]
        else:
            M = np.stack([mesh[n][sc] for n in ("csi", "eta", "zet")], axis=-2)
            n = np.linalg.inv(M)[..., :, 1]
            n = n / np.linalg.norm(n, axis=-1, keepdims=True)
            nor = lambda v: (v * n).sum(-1, keepdims=True) * n
            ub, uc = u0[sc], u0[sc2]
            ref = 2 * (ub - nor(ub)) - (uc - nor(uc)) - nor(ub)
            got = u[sg]
            assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max()
            want[sg] = got
    assert np.array_equal(u, want)


@pytest.mark.parametrize("geom", ["box", "sheared"])
def test_shumann_no_penetration_and_preserved_gradient(geom):
    """on an affine lattice calculateNormal's j vector is the unit lattice vector e of the j axis
    (as for slip).  With a velocity whose part orthogonal to e is linear in the cell index j and
    whose part along e is arbitrary, the Shumann ghost continues the orthogonal profile linearly
    to the ghost cell (the gradient at the wall equals the interior one) and makes the part along
    e averaged across the wall face zero (no penetration)"""
    vecs = AFFINE[geom]
    lay, mesh, params, rng = _setup(affine_nodes((5, 6, 4), vecs), (0, 0, 0), [ZG, ZG, WALL, WALL, ZG, ZG], 17)
    params = dict(params, wall_u=(0, 0, -3, -3))
    m = (lay.mx, lay.my, lay.mz)
    e = vecs[1] / np.linalg.norm(vecs[1])
    par = np.cross(e, [0.3, 0.1, 1.0])
    par = par / np.linalg.norm(par)
    jj = np.arange(lay.my + 6)[None, :, None] - 3.0       # cell index j of the local array
    w = rng.standard_normal(lay.gshape)
    u0 = (0.7 + 0.25 * jj)[..., None] * par + w[..., None] * e
    u = u0.copy()
    O.cartesian_bcs(lay.info(), mesh, params, u)
    for side in (0, 1):
        g, b = (0, 1) if side == 0 else (m[1] - 1, m[1] - 2)
        sg, sb = _plane(lay, 1, g), _plane(lay, 1, b)
        ug = u[sg]
        assert np.abs(ug @ par - (0.7 + 0.25 * g)).max() <= 1e-13
        assert np.abs((ug + u0[sb]) @ e).max() <= 1e-13 * np.abs(w).max()
        assert np.abs(ug - (ug @ par)[..., None] * par - (ug @ e)[..., None] * e).max() <= 1e-13


# ------------------------------------------------------------------------------------------------
# UpdateContravariantBCs
# ------------------------------------------------------------------------------------------------
def _face_plane(lay, a, f):
    """face index f along direction a, transverse range [0, m-1) (the reference's loop range)"""
    m = (lay.mx, lay.my, lay.mz)
    sl = [None, None, None]
    for d in range(3):
        sl[AXIS[d]] = slice(3, 3 + m[d] - 1)
    sl[AXIS[a]] = 3 + f
    return tuple(sl)


def test_contravariant_bcs_patch_rules():
    bc = [FIX, NOSLIP, ZG, SLIP, ZG, WALL]
    lay, mesh, params, rng = _setup(random_nodes((6, 7, 5), 15), (0, 0, 0), bc, 15)
    V0 = rng.standard_normal(lay.gshape + (3,))
    U = rng.standard_normal(lay.gshape + (3,))
    V = V0.copy()
    O.contravariant_bcs(lay.info(), mesh, params, V, U)
    m = (lay.mx, lay.my, lay.mz)
    want = V0.copy()
    area = {0: "icsi", 1: "jeta", 2: "kzet"}
    for q in range(6):
        a, side = q // 2, q % 2
        f = 0 if side == 0 else m[a] - 2
        sf = _face_plane(lay, a, f)
        nb = list(sf)
        nb[AXIS[a]] = sf[AXIS[a]] + 1
        if bc[q] in (NOSLIP, SLIP, WALL):
            want[sf + (a,)] = 0.0
        else:
            v = ((0.5 * (U[tuple(nb)] + U[sf])) * mesh[area[a]][sf]).sum(-1)
            if bc[q] == ZG:
                v = np.minimum(v, 0.0) if side == 0 else np.maximum(v, 0.0)
            want[sf + (a,)] = v
    assert np.abs(V - want).max() <= 1e-14 * np.abs(want).max()
    # nothing but those faces changes
    diff = V != V0
    for q in range(6):
        a, side = q // 2, q % 2
        diff[_face_plane(lay, a, 0 if side == 0 else m[a] - 2) + (a,)] = False
    assert not diff.any()


@pytest.mark.parametrize("geom", ["box", "sheared"])
def test_contravariant_bcs_uniform_flow(geom):
    """interpolated boundary fluxes of a uniform velocity on an affine lattice are S . U"""
    vecs = AFFINE[geom]
    lay, mesh, params, rng = _setup(affine_nodes((5, 6, 4), vecs), (0, 0, 0), [FIX] * 6, 16)
    U0 = np.array([0.7, -0.3, 0.45])
    U = np.broadcast_to(U0, lay.gshape + (3,)).copy()
    V = np.zeros(lay.gshape + (3,))
    O.contravariant_bcs(lay.info(), mesh, params, V, U)
    S, _, _ = affine_geometry(vecs)
    m = (lay.mx, lay.my, lay.mz)
    for a in range(3):
        for f in (0, m[a] - 2):
            # faces with interior transverse indices (the builder's face metrics of the boundary
            # rows across the patch are zero)
            sl = list(_plane(lay, a, f))
            got = V[tuple(sl) + (a,)]
            assert np.abs(got - S[a] @ U0).max() <= 1e-14 * abs(S[a] @ U0) + 1e-15


@pytest.mark.parametrize("ptype", [C.PERIODIC_COPY, C.PERIODIC_WRAP])
def test_contravariant_bcs_periodic_faces(ptype):
    periodic = (ptype, 0, 0)
    bc = [PER, PER, NOSLIP, NOSLIP, ZG, ZG]
    lay, mesh, params, rng = _setup(random_nodes((6, 7, 5), 17), periodic, bc, 17)
    import torch
    from tosca_amd.case import local_to_local
    V0 = local_to_local(torch.as_tensor(rng.standard_normal(lay.gshape + (3,))), lay, periodic).numpy().copy()
    U = rng.standard_normal(lay.gshape + (3,))
    V = V0.copy()
    O.contravariant_bcs(lay.info(), mesh, params, V, U)
    mx = lay.mx
    s0, sR = _face_plane(lay, 0, 0), _face_plane(lay, 0, mx - 2)
    # left face: the right face's flux from before the call; right face: re-interpolated from ucat
    assert np.array_equal(V[s0 + (0,)], V0[sR + (0,)])
    nb = list(sR)
    nb[2] = sR[2] + 1
    v = ((0.5 * (U[tuple(nb)] + U[sR])) * mesh["icsi"][sR]).sum(-1)
    assert np.abs(V[sR + (0,)] - v).max() <= 1e-14 * np.abs(v).max()


# ------------------------------------------------------------------------------------------------
# UpdateTemperatureBCs
# ------------------------------------------------------------------------------------------------
def test_temperature_bcs_patch_rules():
    T_ZG, T_FV, T_FG, T_PER = C.BCT_ZEROGRADIENT, C.BCT_FIXEDVALUE, C.BCT_FIXEDGRADIENT, C.BCT_PERIODIC
    bct = [T_ZG, T_FV, T_FG, T_FG, T_FV, T_ZG]
    lay, mesh, params, rng = _setup(affine_nodes((6, 7, 5), AFFINE["box"]), (0, 0, 0), [ZG] * 6, 18)
    vals = rng.standard_normal(6)
    t0 = 300.0 + rng.standard_normal(lay.gshape)
    t = t0.copy()
    O.temperature_bcs(lay.info(), mesh, params, bct, vals, t)
    m = (lay.mx, lay.my, lay.mz)
    want = t0.copy()
    width = [np.linalg.norm(v) for v in AFFINE["box"]]       # cell width across an i, j, k patch
    for q in range(6):
        a, side = q // 2, q % 2
        sg, sc = _plane(lay, a, 0 if side == 0 else m[a] - 1), _plane(lay, a, 1 if side == 0 else m[a] - 2)
        if bct[q] == T_ZG:
            want[sg] = t0[sc]
        elif bct[q] == T_FV:
            want[sg] = vals[q]
        elif bct[q] == T_FG:
            # box lattice: (1/aj) / |area| is the cell width across the patch
            assert np.allclose((t[sg] - t0[sc]) / width[a], vals[q], rtol=1e-13, atol=1e-12)
            want[sg] = t[sg]
    assert np.array_equal(t, want)


def test_temperature_bcs_surface_temperature():
    """jLeft with the surface-temperature model: the ghost makes the temperature interpolated to
    the wall face (the mean of ghost and boundary cell) equal the surface temperature; other
    patches reject the kind"""
    bct = [C.BCT_ZEROGRADIENT, C.BCT_ZEROGRADIENT, C.BCT_SURFACETEMP, C.BCT_FIXEDVALUE, C.BCT_ZEROGRADIENT,
           C.BCT_ZEROGRADIENT]
    lay, mesh, params, rng = _setup(random_nodes((6, 7, 5), 19), (0, 0, 0), [ZG] * 6, 19)
    vals = np.array([0.0, 0.0, 287.25, 301.0, 0.0, 0.0])
    t0 = 300.0 + rng.standard_normal(lay.gshape)
    t = t0.copy()
    O.temperature_bcs(lay.info(), mesh, params, bct, vals, t)
    sg, sc = _plane(lay, 1, 0), _plane(lay, 1, 1)
    assert np.abs(0.5 * (t[sg] + t0[sc]) - vals[2]).max() <= 1e-12
    m = (lay.mx, lay.my, lay.mz)
    assert (t[_plane(lay, 1, m[1] - 1)] == vals[3]).all()
    from tosca_amd.abi import ToscaError
    for q in (0, 1, 3, 4, 5):
        bad = list(bct)
        bad[2], bad[q] = C.BCT_ZEROGRADIENT, C.BCT_SURFACETEMP
        with pytest.raises(ToscaError):
            O.temperature_bcs(lay.info(), mesh, params, bad, vals, t0.copy())


@pytest.mark.parametrize("periodic", [(1, 0, 2), (2, 1, 0)])
def test_temperature_bcs_periodic(periodic):
    bct = []
    for a in range(3):
        bct += [C.BCT_PERIODIC] * 2 if periodic[a] else [C.BCT_ZEROGRADIENT] * 2
    bc = [PER if periodic[q // 2] else ZG for q in range(6)]
    lay, mesh, params, rng = _setup(random_nodes((6, 7, 5), 19), periodic, bc, 19)
    import torch
    from tosca_amd.case import local_to_local
    t0 = local_to_local(torch.as_tensor(rng.standard_normal(lay.gshape)), lay, periodic).numpy().copy()
    t = t0.copy()
    O.temperature_bcs(lay.info(), mesh, params, bct, np.zeros(6), t)
    m = (lay.mx, lay.my, lay.mz)
    for a in range(3):
        if periodic[a]:
            assert np.array_equal(t[_plane(lay, a, 0)], t0[_plane(lay, a, m[a] - 2)])
            assert np.array_equal(t[_plane(lay, a, m[a] - 1)], t0[_plane(lay, a, 1)])
