"""Pins of the oracle's ExplicitRhsU restatement (oracle_explicit_rhs_u) from facts that do not
depend on its code:

  - contravariantToCartesian inverts the face interpolation of a linear velocity field exactly on
    an affine lattice, and on a random curvilinear mesh equals a per-cell LU solve (numpy) of the
    cell-basis system with the face-averaged fluxes;
  - the no-penetration, periodic and non-resolved-face resets follow the patch rules of
    src/include/inline.h:5262-5355, 5357-5451, 1674-1838, written here as numpy masks;
  - the right-hand side is -dP + bTheta terms + farm + FormU: linear in each source term, with the
    AB2 weights 3/2, -1/2 (ueqn_solvers.c:1304-1347).
"""
import numpy as np
import pytest
import torch

from tosca_amd import case as C
from oracle import oracle as O
from tests._util import np_mesh, to_np, random_nodes
from tests.test_oracle_pins import AFFINE, affine_nodes

NOSLIP, ZG, SLIP, WALL, GEN, PER = (C.BCU_NOSLIP, C.BCU_ZEROGRADIENT, C.BCU_SLIP, C.BCU_WALLFUNCTION,
                                    C.BCU_GENERIC, C.BCU_PERIODIC)


def _setup(nodes, periodic, bc_u, seed=3):
    lay, mesh = C.mesh_metrics(nodes, periodic)
    rng = np.random.Generator(np.random.PCG64(seed))
    gs, glob = lay.gshape, (lay.mz, lay.my, lay.mx, 3)
    f = dict(ucont=rng.standard_normal(gs + (3,)), ucat=rng.standard_normal(gs + (3,)),
             nut=rng.uniform(0.0, 0.05, gs), dp=rng.standard_normal(glob), bt=rng.standard_normal(glob),
             bto=rng.standard_normal(glob), farm=rng.standard_normal(glob), ucont_o=rng.standard_normal(glob))
    bcn = [C.BC_PERIODIC if periodic[a] else C.BC_ZEROGRAD for a in range(3) for _ in range(2)]
    params = C.default_params(periodic, [C.BC_PERIODIC if periodic[a // 2] else C.BC_FIXED for a in range(6)], bcn,
                              div_scheme=C.DIV_CENTRAL, les_model=C.LES_DSM, nu=0.013)
    params = dict(params, bc_u=tuple(bc_u))
    return lay, np_mesh(mesh), f, params


def _run(lay, mesh, f, params, dp=True, bt=True, ab2=1, farm=True, form_mode=0):
    ucont, ucat = f["ucont"].copy(), f["ucat"].copy()
    rhs = np.full((lay.mz, lay.my, lay.mx, 3), np.nan)
    zero = np.zeros_like(f["dp"])
    O.explicit_rhs_u(lay.info(), mesh, params, ucont, ucat, f["nut"], f["dp"] if dp else zero,
                     f["bt"] if bt else None, f["bto"] if bt else None, ab2, f["farm"] if farm else None,
                     f["ucont_o"], rhs, form_mode)
    return ucont, ucat, rhs


def _own(a, lay):
    """owned range [0, m) of a local array, [k][j][i]"""
    return a[3:3 + lay.mz, 3:3 + lay.my, 3:3 + lay.mx]


# ------------------------------------------------------------------------------------------------
# contravariantToCartesian
# ------------------------------------------------------------------------------------------------
@pytest.mark.parametrize("geom", ["box", "sheared"])
def test_contra_to_cart_recovers_linear_field(geom):
    """On an affine lattice every metric is constant; fluxes V = S^q . u(face) of a linear u(face)
    average to S^q . u(cell), so the cell solve must return u(cell) exactly (to round-off)."""
    n = (7, 6, 5)
    lay, mesh, f, params = _setup(affine_nodes(n, AFFINE[geom]), (0, 0, 0), [GEN] * 6)
    gz, gy, gx = lay.gshape
    I = np.arange(gx) - 3
    J = np.arange(gy) - 3
    K = np.arange(gz) - 3
    a0, ai, aj, ak = np.array([0.3, -0.2, 0.5]), np.array([0.7, 0.1, -0.4]), np.array([-0.3, 0.6, 0.2]), \
        np.array([0.25, -0.5, 0.9])
    u_cell = a0 + ai * I[None, None, :, None] + aj * J[None, :, None, None] + ak * K[:, None, None, None]
    S = [mesh["csi"][3, 3, 3], mesh["eta"][3, 3, 3], mesh["zet"][3, 3, 3]]
    ucont = np.zeros_like(f["ucont"])
    for q, d in enumerate(((0, 0, 1), (0, 1, 0), (1, 0, 0))):      # face q between cell c and c + e_q
        u_face = u_cell + 0.5 * (ai * d[2] + aj * d[1] + ak * d[0])
        ucont[..., q] = u_face @ S[q]
    f = dict(f, ucont=ucont)
    _, ucat, _ = _run(lay, mesh, f, params)
    got = _own(ucat, lay)[1:-1, 1:-1, 1:-1]
    want = _own(u_cell, lay)[1:-1, 1:-1, 1:-1]
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()


@pytest.mark.parametrize("seed", [0, 1])
def test_contra_to_cart_matches_lu_solve(seed):
    lay, mesh, f, params = _setup(random_nodes((6, 7, 5), seed), (0, 0, 0), [GEN] * 6, seed=seed)
    ucont_in = f["ucont"]
    _, ucat, _ = _run(lay, mesh, f, params)
    Vo = _own(ucont_in, lay)
    q = np.stack([0.5 * (Vo[1:-1, 1:-1, :-2, 0] + Vo[1:-1, 1:-1, 1:-1, 0]),
                  0.5 * (Vo[1:-1, :-2, 1:-1, 1] + Vo[1:-1, 1:-1, 1:-1, 1]),
                  0.5 * (Vo[:-2, 1:-1, 1:-1, 2] + Vo[1:-1, 1:-1, 1:-1, 2])], axis=-1)
    M = np.stack([_own(mesh[n], lay)[1:-1, 1:-1, 1:-1] for n in ("csi", "eta", "zet")], axis=-2)   # rows
    want = np.linalg.solve(M, q[..., None])[..., 0]
    got = _own(ucat, lay)[1:-1, 1:-1, 1:-1]
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()
    # cells outside the interior keep their values on a non-periodic mesh
    mask = np.ones(lay.gshape, bool)
    mask[4:3 + lay.mz - 1, 4:3 + lay.my - 1, 4:3 + lay.mx - 1] = False
    assert np.array_equal(ucat[mask], f["ucat"][mask])


# ------------------------------------------------------------------------------------------------
# resetNoPenetrationFluxes
# ------------------------------------------------------------------------------------------------
def test_no_penetration_patch_rules():
    bc = [NOSLIP, ZG, SLIP, WALL, ZG, GEN]     # iLeft iRight jLeft jRight kLeft kRight
    lay, mesh, f, params = _setup(random_nodes((6, 7, 5), 4), (0, 0, 0), bc, seed=4)
    ucont, _, _ = _run(lay, mesh, f, params, form_mode=1)    # convective only: no wall stress read
    V0, V = _own(f["ucont"], lay), _own(ucont, lay)
    mx, my, mz = lay.mx, lay.my, lay.mz
    want = V0.copy()
    r = (slice(0, mz - 1), slice(0, my - 1))          # the reference's loops: k < mz-1, j < my-1, i < mx-1
    want[r[0], r[1], 0, 0] = 0.0                                              # iLeft noSlip
    w = want[r[0], r[1], mx - 2, 0]; w[w < 0] = 0.0; want[r[0], r[1], mx - 2, 0] = w     # iRight zeroGradient
    want[:mz - 1, 0, :mx - 1, 1] = 0.0                                        # jLeft slip
    want[:mz - 1, my - 2, :mx - 1, 1] = 0.0                                   # jRight wall function
    w = want[0, :my - 1, :mx - 1, 2]; w[w > 0] = 0.0; want[0, :my - 1, :mx - 1, 2] = w   # kLeft zeroGradient
    assert np.array_equal(V, want)
    # the ghost layers of a non-periodic mesh are not touched
    mask = np.ones(lay.gshape, bool)
    mask[3:3 + mz, 3:3 + my, 3:3 + mx] = False
    assert np.array_equal(ucont[mask], f["ucont"][mask])


# ------------------------------------------------------------------------------------------------
# periodic resets of ucont and ucat
# ------------------------------------------------------------------------------------------------
@pytest.mark.parametrize("periodic", [(1, 0, 2), (2, 1, 0), (1, 1, 1), (2, 2, 2)])
def test_periodic_resets(periodic):
    bc = []
    for a in range(3):
        bc += [PER, PER] if periodic[a] else [NOSLIP, ZG]
    lay, mesh, f, params = _setup(random_nodes((6, 7, 5), 5), periodic, bc, seed=5)
    ucont, ucat, _ = _run(lay, mesh, f, params)
    V, U = _own(ucont, lay), _own(ucat, lay)
    m = (lay.mx, lay.my, lay.mz)
    for a in range(3):
        if not periodic[a]:
            continue
        ax = {0: 2, 1: 1, 2: 0}[a]
        take = lambda arr, i: np.take(arr, i, axis=ax)
        assert np.array_equal(take(V, 0), take(V, m[a] - 2)), a            # face 0 = face m-2
        assert np.array_equal(take(U, 0), take(U, m[a] - 2)), a            # cell 0 = cell m-2
        assert np.array_equal(take(U, m[a] - 1), take(U, 1)), a           # cell m-1 = cell 1
        if periodic[a] == C.PERIODIC_WRAP:
            # type-2 ghosts wrap the owned values: ghost index -1 holds index m-1, on the owned
            # range of the other two axes
            own = [slice(3, 3 + lay.mz), slice(3, 3 + lay.my), slice(3, 3 + lay.mx)]
            lo, hi = list(own), list(own)
            lo[ax], hi[ax] = 2, 3 + m[a] - 1
            assert np.array_equal(ucat[tuple(lo)], ucat[tuple(hi)]), a
            assert np.array_equal(ucont[tuple(lo)], ucont[tuple(hi)]), a


# ------------------------------------------------------------------------------------------------
# right-hand side assembly and resetNonResolvedCellFaces
# ------------------------------------------------------------------------------------------------
def _formu_part(lay, mesh, f, params, form_mode=0):
    """FormU of the reset fields onto a zero rhs, from the oracle's FormU."""
    ucont, ucat = f["ucont"].copy(), f["ucat"].copy()
    rhs = np.zeros((lay.mz, lay.my, lay.mx, 3))
    zero = np.zeros_like(f["dp"])
    O.explicit_rhs_u(lay.info(), mesh, params, ucont, ucat, f["nut"], zero, None, None, 0, None, f["ucont_o"],
                     rhs, form_mode)
    return ucont, ucat


def _nonresolved_mask(lay, params, ucont_o, periodic):
    """True where resetNonResolvedCellFaces zeroes the value (inline.h:1674-1838)."""
    mx, my, mz = lay.mx, lay.my, lay.mz
    bc = params["bc_u"]
    z = np.zeros((mz, my, mx, 3), bool)
    for a, (m, ax) in enumerate(((mx, 2), (my, 1), (mz, 0))):
        idx = np.arange(m)
        shape = [1, 1, 1]; shape[ax] = m
        I = idx.reshape(shape)
        uo = ucont_o[..., a]
        keep0 = (bc[2 * a] == ZG) | (uo < 0)
        keepR = (bc[2 * a + 1] == ZG) | (uo > 0)
        zero_a = ((I == 0) & ~keep0) | (I == m - 1)
        if not periodic[a]:
            zero_a |= (I == m - 2) & ~keepR
        z[..., a] = np.broadcast_to(zero_a, (mz, my, mx))
    return z


@pytest.mark.parametrize("periodic,bc", [((0, 0, 0), [NOSLIP, ZG, SLIP, GEN, ZG, GEN]),
                                         ((1, 0, 2), [PER, PER, ZG, NOSLIP, PER, PER])])
def test_rhs_assembly_and_nonresolved_faces(periodic, bc):
    lay, mesh, f, params = _setup(random_nodes((6, 7, 5), 6), periodic, bc, seed=6)
    _, _, r_all = _run(lay, mesh, f, params, dp=True, bt=True, ab2=1, farm=True)
    _, _, r_none = _run(lay, mesh, f, params, dp=False, bt=False, ab2=0, farm=False)
    _, _, r_ab1 = _run(lay, mesh, f, params, dp=False, bt=True, ab2=0, farm=False)
    zero = _nonresolved_mask(lay, params, f["ucont_o"], periodic)
    # zeroed faces are zero in every run; elsewhere the source terms add linearly
    for r in (r_all, r_none, r_ab1):
        assert np.all(r[zero] == 0.0)
    keep = ~zero
    src = -f["dp"] + 1.5 * f["bt"] - 0.5 * f["bto"] + f["farm"]
    d = (r_all - r_none)[keep]
    assert np.abs(d - src[keep]).max() <= 1e-12 * np.abs(src).max()
    d1 = (r_ab1 - r_none)[keep]
    assert np.abs(d1 - f["bt"][keep]).max() <= 1e-12 * np.abs(f["bt"]).max()
    # FormU part: the oracle's FormU of the reset fields, rhs from zero
    ucont, ucat = _formu_part(lay, mesh, f, params)
    rf = np.zeros_like(r_none)
    O.formu(lay.info(), mesh, params, ucont, ucat, f["nut"], rf, 1.0, 0)
    assert np.array_equal(r_none[keep], rf[keep])


def test_wall_function_patches_zero_the_normal_flux():
    """velocityWallFunction patches (j here) take part in the no-penetration reset."""
    bc = [GEN, GEN, WALL, WALL, GEN, GEN]
    lay, mesh, f, params = _setup(random_nodes((6, 7, 5), 7), (0, 0, 0), bc, seed=7)
    ucont, ucat = f["ucont"].copy(), f["ucat"].copy()
    rhs = np.zeros((lay.mz, lay.my, lay.mx, 3))
    O.explicit_rhs_u(lay.info(), mesh, params, ucont, ucat, f["nut"], f["dp"], None, None, 0, None, f["ucont_o"],
                     rhs, 1)      # formMode 1: convective only, no wall stress needed
    V = _own(ucont, lay)
    assert np.all(V[:lay.mz - 1, 0, :lay.mx - 1, 1] == 0.0)
    assert np.all(V[:lay.mz - 1, lay.my - 2, :lay.mx - 1, 1] == 0.0)
