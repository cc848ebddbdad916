"""Pins of the oracle's FormU at non-periodic boundaries: one-sided transverse differences, QUICK
clipping, slip and wall-function faces, the fp ghost copies, and the metric of the face dot.

The closed forms of tests/test_oracle_pins.py hold on fully periodic lattices only.  Here the
lattice is affine (every metric constant) but bounded, and each expectation is written from a
fact outside the oracle's code:

  * a velocity LINEAR in the lattice coordinates, u = b + sum_q B_q xi_q, has every difference the
    reference takes -- centred (Compute_du_i/j/k's four-point averages, inline.h:2748-2931) or
    one-sided (their corner rows, inline.h:2762-2930) -- equal to B_q exactly, so every viscous face
    flux is nu_eff (grad u + grad u^T) . S exactly, boundary faces and corner rows included;
  * the ghost points those one-sided rows exist to avoid -- points outside the interior along two
    axes, one of them non-periodic, and the points beyond the first ghost layer of a non-periodic
    axis -- are set to NaN: a branch that reads one of them (a centred difference where the
    reference takes a one-sided one, an unclipped QUICK stencil) poisons the result;
  * QUICK is the quadratic through the two upwind cells and the downwind cell (Lagrange weights
    from a Vandermonde solve); at a non-periodic boundary face whose far-upwind cell would lie
    beyond the ghost layer the reference takes u_UU := u_U (inline.h:2215-2227, 5089-5109);
  * slip faces carry no viscous flux with LES on, wall-function faces -tau + nu (grad u + grad u^T).S
    with tau the wall model's (iLWM for both i patches, jLWM for both j patches; k patches have no
    wall-function branch) (ueqn_solvers.c:326-356, 538-568, 748-772);
  * the ghost cell beyond a non-periodic boundary carries the fp of the last interior cell, a
    periodic ghost the fp of the opposite interior cell (ueqn_solvers.c:949-1000);
  * the face dot dots each side's fp with that side's CELL-CENTRED basis (csi, eta, zet), not with
    the face vector, and halves the sum (ueqn_solvers.c:1006-1035): pinned on a randomly perturbed
    lattice, where cell and face vectors differ, with a uniform velocity so that fp = -u0 div(V)
    is known without any flux formula.
"""
from __future__ import annotations

import itertools

import numpy as np
import pytest
import torch

from tosca_amd import case as C
from oracle import oracle as O
from tests._util import np_mesh, to_np, relerr, random_nodes
from tests.test_oracle_pins import AFFINE, affine_nodes, affine_geometry, lagrange_weights

G = C.G
AXN = {0: 2, 1: 1, 2: 0}           # computational direction -> array axis of [k][j][i]


def _m(lay):
    return (lay.mx, lay.my, lay.mz)


def _index_grids(lay):
    """DMDA indices (i, j, k) of every point of the local array, shaped like the array."""
    K, J, I = np.meshgrid(np.arange(lay.mz + 2 * G) - G, np.arange(lay.my + 2 * G) - G,
                          np.arange(lay.mx + 2 * G) - G, indexing="ij")
    return I, J, K


def poison_mask(lay, periodic):
    """Points FormU must not read: beyond the first ghost layer of a non-periodic axis, or outside
    the interior cells along two or more axes with at least one of them non-periodic."""
    I, J, K = _index_grids(lay)
    m = _m(lay)
    idx = (I, J, K)
    deep = np.zeros(I.shape, bool)
    out = []
    for a in range(3):
        x = idx[a]
        o = (x < 1) | (x > m[a] - 2)
        out.append(o)
        if not periodic[a]:
            deep |= (x < 0) | (x > m[a] - 1)
    nout = out[0].astype(int) + out[1] + out[2]
    nonper_out = np.zeros(I.shape, bool)
    for a in range(3):
        if not periodic[a]:
            nonper_out |= out[a]
    return deep | ((nout >= 2) & nonper_out)


def linear_field(lay, periodic, rng):
    """u = b + sum_q B_q xi_q at every point (xi_q: the DMDA index along q, i.e. lattice units),
    B_q = 0 along periodic axes so that periodic ghost copies agree with the linear extension."""
    B = rng.standard_normal((3, 3))              # rows q: d u / d xi_q
    for q in range(3):
        if periodic[q]:
            B[q] = 0.0
    b = rng.standard_normal(3)
    I, J, K = _index_grids(lay)
    u = b + I[..., None] * B[0] + J[..., None] * B[1] + K[..., None] * B[2]
    return u, B


def bounded_case(geom, periodic, div_scheme, bc_u, seed):
    vecs = AFFINE[geom]
    n = (7, 6, 5)
    lay, mesh = C.mesh_metrics(affine_nodes(n, vecs), periodic)
    S, vol, gxi = affine_geometry(vecs)
    nm = np_mesh(mesh)
    # the face vectors and Jacobians FormU reads (faces 0 .. m-2 along their axis, interior
    # transverse rows) and the cell bases of the owned points are the lattice's
    I, J, K = _index_grids(lay)
    m = _m(lay)
    idx = (I, J, K)
    for q, names in ((0, ("icsi", "ieta", "izet", "iaj")), (1, ("jcsi", "jeta", "jzet", "jaj")),
                     (2, ("kcsi", "keta", "kzet", "kaj"))):
        read = np.ones(I.shape, bool)
        for a in range(3):
            lo, hi = (0, m[a] - 2) if a == q else (1, m[a] - 2)
            read &= (idx[a] >= lo) & (idx[a] <= hi)
        for b, name in enumerate(names[:3]):
            assert np.abs(nm[name][read] - S[b]).max() <= 1e-12, name
        assert np.abs(nm[names[3]][read] * vol - 1.0).max() <= 1e-12, names[3]
    own = (slice(G, G + lay.mz), slice(G, G + lay.my), slice(G, G + lay.mx))
    for name, a in (("csi", 0), ("eta", 1), ("zet", 2)):
        assert np.abs(nm[name][own] - S[a]).max() <= 1e-12, name
    rng = np.random.Generator(np.random.PCG64(seed))
    u, B = linear_field(lay, periodic, rng)
    bad = poison_mask(lay, periodic)
    u[bad] = np.nan
    V = rng.standard_normal(lay.gshape + (3,))
    nut0 = 0.021
    nut = np.full(lay.gshape, nut0)
    nut[bad] = np.nan
    tau_i = rng.standard_normal((3, lay.mz, lay.my))
    tau_j = rng.standard_normal((3, lay.mz, lay.mx))
    bcn = [C.BC_PERIODIC if periodic[a // 2] else C.BC_ZEROGRAD for a in range(6)]
    params = C.default_params(periodic, [C.BC_PERIODIC if periodic[a // 2] else C.BC_ZEROGRAD for a in range(6)],
                              bcn, div_scheme=div_scheme, les_model=C.LES_SMAGORINSKY, nu=0.013)
    params = dict(params, bc_u=tuple(bc_u), tau_wall_i=tau_i.ctypes.data, tau_wall_j=tau_j.ctypes.data)
    return lay, nm, S, vol, gxi, u, B, V, nut0, nut, tau_i, tau_j, params


def expected_rhs(lay, periodic, S, vol, gxi, u, B, V, nu, nut0, tau_i, tau_j, bc_u, quick, mode, scale):
    """FormU written from the facts in the module docstring, face by face."""
    m = _m(lay)
    Gm = sum(np.outer(B[q], gxi[q]) for q in range(3))          # grad u [m][n]
    T = Gm + Gm.T
    wpos = lagrange_weights([-1, 0, 1], 0.5)
    wneg = lagrange_weights([0, 1, 2], 0.5)

    def U(i, j, k):
        return u[k + G, j + G, i + G]

    def face_flux(a, x):
        """flux (div + visc) at face x (DMDA index of its left cell) along a"""
        f = x[a]
        if periodic[a] and f == 0:
            x = list(x)
            x[a] = m[a] - 2
            f = m[a] - 2
        Vf = V[x[2] + G, x[1] + G, x[0] + G, a]
        out = np.zeros(3)
        if mode != 2:
            def cell(s):
                y = list(x)
                y[a] = x[a] + s
                return U(*y)
            if quick:
                if Vf > 0:
                    s3 = (-1, 0, 1)
                    if not periodic[a] and f - 1 < 0:
                        s3 = (0, 0, 1)                      # u_UU := u_U (left boundary face)
                    uf = sum(w * cell(s) for w, s in zip(wpos, s3))
                else:
                    s3 = (0, 1, 2)
                    if not periodic[a] and f + 2 > m[a] - 1:
                        s3 = (0, 1, 1)                      # u_UU := u_U (right boundary face)
                    uf = sum(w * cell(s) for w, s in zip(wneg, s3))
            else:
                uf = 0.5 * (cell(0) + cell(1))
            out += -Vf * uf
        if mode != 1:
            side = 0 if f == 0 else (1 if f == m[a] - 2 else -1)
            kind = bc_u[2 * a + side] if side >= 0 and not periodic[a] else None
            nueff = nu + nut0
            base = np.zeros(3)
            if kind == C.BCU_WALLFUNCTION and a < 2:
                tau = tau_i[:, x[2], x[1]] if a == 0 else tau_j[:, x[2], x[0]]
                base = -tau
                nueff = nu
            elif kind == C.BCU_SLIP:
                nueff = 0.0
            out += base + nueff * (T @ S[a])
        return out

    # fp of the interior cells
    fp = {}
    for k, j, i in itertools.product(range(1, m[2] - 1), range(1, m[1] - 1), range(1, m[0] - 1)):
        x = (i, j, k)
        s = np.zeros(3)
        for a in range(3):
            lo = list(x)
            lo[a] -= 1
            s += face_flux(a, x) - face_flux(a, tuple(lo))
        fp[x] = s

    def fp_at(x):
        y = list(x)
        for a in range(3):
            if y[a] == m[a] - 1:
                y[a] = 1 if periodic[a] else m[a] - 2
        return fp[tuple(y)]

    rhs = np.zeros((lay.mz, lay.my, lay.mx, 3))
    for (i, j, k), f0 in fp.items():
        for a in range(3):
            up = [i, j, k]
            up[a] += 1
            rhs[k, j, i, a] = scale / vol * 0.5 * (S[a] @ f0 + S[a] @ fp_at(tuple(up)))
    return rhs


BCS = {
    "walls": (C.BCU_WALLFUNCTION, C.BCU_SLIP, C.BCU_WALLFUNCTION, C.BCU_SLIP, C.BCU_SLIP, C.BCU_WALLFUNCTION),
    "swapped": (C.BCU_SLIP, C.BCU_WALLFUNCTION, C.BCU_SLIP, C.BCU_WALLFUNCTION, C.BCU_WALLFUNCTION, C.BCU_SLIP),
    "generic": (C.BCU_NOSLIP, C.BCU_ZEROGRADIENT, C.BCU_FIXEDVALUE, C.BCU_NOSLIP, C.BCU_GENERIC, C.BCU_ZEROGRADIENT),
}
PERS = [(0, 0, 0), (2, 0, 0), (0, 1, 0), (0, 0, 2), (1, 0, 2)]


@pytest.mark.parametrize("geom", sorted(AFFINE))
@pytest.mark.parametrize("periodic", PERS)
@pytest.mark.parametrize("scheme", [C.DIV_CENTRAL, C.DIV_QUICK])
@pytest.mark.parametrize("bcs", sorted(BCS))
@pytest.mark.parametrize("mode", [0, 1, 2])
def test_formu_bounded_linear(geom, periodic, scheme, bcs, mode):
    bc_u = tuple(C.BCU_PERIODIC if periodic[q // 2] else BCS[bcs][q] for q in range(6))
    lay, nm, S, vol, gxi, u, B, V, nut0, nut, tau_i, tau_j, params = bounded_case(geom, periodic, scheme, bc_u, 31)
    rng = np.random.Generator(np.random.PCG64(77))
    rhs0 = rng.standard_normal((lay.mz, lay.my, lay.mx, 3))
    got = rhs0.copy()
    O.formu(lay.info(), nm, params, V, u, nut, got, 0.7, mode)
    d = got - rhs0
    assert np.isfinite(d).all(), "FormU read a point the reference's boundary rules never read"
    e = expected_rhs(lay, periodic, S, vol, gxi, u, B, V, params["nu"], nut0, tau_i, tau_j, bc_u,
                     scheme == C.DIV_QUICK, mode, 0.7)
    inner = (slice(1, -1),) * 3
    assert np.abs(d[~_interior_mask(lay)]).max() == 0.0          # only the interior cells' faces change
    err = relerr(d[inner], e[inner])
    assert err <= 1e-12, err


def _interior_mask(lay):
    mk = np.zeros((lay.mz, lay.my, lay.mx), bool)
    mk[1:-1, 1:-1, 1:-1] = True
    return mk


@pytest.mark.parametrize("periodic", [(0, 0, 0), (1, 2, 0), (2, 1, 1)])
@pytest.mark.parametrize("scheme", [C.DIV_CENTRAL, C.DIV_QUICK])
def test_face_dot_uses_cell_centred_bases(periodic, scheme):
    lay, mesh = C.mesh_metrics(random_nodes((7, 6, 5), 5, amp=0.2), periodic)
    nm = np_mesh(mesh)
    m = _m(lay)
    rng = np.random.Generator(np.random.PCG64(8))
    u0 = np.array([0.7, -1.3, 0.4])
    u = np.broadcast_to(u0, lay.gshape + (3,)).copy()
    V = rng.standard_normal(lay.gshape + (3,))
    nut = np.zeros(lay.gshape)
    bcn = [C.BC_PERIODIC if periodic[a // 2] else C.BC_ZEROGRAD for a in range(6)]
    params = C.default_params(periodic, bcn, bcn, div_scheme=scheme, les_active=0, nu=0.01)
    # on the perturbed lattice cell and face vectors differ (otherwise the pin would not separate them)
    own = (slice(G + 1, G + lay.mz - 1), slice(G + 1, G + lay.my - 1), slice(G + 1, G + lay.mx - 1))
    assert np.abs(nm["icsi"][own] - nm["csi"][own]).max() > 1e-3
    got = np.zeros((lay.mz, lay.my, lay.mx, 3))
    O.formu(lay.info(), nm, params, V, u, nut, got, 1.3, 1)

    def Vf(a, x):
        y = list(x)
        if periodic[a] and y[a] == 0:
            y[a] = m[a] - 2
        return V[y[2] + G, y[1] + G, y[0] + G, a]

    fp = {}
    for k, j, i in itertools.product(range(1, m[2] - 1), range(1, m[1] - 1), range(1, m[0] - 1)):
        div = 0.0
        for a in range(3):
            lo = [i, j, k]
            lo[a] -= 1
            div += Vf(a, (i, j, k)) - Vf(a, tuple(lo))
        fp[(i, j, k)] = -u0 * div                    # uniform u: every face value is u0
    names = ("csi", "eta", "zet")
    jac = ("iaj", "jaj", "kaj")
    expect = np.zeros_like(got)
    for (i, j, k), f0 in fp.items():
        for a in range(3):
            up = [i, j, k]
            up[a] += 1
            src = list(up)
            if up[a] == m[a] - 1:
                src[a] = 1 if periodic[a] else m[a] - 2
            f1 = fp[tuple(src)]
            cl = nm[names[a]][k + G, j + G, i + G]
            cr = nm[names[a]][up[2] + G, up[1] + G, up[0] + G]   # the right cell's own basis (ghost too)
            expect[k, j, i, a] = 1.3 * nm[jac[a]][k + G, j + G, i + G] * 0.5 * (cl @ f0 + cr @ f1)
    assert relerr(got, expect) <= 1e-13, relerr(got, expect)
