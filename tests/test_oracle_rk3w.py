"""Pins of the oracle's explicit Wray Runge-Kutta predictor (oracle_ueqn_rk3w: UeqnRK3CN_W's
low-storage scheme, src/solvers/ueqn_solvers.c:1508-1583, coefficients :1518-1520, with the whole
right-hand side explicit) and of its tangent (oracle_rk3w_jvp), from facts outside their code:

  - with the sources off, no LES and form mode 2, ExplicitRhsU is A(P u): P the face resets
    (no-penetration on the walls, face 0 <- face m-2 on periodic axes), A linear.  On such a
    system the reset state y = P u follows y' = y + dt (beta_i B y - delta_i B y_prev) with
    B = P A P, and every three-stage third-order Runge-Kutta step, Wray's included, is exactly the
    cubic Taylor polynomial sum_{n<=3} (dt B)^n / n!; ten stages are three such steps followed by
    the first stage of a fourth, I + 8/15 dt B.  B^n y is built by repeated single ExplicitRhsU
    calls, so a wrong beta, delta or stage order fails;
  - a uniform free stream with constant sources S.f on a periodic box advances by dt f per step
    (the stage weights beta - delta sum to one): 10 stages give 3 + 8/15 steps;
  - the tangent's value channel is the double oracle bit for bit, and its tangent matches central
    finite differences of the double oracle in the initial Ucont and in nu_t (LES on, both schemes,
    periodic and walled axes).
"""
import math

import numpy as np
import pytest
import torch

from tosca_amd import case as C
from oracle import oracle as O
from tests._util import np_mesh, random_nodes, random_case

G = 3


def _owned(a, lay):
    return a[G:G + lay.mz, G:G + lay.my, G:G + lay.mx]


def _to_local(glob, lay, periodic):
    loc = torch.zeros(lay.gshape + (3,), dtype=torch.float64)
    C.owned(loc, lay)[...] = torch.as_tensor(glob)
    return C.local_to_local(loc, lay, periodic).numpy().copy()


def _P(u, lay, periodic, bc_u):
    """the face resets of ExplicitRhsU on a global vector, written from their definition: walls
    (noSlip / slip / wall function) carry no normal flux; face 0 of a periodic axis is face m-2"""
    u = u.copy()
    dims = {0: 2, 1: 1, 2: 0}                     # axis (i, j, k) -> array dim of [k][j][i]
    m = {0: lay.mx, 1: lay.my, 2: lay.mz}
    walls = (C.BCU_NOSLIP, C.BCU_SLIP, C.BCU_WALLFUNCTION)
    for a in range(3):
        d = dims[a]
        for side, idx in ((0, 0), (1, m[a] - 2)):
            if bc_u[2 * a + side] in walls:
                sl = [slice(0, n - 1) for n in (lay.mz, lay.my, lay.mx)] + [a]
                sl[d] = idx
                u[tuple(sl)] = 0.0
    for a in range(3):
        if periodic[a]:
            d = dims[a]
            src = [slice(None)] * 4
            dst = [slice(None)] * 4
            src[d], dst[d] = m[a] - 2, 0
            u[tuple(dst)] = u[tuple(src)]
    return u


def _setup(periodic, walls, seed, nu=0.2):
    lay, mesh = C.mesh_metrics(random_nodes((7, 6, 5), seed), periodic)
    rng = np.random.Generator(np.random.PCG64(600 + seed))
    bcs = [C.BC_PERIODIC if periodic[a // 2] else C.BC_FIXED for a in range(6)]
    bcn = [C.BC_PERIODIC if periodic[a // 2] else C.BC_ZEROGRAD for a in range(6)]
    params = C.default_params(periodic, bcs, bcn, div_scheme=C.DIV_CENTRAL, nu=nu, les_active=0)
    bc_u = tuple(C.BCU_PERIODIC if periodic[q // 2] else walls[q] for q in range(6))
    params = dict(params, bc_u=bc_u)
    u0 = _P(rng.standard_normal((lay.mz, lay.my, lay.mx, 3)), lay, periodic, bc_u)
    return lay, np_mesh(mesh), params, u0


def _A(lay, mesh, params, periodic, u):
    """one ExplicitRhsU of the global u (Ucont_o = u), no sources, form mode 2"""
    glob = (lay.mz, lay.my, lay.mx, 3)
    V = _to_local(u, lay, periodic)
    rhs = np.zeros(glob)
    O.explicit_rhs_u(lay.info(), mesh, params, V, np.zeros(lay.gshape + (3,)), np.zeros(lay.gshape), np.zeros(glob),
                     None, None, 0, None, u, rhs, 2)
    return rhs


CASES = [((1, 1, 1), None), ((2, 2, 2), None), ((2, 0, 1), (C.BCU_NOSLIP, C.BCU_SLIP)),
         ((0, 0, 0), (C.BCU_SLIP, C.BCU_NOSLIP)), ((0, 2, 0), (C.BCU_NOSLIP, C.BCU_NOSLIP))]


def _rk3w(lay, mesh, params, periodic, u0, dt, n_stages, dp=None, btheta=None, form_mode=2):
    glob = (lay.mz, lay.my, lay.mx, 3)
    V = _to_local(u0, lay, periodic)
    uo, rhs, rhs_o = (np.full(glob, np.nan) for _ in range(3))
    O.ueqn_rk3w(lay.info(), mesh, params, V, np.zeros(lay.gshape + (3,)), np.zeros(lay.gshape),
                np.zeros(glob) if dp is None else dp, btheta, None, 0, None, dt, n_stages, uo, rhs, rhs_o, form_mode)
    return _owned(V, lay).copy(), uo, rhs, rhs_o


@pytest.mark.parametrize("periodic,kinds", CASES)
def test_rk3w_is_the_cubic_taylor_polynomial(periodic, kinds):
    seed = CASES.index((periodic, kinds))
    walls = tuple((kinds or (0, 0))[q % 2] for q in range(6))
    lay, mesh, params, y0 = _setup(periodic, walls, seed)
    bc = params["bc_u"]
    P = lambda u: _P(u, lay, periodic, bc)
    B = lambda y: P(_A(lay, mesh, params, periodic, y))
    dt = 0.6
    step = lambda y: sum(dt ** n / math.factorial(n) * t for n, t in enumerate(_powers(B, y, 3)))
    one = step(y0)
    # the cubic term is visible at this dt
    assert np.abs(dt ** 3 / 6 * _powers(B, y0, 3)[3]).max() > 1e-4 * np.abs(one).max()
    got, uo, rhs, rhs_o = _rk3w(lay, mesh, params, periodic, y0, dt, 3)
    assert np.abs(P(got) - one).max() <= 1e-12 * np.abs(one).max()
    assert np.array_equal(uo, y0)            # Ucont_o: the state at the step's start
    assert np.array_equal(rhs_o, rhs)        # Rhs_o: the last stage's Rhs
    # ten stages: three steps and the first stage of a fourth
    y3 = step(step(one))
    want = y3 + dt * 8.0 / 15.0 * B(y3)
    got10, *_ = _rk3w(lay, mesh, params, periodic, y0, dt, 10)
    assert np.abs(P(got10) - want).max() <= 1e-11 * np.abs(want).max()


def _powers(B, y, n):
    out = [y]
    for _ in range(n):
        out.append(B(out[-1]))
    return out


def _free_stream(lay, mesh, U0):
    S = {q: mesh[n] for q, n in enumerate(("icsi", "jeta", "kzet"))}
    return np.stack([_owned(S[q], lay) @ U0 for q in range(3)], axis=-1)


@pytest.mark.parametrize("periodic", [(1, 1, 1), (2, 2, 2)])
def test_rk3w_constant_sources_on_free_stream(periodic):
    from tests.test_oracle_pins import AFFINE, affine_nodes
    lay, mesh_t = C.mesh_metrics(affine_nodes((6, 7, 5), AFFINE["sheared"]), periodic)
    mesh = np_mesh(mesh_t)
    bcs = [C.BC_PERIODIC] * 6
    params = C.default_params(periodic, bcs, bcs, div_scheme=C.DIV_CENTRAL, nu=0.05, les_active=0)
    U0, f, bt = np.array([1.3, -0.4, 0.7]), np.array([0.2, 0.5, -0.3]), np.array([0.0, 0.1, 0.4])
    u0 = _free_stream(lay, mesh, U0)
    dp, btheta = -_free_stream(lay, mesh, f), _free_stream(lay, mesh, bt)
    dt = 0.3
    sl = (slice(1, lay.mz - 1), slice(1, lay.my - 1), slice(1, lay.mx - 1))
    for n_stages, steps in ((3, 1.0), (6, 2.0), (10, 3.0 + 8.0 / 15.0)):
        got, *_ = _rk3w(lay, mesh, params, periodic, u0, dt, n_stages, dp=dp, btheta=btheta, form_mode=0)
        want = _free_stream(lay, mesh, U0 + steps * dt * (f + bt))
        assert np.abs(got[sl] - want[sl]).max() <= 1e-12 * np.abs(want).max(), n_stages


def _les_case(periodic, scheme, seed):
    lay, mesh, f, params = random_case(n=(6, 5, 7), periodic=periodic, seed=seed, div_scheme=scheme,
                                       bcu_nonper=(C.BCU_NOSLIP, C.BCU_SLIP, C.BCU_ZEROGRADIENT))
    rng = np.random.Generator(np.random.PCG64(700 + seed))
    glob = (lay.mz, lay.my, lay.mx, 3)
    x = dict(ucont=f["ucont"].numpy().copy(), ucat=f["ucat"].numpy().copy(), nut=f["nut"].numpy().copy(),
             dp=rng.standard_normal(glob), bt=rng.standard_normal(glob), bto=rng.standard_normal(glob),
             farm=rng.standard_normal(glob))
    return lay, np_mesh(mesh), dict(params, nu=0.05), x, rng


@pytest.mark.parametrize("periodic,scheme", [((2, 0, 1), C.DIV_CENTRAL), ((0, 1, 2), C.DIV_QUICK),
                                             ((0, 0, 0), C.DIV_CENTRAL)])
def test_rk3w_tangent(periodic, scheme):
    lay, mesh, params, x, rng = _les_case(periodic, scheme, 3 + periodic[0])
    info = lay.info()
    glob = (lay.mz, lay.my, lay.mx, 3)
    dt, ns = 0.05, 4
    args = (x["nut"], x["dp"], x["bt"], x["bto"], 1, x["farm"], dt, ns)
    # value channel: the double oracle bit for bit
    uf, _ = O.rk3w_jvp(info, mesh, params, x["ucont"], x["ucat"], *args)
    V, U = x["ucont"].copy(), x["ucat"].copy()
    uo, rhs, rhs_o = (np.zeros(glob) for _ in range(3))
    O.ueqn_rk3w(info, mesh, params, V, U, x["nut"], x["dp"], x["bt"], x["bto"], 1, x["farm"], dt, ns, uo, rhs, rhs_o)
    assert np.array_equal(uf, _owned(V, lay))

    def final(u0, nut):
        V = _to_local(u0, lay, periodic)
        # the caller's ghosts outside periodic axes are kept (never read)
        V = np.where(np.isnan(V), 0.0, V)
        V2 = x["ucont"].copy()
        _owned(V2, lay)[...] = _owned(V, lay)
        u = V2
        uo_, r_, ro_ = (np.zeros(glob) for _ in range(3))
        O.ueqn_rk3w(info, mesh, params, u, x["ucat"].copy(), nut, x["dp"], x["bt"], x["bto"], 1, x["farm"], dt, ns,
                    uo_, r_, ro_)
        return _owned(u, lay).copy()

    u0 = _owned(x["ucont"], lay).copy()
    du = rng.standard_normal(glob)
    dn = rng.standard_normal(lay.gshape) * 0.01
    for name, d_u, d_n in (("ucont", du, None), ("nut", None, dn)):
        _, duf = O.rk3w_jvp(info, mesh, params, x["ucont"], x["ucat"], *args, d_ucont=d_u, d_nut=d_n)
        eps = 1e-6
        if d_u is not None:
            fd = (final(u0 + eps * d_u, x["nut"]) - final(u0 - eps * d_u, x["nut"])) / (2 * eps)
        else:
            fd = (final(u0, x["nut"] + eps * d_n) - final(u0, x["nut"] - eps * d_n)) / (2 * eps)
        err = np.abs(duf - fd).max() / np.abs(duf).max()
        print(f"  rk3w tangent {name}: max|J d - FD| / max|J d| = {err:.2e}")
        assert np.abs(duf).max() > 0
        assert err <= 1e-6, name
