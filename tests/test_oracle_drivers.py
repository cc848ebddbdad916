"""Pins of the oracle's explicit momentum predictors and IMEX operator (oracle_ueqn_fe,
oracle_ueqn_rk4, oracle_imex_matvec; src/solvers/ueqn_solvers.c:1411-1504, 1739-1801) from facts
outside their code:

  - with the temperature, wind-farm and pressure sources off, no LES and form mode 2, ExplicitRhsU
    is a linear map A of the global Ucont (viscous FormU of the converted velocity, with face
    resets that are fixed masks for noSlip/slip walls and a fixed Ucont_o).  On such a linear
    autonomous system forward Euler is I + dt A and the classical Runge-Kutta step is exactly the
    Taylor polynomial sum_{n<=4} (dt A)^n / n!; A^n u is built here by repeated single calls of
    the oracle's ExplicitRhsU, so a wrong stage weight or stage time fails;
  - constant sources on top only shift the result by dt times the source (RK4 weights sum to 1)
    when A of the shift vanishes: a uniform free stream with sources S.f on a periodic box;
  - the IMEX operator is the identity at coef = 0, affine in coef with slope -Visc, Visc equal to
    the form-mode-2 ExplicitRhsU without sources, and the identity on a free stream.
"""
import math

import numpy as np
import pytest
import torch

from tosca_amd import case as C
from oracle import oracle as O
from tests._util import np_mesh, random_nodes

G = 3


def _owned(a, lay):
    return a[G:G + lay.mz, G:G + lay.my, G:G + lay.mx]


def _to_local(glob, lay, periodic):
    """DMGlobalToLocal into a zero local array: owned copy and the type-2 ghost wrap"""
    loc = torch.zeros(lay.gshape + (3,), dtype=torch.float64)
    C.owned(loc, lay)[...] = torch.as_tensor(glob)
    return C.local_to_local(loc, lay, periodic).numpy().copy()


def _setup(periodic, walls, seed, nu=0.2):
    lay, mesh = C.mesh_metrics(random_nodes((7, 6, 5), seed), periodic)
    rng = np.random.Generator(np.random.PCG64(300 + seed))
    bcs = [C.BC_PERIODIC if periodic[a // 2] else C.BC_FIXED for a in range(6)]
    bcn = [C.BC_PERIODIC if periodic[a // 2] else C.BC_ZEROGRAD for a in range(6)]
    params = C.default_params(periodic, bcs, bcn, div_scheme=C.DIV_CENTRAL, nu=nu, les_active=0)
    bc_u = tuple(C.BCU_PERIODIC if periodic[q // 2] else walls[q] for q in range(6))
    params = dict(params, bc_u=bc_u)
    u0 = rng.standard_normal((lay.mz, lay.my, lay.mx, 3))
    return lay, np_mesh(mesh), params, u0, rng


def _A(lay, mesh, params, periodic, u, uo):
    """one ExplicitRhsU of the global u, no sources, form mode 2"""
    glob = (lay.mz, lay.my, lay.mx, 3)
    V = _to_local(u, lay, periodic)
    U = np.zeros(lay.gshape + (3,))
    rhs = np.zeros(glob)
    O.explicit_rhs_u(lay.info(), mesh, params, V, U, np.zeros(lay.gshape), np.zeros(glob), None, None, 0, None, uo,
                     rhs, 2)
    return rhs


CASES = [((1, 1, 1), None), ((2, 2, 2), None), ((2, 0, 1), (C.BCU_NOSLIP, C.BCU_SLIP)),
         ((0, 0, 0), (C.BCU_SLIP, C.BCU_NOSLIP))]


@pytest.mark.parametrize("periodic,kinds", CASES)
def test_rk4_is_the_taylor_polynomial(periodic, kinds):
    seed = CASES.index((periodic, kinds))
    walls = tuple((kinds or (0, 0))[q % 2] for q in range(6))
    lay, mesh, params, u0, rng = _setup(periodic, walls, seed)
    glob = (lay.mz, lay.my, lay.mx, 3)
    dt = 0.7
    # Taylor polynomial from repeated applications of A
    terms, t = [u0], u0
    for n in range(1, 5):
        t = _A(lay, mesh, params, periodic, t, u0)
        terms.append(t)
    taylor = sum(dt ** n / math.factorial(n) * terms[n] for n in range(5))
    # the size of the highest term: the pin must see it
    assert np.abs(dt ** 4 / 24 * terms[4]).max() > 1e-4 * np.abs(taylor).max()
    V = _to_local(u0, lay, periodic)
    U = np.zeros(lay.gshape + (3,))
    rhs, utmp = np.zeros(glob), np.full(glob, np.nan)
    zero = np.zeros(glob)
    O.ueqn_rk4(lay.info(), mesh, params, V, U, np.zeros(lay.gshape), zero, None, None, 0, None, u0, dt, rhs, utmp, 2)
    got = _owned(V, lay)
    assert np.abs(got - taylor).max() <= 1e-12 * np.abs(taylor).max()
    assert np.array_equal(utmp, got)
    # forward Euler: I + dt A
    V = _to_local(u0, lay, periodic)
    O.ueqn_fe(lay.info(), mesh, params, V, np.zeros(lay.gshape + (3,)), np.zeros(lay.gshape), zero, None, None, 0,
              None, u0, dt, rhs, 2)
    fe = u0 + dt * terms[1]
    assert np.abs(_owned(V, lay) - fe).max() <= 1e-13 * np.abs(fe).max()
    assert np.array_equal(rhs, _owned(V, lay))


def _free_stream(lay, mesh, U0):
    """face fluxes S.U0 of a uniform velocity (global) and the velocity itself (local)"""
    S = {q: mesh[n] for q, n in enumerate(("icsi", "jeta", "kzet"))}
    V = np.stack([_owned(S[q], lay) @ U0 for q in range(3)], axis=-1)
    return V


@pytest.mark.parametrize("periodic", [(1, 1, 1), (2, 2, 2)])
def test_constant_sources_on_free_stream(periodic):
    """uniform U0 on a periodic affine box: A(free stream) = 0 and a source S.f keeps the flow a
    free stream, so RK4 and FE both give U0 + dt f in flux form, whatever the stage weights as long
    as they sum to one"""
    from tests.test_oracle_pins import AFFINE, affine_nodes
    lay, mesh_t = C.mesh_metrics(affine_nodes((6, 7, 5), AFFINE["sheared"]), periodic)
    mesh = np_mesh(mesh_t)
    bcs = [C.BC_PERIODIC] * 6
    params = dict(C.default_params(periodic, bcs, bcs, div_scheme=C.DIV_CENTRAL, nu=0.05, les_active=0))
    U0, f, bt = np.array([1.3, -0.4, 0.7]), np.array([0.2, 0.5, -0.3]), np.array([0.0, 0.1, 0.4])
    glob = (lay.mz, lay.my, lay.mx, 3)
    u0 = _free_stream(lay, mesh, U0)
    dp = -_free_stream(lay, mesh, f)              # -dP adds +S.f
    btheta = _free_stream(lay, mesh, bt)
    dt = 0.3
    want = _free_stream(lay, mesh, U0 + dt * (f + bt))
    for driver in ("fe", "rk4"):
        V = _to_local(u0, lay, periodic)
        U = np.zeros(lay.gshape + (3,))
        rhs = np.zeros(glob)
        if driver == "fe":
            O.ueqn_fe(lay.info(), mesh, params, V, U, np.zeros(lay.gshape), dp, btheta, None, 0, None, u0, dt, rhs, 0)
        else:
            O.ueqn_rk4(lay.info(), mesh, params, V, U, np.zeros(lay.gshape), dp, btheta, None, 0, None, u0, dt, rhs,
                       np.zeros(glob), 0)
        got = _owned(V, lay)
        # faces 1 .. m-2; face 0 of a periodic axis is left to the next call's periodic reset and
        # face m-1 is unused (resetNonResolvedCellFaces)
        sl = (slice(1, lay.mz - 1), slice(1, lay.my - 1), slice(1, lay.mx - 1))
        assert np.abs(got[sl] - want[sl]).max() <= 1e-12 * np.abs(want).max(), driver


@pytest.mark.parametrize("periodic,kinds", CASES[1:])
def test_imex_operator(periodic, kinds):
    walls = tuple((kinds or (0, 0))[q % 2] for q in range(6))
    seed = 10 + CASES.index((periodic, kinds))
    lay, mesh, params, v, rng = _setup(periodic, walls, seed)
    glob = (lay.mz, lay.my, lay.mx, 3)
    uo = rng.standard_normal(glob)

    def apply(coef):
        av = np.full(glob, np.nan)
        O.imex_matvec(lay.info(), mesh, params, v, np.zeros(lay.gshape + (3,)), np.zeros(lay.gshape + (3,)),
                      np.zeros(lay.gshape), uo, coef, av)
        return av

    assert np.array_equal(apply(0.0), v)
    a1, a2 = apply(0.4), apply(-1.1)
    visc = _A(lay, mesh, params, periodic, v, uo)
    assert np.abs(a1 - (v - 0.4 * visc)).max() <= 1e-13 * np.abs(v).max()
    assert np.abs(a2 - (v + 1.1 * visc)).max() <= 1e-13 * np.abs(v).max()
    assert np.abs(visc).max() > 1e-3


def test_imex_identity_on_free_stream():
    from tests.test_oracle_pins import AFFINE, affine_nodes
    periodic = (2, 1, 2)
    lay, mesh_t = C.mesh_metrics(affine_nodes((6, 7, 5), AFFINE["sheared"]), periodic)
    mesh = np_mesh(mesh_t)
    bcs = [C.BC_PERIODIC] * 6
    params = C.default_params(periodic, bcs, bcs, div_scheme=C.DIV_CENTRAL, nu=0.05, les_active=0)
    v = _free_stream(lay, mesh, np.array([0.3, 1.2, -0.8]))
    glob = (lay.mz, lay.my, lay.mx, 3)
    av = np.zeros(glob)
    O.imex_matvec(lay.info(), mesh, params, v, np.zeros(lay.gshape + (3,)), np.zeros(lay.gshape + (3,)),
                  np.zeros(lay.gshape), v, 0.9, av)
    sl = (slice(0, lay.mz - 1), slice(0, lay.my - 1), slice(0, lay.mx - 1))
    assert np.abs(av[sl] - v[sl]).max() <= 1e-12 * np.abs(v).max()
