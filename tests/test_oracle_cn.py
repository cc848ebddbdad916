"""Pins of the oracle's Crank-Nicolson residual (oracle_cn_rhs, oracle_snes_func_eval;
src/solvers/ueqn_solvers.c:1087-1272) from facts outside their code:

  - the constant right-hand side is -dP + the AB2 buoyancy weights 3/2, -1/2 + the explicit
    Crank-Nicolson half 1/2 Rhs_o + farm (first step: bTheta with weight 1, no Rhs_o);
  - on the first step (backward Euler) the residual is dt R(u) - u + u_o, with R the oracle's
    ExplicitRhsU of u with the same pressure gradient (pinned on its own);
  - on a periodic free stream with a uniform source S.f, u = u_o + dt bU is a root of the residual
    at every resolved face, for both step kinds (FormU of a free stream vanishes).
"""
import numpy as np
import pytest
import torch

from tosca_amd import case as C
from oracle import oracle as O
from tests._util import np_mesh, random_nodes
from tests.test_oracle_pins import AFFINE, affine_nodes

G = 3


def _to_local(glob, lay, periodic, template=None):
    loc = torch.zeros(lay.gshape + (3,), dtype=torch.float64) if template is None else torch.as_tensor(template.copy())
    C.owned(loc, lay)[...] = torch.as_tensor(glob)
    return C.local_to_local(loc, lay, periodic).numpy().copy()


@pytest.mark.parametrize("ab2", [0, 1])
def test_constant_rhs(ab2):
    lay, _ = C.mesh_metrics(random_nodes((5, 4, 6), 1), (0, 0, 0))
    rng = np.random.Generator(np.random.PCG64(3))
    glob = (lay.mz, lay.my, lay.mx, 3)
    dp, bt, bto, ro, farm = (rng.standard_normal(glob) for _ in range(5))
    params = C.default_params((0, 0, 0), [C.BC_FIXED] * 6, [C.BC_ZEROGRAD] * 6)
    bU = np.full(glob, np.nan)
    O.cn_rhs(lay.info(), params, dp, bt, bto, ab2, ro, farm, bU)
    want = -dp + (1.5 * bt - 0.5 * bto + 0.5 * ro if ab2 else bt) + farm
    assert np.abs(bU - want).max() <= 1e-14 * np.abs(want).max()


@pytest.mark.parametrize("periodic", [(0, 0, 0), (2, 0, 1), (1, 1, 1)])
def test_first_step_residual_is_backward_euler(periodic):
    lay, mesh_t = C.mesh_metrics(random_nodes((7, 6, 5), 4), periodic)
    mesh = np_mesh(mesh_t)
    rng = np.random.Generator(np.random.PCG64(12))
    bcs = [C.BC_PERIODIC if periodic[a // 2] else C.BC_FIXED for a in range(6)]
    bcn = [C.BC_PERIODIC if periodic[a // 2] else C.BC_ZEROGRAD for a in range(6)]
    params = C.default_params(periodic, bcs, bcn, div_scheme=C.DIV_QUICK, nu=0.02, les_model=C.LES_DSM)
    params = dict(params, bc_u=tuple(C.BCU_PERIODIC if periodic[q // 2] else C.BCU_SLIP for q in range(6)))
    glob = (lay.mz, lay.my, lay.mx, 3)
    u, uo, dp = rng.standard_normal(glob), rng.standard_normal(glob), rng.standard_normal(glob)
    nut = rng.uniform(0.0, 0.05, lay.gshape)
    ucat0 = rng.standard_normal(lay.gshape + (3,))
    dt = 0.3
    bU = np.zeros(glob)
    O.cn_rhs(lay.info(), params, dp, None, None, 0, None, None, bU)
    F = np.zeros(glob)
    O.snes_func_eval(lay.info(), mesh, params, u, np.zeros(lay.gshape + (3,)), ucat0.copy(), nut, bU, uo, dt, 0, F)
    R = np.zeros(glob)
    O.explicit_rhs_u(lay.info(), mesh, params, _to_local(u, lay, periodic), ucat0.copy(), nut, dp, None, None, 0,
                     None, uo, R, 0)
    want = dt * R - u + uo
    assert np.abs(F - want).max() <= 1e-13 * np.abs(want).max()


def _free_stream(lay, mesh, U0):
    S = [mesh[n][G:G + lay.mz, G:G + lay.my, G:G + lay.mx] for n in ("icsi", "jeta", "kzet")]
    return np.stack([S[q] @ U0 for q in range(3)], axis=-1)


@pytest.mark.parametrize("ab2", [0, 1])
def test_free_stream_root(ab2):
    periodic = (2, 1, 2)
    lay, mesh_t = C.mesh_metrics(affine_nodes((6, 7, 5), AFFINE["sheared"]), periodic)
    mesh = np_mesh(mesh_t)
    bcs = [C.BC_PERIODIC] * 6
    params = C.default_params(periodic, bcs, bcs, div_scheme=C.DIV_CENTRAL, nu=0.05, les_active=0)
    glob = (lay.mz, lay.my, lay.mx, 3)
    uo = _free_stream(lay, mesh, np.array([0.9, -0.3, 0.4]))
    dp = -_free_stream(lay, mesh, np.array([0.2, 0.1, -0.5]))
    rhs_o = _free_stream(lay, mesh, np.array([-0.1, 0.3, 0.2]))
    bU = np.zeros(glob)
    O.cn_rhs(lay.info(), params, dp, None, None, ab2, rhs_o, None, bU)
    dt = 0.25
    u = uo + dt * bU
    F = np.full(glob, np.nan)
    O.snes_func_eval(lay.info(), mesh, params, u, np.zeros(lay.gshape + (3,)), np.zeros(lay.gshape + (3,)),
                     np.zeros(lay.gshape), bU, uo, dt, ab2, F)
    sl = (slice(1, lay.mz - 1), slice(1, lay.my - 1), slice(1, lay.mx - 1))
    assert np.abs(F[sl]).max() <= 1e-12 * np.abs(u).max()
