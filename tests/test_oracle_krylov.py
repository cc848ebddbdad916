"""Pins of the oracle's GMRES restatement (PETSc 3.15 KSPGMRES as kspIMEX runs it) and of the IMEX
predictors built on it (oracle_gmres_dense, oracle_imex_solve, oracle_ueqn_rk3cn, oracle_ueqn_abcn;
src/ueqn.c:217-282, src/solvers/ueqn_solvers.c:1508-1737), from facts outside their code:

  - Krylov theory: on a diagonalisable matrix with k distinct eigenvalues, full GMRES reaches the
    exact solution in at most k iterations; with b an eigenvector it stops after one (happy
    breakdown); from the exact solution it takes none;
  - the returned x agrees with numpy.linalg.solve, its true residual is below atol, and the
    GMRES residual estimate equals the true residual norm (no preconditioner);
  - restarts converge too (more iterations), the iteration limit is reported as reason -3;
  - the IMEX solve returns x with || x - coef Visc(x) - b || <= atol, checked with the oracle's
    IMEX operator;
  - both IMEX Runge-Kutta schemes are consistent: on a periodic free stream with uniform sources
    (no convection or diffusion), three stages advance U0 by exactly dt times the source, which
    requires the stage weights to sum to one; ABCN likewise, by 2 dt times it on a first step
    (the reference's doubled explicit term).
"""
import numpy as np
import pytest
import torch

from tosca_amd import case as C
from oracle import oracle as O
from tests._util import np_mesh, random_nodes
from tests.test_oracle_pins import AFFINE, affine_nodes

G = 3


def _diag_problem(n=40, k=5, seed=1):
    rng = np.random.Generator(np.random.PCG64(seed))
    vals = np.linspace(1.0, 3.0, k)
    d = vals[rng.integers(0, k, n)]
    d[:k] = vals                                   # every eigenvalue present
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    A = Q @ np.diag(d) @ Q.T
    b = rng.standard_normal(n)
    return A, b


def test_distinct_eigenvalues_bound_the_iterations():
    A, b = _diag_problem()
    x = np.zeros(len(b))
    its, res, reason = O.gmres_dense(A, b, x, restart=30, atol=1e-11)
    assert reason > 0 and its <= 5
    assert np.abs(x - np.linalg.solve(A, b)).max() <= 1e-9


def test_random_nonsymmetric_system():
    rng = np.random.Generator(np.random.PCG64(7))
    n = 60
    A = np.eye(n) + 0.4 * rng.standard_normal((n, n)) / np.sqrt(n)
    b = rng.standard_normal(n)
    x = 0.1 * rng.standard_normal(n)
    its, res, reason = O.gmres_dense(A, b, x, restart=30, atol=1e-10)
    true = np.linalg.norm(b - A @ x)
    assert reason > 0 and true <= 1e-10 * 1.001
    assert abs(res - true) <= 1e-12 * np.linalg.norm(b)
    assert np.abs(x - np.linalg.solve(A, b)).max() <= 1e-9


def test_restarts_and_iteration_limit():
    rng = np.random.Generator(np.random.PCG64(8))
    n = 50
    A = np.eye(n) + 0.6 * rng.standard_normal((n, n)) / np.sqrt(n)
    b = rng.standard_normal(n)
    x_full = np.zeros(n)
    its_full, _, r_full = O.gmres_dense(A, b, x_full, restart=50, atol=1e-10)
    x_rst = np.zeros(n)
    its_rst, _, r_rst = O.gmres_dense(A, b, x_rst, restart=4, atol=1e-10)
    assert r_full > 0 and r_rst > 0 and its_rst > its_full
    assert np.abs(x_rst - x_full).max() <= 1e-8
    x = np.zeros(n)
    its, res, reason = O.gmres_dense(A, b, x, restart=4, atol=1e-10, maxit=3)
    assert reason == -3 and its == 3


def test_breakdowns_and_exact_guess():
    A, _ = _diag_problem()
    w, Q = np.linalg.eigh(A)
    b = Q[:, 0] * 2.0
    x = np.zeros(len(b))
    its, res, reason = O.gmres_dense(A, b, x, atol=1e-12)
    assert its == 1 and reason > 0
    assert np.abs(x - b / w[0]).max() <= 1e-12
    x2 = np.linalg.solve(A, b)
    its, res, reason = O.gmres_dense(A, b, x2, atol=1e-10)
    assert its == 0 and reason > 0


def _residual_history(A, b, x0, kmax):
    """GMRES residual estimates after 1..kmax iterations of one cycle (the iteration limit stops
    each run; atol = rtol = 0 never stops it earlier)."""
    out = []
    for k in range(1, kmax + 1):
        x = x0.copy()
        its, res, reason = O.gmres_dense(A, b, x, restart=kmax + 5, rtol=0.0, atol=0.0, maxit=k)
        assert its == k and reason == -3
        out.append(res)
    return out


def test_relative_tolerance_is_taken_against_the_rhs_norm():
    """kspIMEX has a nonzero initial guess (src/ueqn.c:276), so PETSc's KSPConvergedDefault measures
    rtol against ||b||, not against ||r0||.  With a guess far from the solution (||r0|| = 300 ||b||)
    the two rules stop at different iterations; the solve must stop at the first iteration whose
    residual estimate is <= rtol ||b||, with reason KSP_CONVERGED_RTOL (2)."""
    rng = np.random.Generator(np.random.PCG64(21))
    n = 40
    A = np.eye(n) + 0.5 * rng.standard_normal((n, n)) / np.sqrt(n)
    b = rng.standard_normal(n)
    x0 = 300.0 * np.linalg.norm(b) / np.sqrt(n) * rng.standard_normal(n)
    hist = _residual_history(A, b, x0, 12)
    bn, r0 = np.linalg.norm(b), np.linalg.norm(b - A @ x0)
    assert r0 > 100 * bn
    k = 6
    rtol = hist[k - 1] / bn * (1 + 1e-9)                 # met first at iteration k by ||b||
    assert hist[k - 2] > rtol * bn
    k_r0 = next(i + 1 for i, r in enumerate(hist) if r <= rtol * r0)
    assert k_r0 < k                                      # the ||r0|| rule would stop earlier
    x = x0.copy()
    its, res, reason = O.gmres_dense(A, b, x, restart=30, rtol=rtol, atol=0.0)
    assert its == k and reason == 2
    assert np.linalg.norm(b - A @ x) <= rtol * bn * (1 + 1e-6)


def test_atol_reason_and_zero_rhs():
    """Converging under atol reports KSP_CONVERGED_ATOL (3); with b = 0 and a nonzero guess the
    relative test falls back to ||r0||."""
    rng = np.random.Generator(np.random.PCG64(22))
    n = 30
    A = np.eye(n) + 0.4 * rng.standard_normal((n, n)) / np.sqrt(n)
    b = rng.standard_normal(n)
    x = np.zeros(n)
    its, res, reason = O.gmres_dense(A, b, x, restart=30, rtol=1e-30, atol=1e-8)
    assert reason == 3 and res < 1e-8
    z = np.zeros(n)
    x0 = rng.standard_normal(n)
    hist = _residual_history(A, z, x0, 8)
    r0 = np.linalg.norm(A @ x0)
    k = 4
    rtol = hist[k - 1] / r0 * (1 + 1e-9)
    x = x0.copy()
    its, res, reason = O.gmres_dense(A, z, x, restart=30, rtol=rtol, atol=0.0)
    assert its == k and reason == 2


def test_nan_residual_reason():
    """A NaN residual is KSP_DIVERGED_NANORINF (-9)."""
    n = 10
    A = np.eye(n)
    A[3, 3] = np.nan
    x = np.zeros(n)
    its, res, reason = O.gmres_dense(A, np.ones(n), x, restart=5, atol=1e-8)
    assert reason == -9


def _imex_case(periodic, seed):
    lay, mesh = C.mesh_metrics(random_nodes((7, 6, 5), seed), periodic)
    rng = np.random.Generator(np.random.PCG64(400 + seed))
    bcs = [C.BC_PERIODIC if periodic[a // 2] else C.BC_FIXED for a in range(6)]
    bcn = [C.BC_PERIODIC if periodic[a // 2] else C.BC_ZEROGRAD for a in range(6)]
    params = C.default_params(periodic, bcs, bcn, div_scheme=C.DIV_CENTRAL, nu=0.05, les_active=0)
    params = dict(params, bc_u=tuple(C.BCU_PERIODIC if periodic[q // 2] else C.BCU_NOSLIP for q in range(6)))
    return lay, np_mesh(mesh), params, rng


@pytest.mark.parametrize("periodic", [(1, 1, 1), (2, 0, 1), (0, 0, 0)])
def test_imex_solve_residual(periodic):
    lay, mesh, params, rng = _imex_case(periodic, 3 + sum(periodic))
    glob = (lay.mz, lay.my, lay.mx, 3)
    b = rng.standard_normal(glob)
    uo = rng.standard_normal(glob)
    x = b.copy()
    coef = 0.8
    work = lambda: (np.zeros(lay.gshape + (3,)), np.zeros(lay.gshape + (3,)))
    its, res, reason = O.imex_solve(lay.info(), mesh, params, b, x, *work(), np.zeros(lay.gshape), uo, coef,
                                    atol=1e-9)
    assert reason > 0 and its > 1
    ax = np.zeros(glob)
    O.imex_matvec(lay.info(), mesh, params, x, *work(), np.zeros(lay.gshape), uo, coef, ax)
    true = np.linalg.norm((ax - b).ravel())
    assert true <= 1e-9 * 1.001
    assert abs(res - true) <= 1e-10


def _free_stream(lay, mesh, U0):
    S = [mesh[n][G:G + lay.mz, G:G + lay.my, G:G + lay.mx] for n in ("icsi", "jeta", "kzet")]
    return np.stack([S[q] @ U0 for q in range(3)], axis=-1)


@pytest.mark.parametrize("scheme", [0, 1])
@pytest.mark.parametrize("periodic", [(1, 1, 1), (2, 2, 2)])
def test_rk3cn_consistency_on_free_stream(scheme, periodic):
    lay, mesh_t = C.mesh_metrics(affine_nodes((6, 7, 5), AFFINE["sheared"]), periodic)
    mesh = np_mesh(mesh_t)
    bcs = [C.BC_PERIODIC] * 6
    params = C.default_params(periodic, bcs, bcs, div_scheme=C.DIV_CENTRAL, nu=0.05, les_active=0)
    U0, f = np.array([1.3, -0.4, 0.7]), np.array([0.2, 0.5, -0.3])
    glob = (lay.mz, lay.my, lay.mx, 3)
    u0 = _free_stream(lay, mesh, U0)
    loc = torch.zeros(lay.gshape + (3,), dtype=torch.float64)
    C.owned(loc, lay)[...] = torch.as_tensor(u0)
    V = C.local_to_local(loc, lay, periodic).numpy().copy()
    dp = -_free_stream(lay, mesh, f)
    dt = 0.3
    rhs = np.zeros(glob)
    its, res, reason = O.ueqn_rk3cn(lay.info(), mesh, params, scheme, V, np.zeros(lay.gshape + (3,)),
                                    np.zeros(lay.gshape), dp, None, None, 0, None, u0, dt, rhs, atol=1e-12)
    assert reason > 0
    want = _free_stream(lay, mesh, U0 + dt * f)
    got = V[G:G + lay.mz, G:G + lay.my, G:G + lay.mx]
    sl = (slice(1, lay.mz - 1), slice(1, lay.my - 1), slice(1, lay.mx - 1))
    assert np.abs(got[sl] - want[sl]).max() <= 1e-12 * np.abs(want).max()


@pytest.mark.parametrize("ab2", [0, 1])
def test_abcn_on_free_stream(ab2):
    """ABCN on a periodic free stream with a uniform source S.f: no convection, no diffusion, so
    one step gives U0 + dt f, except on the first step (ab2 == 0), where the reference adds the
    explicit right-hand side twice (ueqn_solvers.c:1694-1695): U0 + 2 dt f"""
    periodic = (2, 1, 2)
    lay, mesh_t = C.mesh_metrics(affine_nodes((6, 7, 5), AFFINE["sheared"]), periodic)
    mesh = np_mesh(mesh_t)
    bcs = [C.BC_PERIODIC] * 6
    params = C.default_params(periodic, bcs, bcs, div_scheme=C.DIV_CENTRAL, nu=0.05, les_active=0)
    U0, f = np.array([0.6, 1.1, -0.4]), np.array([-0.2, 0.3, 0.5])
    glob = (lay.mz, lay.my, lay.mx, 3)
    u0 = _free_stream(lay, mesh, U0)
    loc = torch.zeros(lay.gshape + (3,), dtype=torch.float64)
    C.owned(loc, lay)[...] = torch.as_tensor(u0)
    V = C.local_to_local(loc, lay, periodic).numpy().copy()
    dp = -_free_stream(lay, mesh, f)
    dt = 0.2
    conv_o = np.zeros(glob)
    rhs = np.zeros(glob)
    bt = np.zeros(glob) if ab2 else None
    its, res, reason = O.ueqn_abcn(lay.info(), mesh, params, V, np.zeros(lay.gshape + (3,)), np.zeros(lay.gshape), dp,
                                   bt, bt, ab2, None, u0, dt, rhs, conv_o, atol=1e-12)
    assert reason > 0
    want = _free_stream(lay, mesh, U0 + (1 if ab2 else 2) * dt * f)
    got = V[G:G + lay.mz, G:G + lay.my, G:G + lay.mx]
    sl = (slice(1, lay.mz - 1), slice(1, lay.my - 1), slice(1, lay.mx - 1))
    assert np.abs(got[sl] - want[sl]).max() <= 1e-12 * np.abs(want).max()
    assert np.abs(conv_o[sl]).max() <= 1e-12 * np.abs(u0).max()     # the stored convection of a free stream
