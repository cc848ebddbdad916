"""Pins of the oracle's FormT restatement (oracle_formt, src/solvers/teqn_solvers.c:6-592) from facts
that do not depend on its code:

  - advection: WENO3 reproduces a linear field exactly, so on an affine lattice the advective
    budget is -sum_q U^q (grad T . e_q) / V; across a step it is upwind-ENO, the jump lands in the
    one cell downstream of it, on both flow directions, and across a jump in a ramp each face
    takes the smooth candidate's linear reconstruction;
  - diffusion: on an affine (sheared) lattice the discrete operator is exact for quadratics, so
    the viscous budget is kappa lap(T) = 2 kappa tr(Q) for T = x.Qx + b.x, with kappa = nu/Pr
    without LES, div(k_t grad T) with the face-averaged k_t (linear in space) for the models that
    carry one (nu_t ignored), and
    nu_t (1 + 2 l / delta) with the stability length l = min(delta, 7.6 nu_t/delta
    sqrt(T_ref/(9.81 dT/dz dz))) for the others, l = delta when unstable;
  - thetaWallFunction patches: a uniform T with no flow leaves only the wall heat flux, q.n/dz in
    the wall cells, indexed [k][j] (i patches) and [k][i] (j patches), and nothing without LES;
  - formMode 1 + 2 = 0, linear in scale, and a fully periodic domain conserves: sum rhs/aj = 0
    (the periodic face reset makes the fluxes of face 0 and face m-2 one and the same).
"""
import numpy as np
import pytest

from tosca_amd import case as C
from oracle import oracle as O
from tests._util import np_mesh, random_nodes
from tests.test_oracle_pins import AFFINE, affine_nodes

G = 3


def _params(periodic=(0, 0, 0), les=0, model=C.LES_DSM, nu=0.0, t_ref=300.0):
    bcs = [C.BC_PERIODIC if periodic[a // 2] else C.BC_FIXED for a in range(6)]
    bcn = [C.BC_PERIODIC if periodic[a // 2] else C.BC_ZEROGRAD for a in range(6)]
    return C.default_params(periodic, bcs, bcn, les_model=model, nu=nu, t_ref=t_ref, les_active=les)


def _idx(lay):
    """local index arrays (k, j, i) of every local point, ghosts included"""
    k, j, i = np.meshgrid(np.arange(lay.mz + 2 * G) - G, np.arange(lay.my + 2 * G) - G,
                          np.arange(lay.mx + 2 * G) - G, indexing="ij")
    return k, j, i


def _affine(n, geom, periodic=(0, 0, 0)):
    vecs = AFFINE[geom]
    lay, mesh = C.mesh_metrics(affine_nodes(n, vecs), periodic)
    cent = mesh["cent"].numpy()
    ei, ej, ek = vecs
    c0 = cent[G + 1, G + 1, G + 1] - ei - ej - ek
    k, j, i = _idx(lay)
    X = c0 + i[..., None] * ei + j[..., None] * ej + k[..., None] * ek   # cell centres, extended to ghosts
    assert np.abs(X[G + 1:G + lay.mz - 1, G + 1:G + lay.my - 1, G + 1:G + lay.mx - 1] -
                  cent[G + 1:G + lay.mz - 1, G + 1:G + lay.my - 1, G + 1:G + lay.mx - 1]).max() < 1e-12
    return lay, np_mesh(mesh), X, vecs


def _run(lay, mesh, params, ucont, theta, nut=None, kt=None, tp=None, scale=1.0, mode=0):
    gs = lay.gshape
    nut = np.zeros(gs) if nut is None else nut
    kt = np.zeros(gs) if kt is None else kt
    tp = dict(pr=0.7, wall_t=(0, 0, 0, 0), q_wall_i=None, q_wall_j=None) if tp is None else tp
    rhs = np.full((lay.mz, lay.my, lay.mx), np.nan)
    O.formt(lay.info(), mesh, params, tp, ucont, theta, nut, kt, rhs, scale, mode)
    return rhs


def _inner(a, lay, d=2):
    """global cells at distance >= d from every boundary, [k][j][i]"""
    return a[d:lay.mz - d, d:lay.my - d, d:lay.mx - d]


# ------------------------------------------------------------------------------------------------
# advection
# ------------------------------------------------------------------------------------------------
@pytest.mark.parametrize("geom", sorted(AFFINE))
def test_linear_field_advection_exact(geom):
    lay, mesh, X, (ei, ej, ek) = _affine((9, 8, 7), geom)
    grad = np.array([0.7, -1.3, 0.4])
    theta = 300.0 + X @ grad
    U = np.array([0.9, -1.4, 0.6])     # fluxes through the i, j, k faces
    ucont = np.broadcast_to(U, lay.gshape + (3,)).copy()
    rhs = _run(lay, mesh, _params(), ucont, theta, mode=1)
    vol = abs(np.linalg.det(np.stack([ei, ej, ek])))
    want = -(U[0] * grad @ ei + U[1] * grad @ ej + U[2] * grad @ ek) / vol
    got = _inner(rhs, lay)
    assert np.abs(got - want).max() <= 1e-11 * abs(want)


@pytest.mark.parametrize("u", [1.3, -0.8])
def test_step_is_upwind_eno(u):
    """T = 1 for i > s, 0 otherwise, flux u through i faces only: the advective budget times the
    volume is -u in the cell just downstream of the step (s + 1 for u > 0, s for u < 0) and zero
    elsewhere, to the 1e-6 WENO epsilon"""
    lay, mesh, X, (ei, ej, ek) = _affine((7, 12, 6), "box")
    k, j, i = _idx(lay)
    s = 5
    theta = np.where(i > s, 1.0, 0.0)
    ucont = np.zeros(lay.gshape + (3,))
    ucont[..., 0] = u
    rhs = _run(lay, mesh, _params(), ucont, theta, mode=1)
    vol = abs(np.linalg.det(np.stack([ei, ej, ek])))
    row = rhs[1:-1, 1:-1, 2:lay.mx - 2] * vol          # i = 2 .. mx-3
    want = np.zeros(row.shape)
    down = (s + 1) if u > 0 else s
    want[..., down - 2] = -u
    assert np.abs(row - want).max() <= 1e-9


@pytest.mark.parametrize("u", [1.3, -0.8])
def test_ramp_jump_takes_smooth_stencil(u):
    """T = i for i <= s and 1000 + i beyond: at each face one of the two WENO3 candidates is smooth
    and its linear reconstruction (fC + fR)/2 or (3 fC - fL)/2 wins to ~1e-12, so the budget
    times -V/u is the exact cell difference: 1 everywhere but 1001 in the downstream cell"""
    lay, mesh, X, (ei, ej, ek) = _affine((7, 12, 6), "box")
    k, j, i = _idx(lay)
    s = 5
    theta = np.where(i > s, 1000.0 + i, 1.0 * i)
    ucont = np.zeros(lay.gshape + (3,))
    ucont[..., 0] = u
    rhs = _run(lay, mesh, _params(), ucont, theta, mode=1)
    vol = abs(np.linalg.det(np.stack([ei, ej, ek])))
    row = rhs[1:-1, 1:-1, 2:lay.mx - 2] * vol / (-u)
    want = np.ones(row.shape)
    want[..., ((s + 1) if u > 0 else s) - 2] = 1001.0
    assert np.abs(row - want).max() <= 1e-7


# ------------------------------------------------------------------------------------------------
# diffusion
# ------------------------------------------------------------------------------------------------
Q = np.array([[0.30, 0.05, -0.02], [0.05, -0.10, 0.04], [-0.02, 0.04, 0.22]])
B = np.array([0.1, -0.2, 0.3])


def _quadratic(X):
    return 290.0 + np.einsum("...a,ab,...b->...", X, Q, X) + X @ B


@pytest.mark.parametrize("geom", sorted(AFFINE))
def test_laplacian_exact_without_les(geom):
    lay, mesh, X, _ = _affine((9, 8, 7), geom)
    nu, pr = 0.021, 0.7
    ucont = np.zeros(lay.gshape + (3,))
    rhs = _run(lay, mesh, _params(nu=nu), ucont, _quadratic(X), tp=dict(pr=pr, wall_t=(0, 0, 0, 0)), mode=2)
    want = nu / pr * 2.0 * np.trace(Q)
    assert np.abs(_inner(rhs, lay) - want).max() <= 1e-11 * abs(want)


@pytest.mark.parametrize("geom", sorted(AFFINE))
def test_linear_field_no_diffusion_one_sided_edges(geom):
    """A linear T carries a uniform diffusive flux, so every interior budget vanishes, boundary
    cells included; T is NaN on the mesh edges (points with two or more indices on a boundary or
    in the ghosts), which only the two-point one-sided transverse differences of the boundary
    face rows avoid reading"""
    lay, mesh, X, _ = _affine((7, 8, 6), geom)
    k, j, i = _idx(lay)
    nb = sum(((a < 1) | (a > m - 2)).astype(int) for a, m in ((k, lay.mz), (j, lay.my), (i, lay.mx)))
    theta = 300.0 + X @ np.array([0.7, -1.3, 0.4])
    theta[nb >= 2] = np.nan
    nu, pr = 0.021, 0.7
    rhs = _run(lay, mesh, _params(nu=nu), np.zeros(lay.gshape + (3,)), theta, tp=dict(pr=pr, wall_t=(0, 0, 0, 0)),
               mode=2)
    r = rhs[1:-1, 1:-1, 1:-1]
    assert np.isfinite(r).all()
    assert np.abs(r).max() <= 1e-11 * nu / pr * 1.5


@pytest.mark.parametrize("model", [C.LES_DSM, C.LES_DPASD, C.LES_AMD])
def test_kt_models_use_face_averaged_kt(model):
    """k_t linear in space, T quadratic: face-averaged k_t is exact at face centres and the flux is
    quadratic, so the budget is div(k_t grad T) = grad k_t . grad T + k_t lap T at the cell centre"""
    lay, mesh, X, _ = _affine((9, 8, 7), "sheared")
    rng = np.random.Generator(np.random.PCG64(4))
    kvec = np.array([0.002, -0.003, 0.004])
    kt = 0.037 + X @ kvec
    nut = rng.uniform(0.0, 0.2, lay.gshape)             # must not enter
    rhs = _run(lay, mesh, _params(les=1, model=model, nu=0.5), np.zeros(lay.gshape + (3,)), _quadratic(X),
               nut=nut, kt=kt, mode=2)
    Xc = X[G:G + lay.mz, G:G + lay.my, G:G + lay.mx]
    want = (2.0 * Xc @ Q + B) @ kvec + (0.037 + Xc @ kvec) * 2.0 * np.trace(Q)
    assert np.abs(_inner(rhs, lay) - _inner(want, lay)).max() <= 1e-11 * np.abs(want).max()


@pytest.mark.parametrize("s", [-0.02, 0.0004, 0.3])
def test_turbulent_prandtl_closure(s):
    """T = a x^2 + s z on a box lattice (x along k, z along j), uniform nu_t: the budget is
    2 a kappa with kappa = nu_t / Prt(dT/dz dz, delta), delta the cube root of the cell volume"""
    lay, mesh, X, (ei, ej, ek) = _affine((9, 8, 7), "box")
    a, nut0, t_ref = 0.45, 1e-3, 300.0
    theta = 300.0 + a * X[..., 0] ** 2 + s * X[..., 2]
    nut = np.full(lay.gshape, nut0)
    rhs = _run(lay, mesh, _params(les=1, model=C.LES_SMAGORINSKY, nu=0.5, t_ref=t_ref), np.zeros(lay.gshape + (3,)),
               theta, nut=nut, mode=2)
    dz = ej[2]
    delta = abs(np.linalg.det(np.stack([ei, ej, ek]))) ** (1.0 / 3.0)
    if s > 0:      # stable: g . grad T < 0
        l = min(delta, 7.6 * nut0 / delta * np.sqrt(t_ref / (9.81 * s * dz)))
    else:
        l = delta
    kappa = nut0 * (1.0 + 2.0 * l / delta)
    want = 2.0 * a * kappa
    got = _inner(rhs, lay)
    assert np.abs(got - want).max() <= 1e-11 * abs(want)
    if s == 0.3:
        assert l < 0.2 * delta      # the stability length is the active branch
    else:
        assert l == delta


# ------------------------------------------------------------------------------------------------
# wall heat flux
# ------------------------------------------------------------------------------------------------
@pytest.mark.parametrize("les", [1, 0])
def test_wall_heat_flux(les):
    lay, mesh, X, (ei, ej, ek) = _affine((6, 7, 5), "box")
    rng = np.random.Generator(np.random.PCG64(9))
    qi = rng.standard_normal((3, lay.mz, lay.my))
    qj = rng.standard_normal((3, lay.mz, lay.mx))
    tp = dict(pr=0.7, wall_t=(1, 1, 1, 1), q_wall_i=qi, q_wall_j=qj)
    theta = np.full(lay.gshape, 301.0)
    rhs = _run(lay, mesh, _params(les=les, model=C.LES_DSM, nu=0.01), np.zeros(lay.gshape + (3,)), theta, tp=tp)
    r = rhs[1:-1, 1:-1, 1:-1]
    if not les:
        assert np.abs(r).max() == 0.0
        return
    dy, dz = ei[1], ej[2]
    want = np.zeros_like(r)
    kk = np.arange(1, lay.mz - 1)
    # i walls (normal along y): q_wall_i[:, k, j], cells i = 1 and i = mx-2
    want[:, :, 0] += qi[1][kk][:, 1:-1] / dy
    want[:, :, -1] += qi[1][kk][:, 1:-1] / dy
    # j walls (normal along z): q_wall_j[:, k, i], cells j = 1 and j = my-2
    want[:, 0, :] += qj[2][kk][:, 1:-1] / dz
    want[:, -1, :] += qj[2][kk][:, 1:-1] / dz
    assert np.abs(r - want).max() <= 1e-13 * np.abs(want).max()


# ------------------------------------------------------------------------------------------------
# structure
# ------------------------------------------------------------------------------------------------
def _random(periodic, seed, les=1, model=C.LES_DSM):
    lay, mesh = C.mesh_metrics(random_nodes((7, 6, 5), seed), periodic)
    rng = np.random.Generator(np.random.PCG64(50 + seed))
    gs = lay.gshape
    f = dict(ucont=rng.standard_normal(gs + (3,)), theta=300.0 + rng.standard_normal(gs),
             nut=rng.uniform(0.0, 0.1, gs), kt=rng.uniform(0.0, 0.1, gs))
    return lay, np_mesh(mesh), f, _params(periodic, les=les, model=model, nu=0.01)


@pytest.mark.parametrize("model", [C.LES_DSM, C.LES_SMAGORINSKY])
def test_form_modes_and_scale(model):
    lay, mesh, f, params = _random((2, 0, 1), 1, model=model)
    r = {m: _run(lay, mesh, params, f["ucont"], f["theta"], f["nut"], f["kt"], mode=m) for m in (0, 1, 2)}
    r5 = _run(lay, mesh, params, f["ucont"], f["theta"], f["nut"], f["kt"], scale=-2.5)
    ins = (slice(1, -1),) * 3
    assert np.abs(r[1] + r[2] - r[0])[ins].max() <= 1e-13 * np.abs(r[0][ins]).max()
    assert np.abs(r5 + 2.5 * r[0])[ins].max() <= 1e-14 * np.abs(r5[ins]).max()
    assert np.abs(r[1][ins]).max() > 0 and np.abs(r[2][ins]).max() > 0
    # boundary cells of rhs are left untouched
    assert np.isnan(r[0][0]).all() and np.isnan(r[0][:, :, -1]).all()


@pytest.mark.parametrize("periodic", [(1, 1, 1), (2, 2, 2), (1, 2, 1), (2, 1, 2)])
@pytest.mark.parametrize("les", [0, 1])
def test_periodic_conservation(periodic, les):
    lay, mesh, f, params = _random(periodic, 7 + sum(periodic), les=les)
    rhs = _run(lay, mesh, params, f["ucont"], f["theta"], f["nut"], f["kt"])
    vol_rhs = rhs[1:-1, 1:-1, 1:-1] / mesh["aj"][G + 1:G + lay.mz - 1, G + 1:G + lay.my - 1, G + 1:G + lay.mx - 1]
    assert abs(vol_rhs.sum()) <= 1e-12 * np.abs(vol_rhs).sum()


def test_unsupported_models_rejected():
    lay, mesh, f, params = _random((0, 0, 0), 3)
    for model in (C.LES_STABILITY, C.LES_BV, C.LES_BAMD):
        with pytest.raises(RuntimeError):
            _run(lay, mesh, dict(params, les_model=model), f["ucont"], f["theta"], f["nut"], f["kt"])
