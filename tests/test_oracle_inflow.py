"""Pins of the oracle's power-law and mapped inflow-database ghosts of UpdateCartesianBCs
(src/boundary.c:1228-1622) and of the helpers they call (findInterpolationWeights(WithExtrap),
scaleHyperTang{Bot,Top}, src/include/inline.h:5873-5999), against facts outside the oracle's code:
  - the two-point interpolation: it reproduces every linear function at any query point inside the
    table (whichever two points it picks), and outside the table it returns the end value with
    weights 1/2, 1/2.  Its choice of points is not always the two nearest: the first pass compares
    point 1 with itself (tol1 = 0) and never makes it the leader, so on the table 0, 1, ..., 9 the
    query 1.3 interpolates between points 0 and 2 (weights 0.35, 0.65), worked by hand from
    inline.h:5895-5937; the oracle keeps that choice, and 3.25 takes points 3 and 4;
  - the mapped plane (typeU 3): a plane whose entries encode their own indices shows which entry
    each ghost reads: ((j - 1) mod n1) + 1, ((i - 1) mod n2) + 1 inside the periods, the last row /
    column past them; the side-ghost rows and columns of the plane (NaN) are never read;
  - the top blend: far below the blending height (7 (h - cb) / delta <= -35, tanh = -1 exactly)
    the plane value is kept unchanged, far above it the ghost is the height profile, which then
    reproduces a linear profile inside its table and takes its top value above it, and at h = cb
    the ghost is the mean of the two;
  - the interpolated plane (typeU 4): stencils and weights that interpolate a linear plane
    bilinearly reproduce it at the stencil's target point; with spline, v and w come from the
    6-point tables (_2 and _1, checked with one-hot weights); with shift2 the ghost is the mix
    (1 - w) left + w right of the columns y_ids - 1 and y_ids;
  - the power law: log |u| / |uref| against log(h / href) has slope 0.107027 at every cell, and at
    h = href the ghost is uref (1 + noise / 10000 uprime_rms).
"""
import numpy as np
import pytest

from oracle import oracle as O
from tosca_amd import case as C
from tests._util import np_mesh, to_np, random_nodes


def _setup(shape=(5, 9, 8), seed=63):
    lay, mesh_t = C.mesh_metrics(random_nodes(shape, seed), (0, 0, 0))
    params = C.default_params((0, 0, 0), [C.BC_FIXED] * 6, [C.BC_ZEROGRAD] * 6, div_scheme=C.DIV_CENTRAL,
                              les_model=C.LES_DSM, nu=0.01)
    params = dict(params, bc_u=(C.BCU_ZEROGRADIENT, C.BCU_ZEROGRADIENT, C.BCU_ZEROGRADIENT, C.BCU_ZEROGRADIENT,
                                C.BCU_GENERIC, C.BCU_ZEROGRADIENT))
    return lay, np_mesh(mesh_t), params, to_np(mesh_t["cent"])


def _ghost(lay, mesh, params, inlet):
    """The kLeft ghost plane (k = 0) at the interior (j, i), shape (my - 2, mx - 2, 3)."""
    U = np.random.Generator(np.random.PCG64(1)).standard_normal(lay.gshape + (3,))
    O.cartesian_bcs(lay.info(), mesh, params, U, None, None, inlet)
    return U[3, 4:lay.my + 2, 4:lay.mx + 2]


def _cent_with_heights(lay, cent, h):
    """cent with z of the kLeft boundary cells (k = 1) set to h[j - 1, i - 1] (ground at 0)."""
    c = cent.copy()
    c[4, 4:lay.my + 2, 4:lay.mx + 2, 2] = h
    return c


def test_interp_weights_linear_nearest_and_ends():
    rng = np.random.Generator(np.random.PCG64(3))
    for trial in range(200):
        pv = np.sort(rng.uniform(-5.0, 40.0, 10))
        a, b = rng.standard_normal(2)
        p = rng.uniform(pv[0], pv[-1])
        (w0, w1), (l0, l1) = O.interp_weights_extrap(pv, p)
        assert abs(w0 * (a * pv[l0] + b) + w1 * (a * pv[l1] + b) - (a * p + b)) <= 1e-12 * (abs(a) * 40 + abs(b) + 1)
        assert abs(w0 + w1 - 1.0) <= 1e-13
        assert pv[l0] < pv[l1]
    pv = np.linspace(0.0, 9.0, 10)
    (w0, w1), (l0, l1) = O.interp_weights_extrap(pv, 1.3)
    assert (l0, l1) == (0, 2) and abs(w0 - 0.35) < 1e-15 and abs(w1 - 0.65) < 1e-15
    assert O.interp_weights_extrap(pv, 9.0) == ((0.5, 0.5), (9, 9))
    assert O.interp_weights_extrap(pv, 12.0) == ((0.5, 0.5), (9, 9))
    assert O.interp_weights_extrap(pv, -1.0) == ((0.5, 0.5), (0, 0))
    (w0, w1), (l0, l1) = O.interp_weights_extrap(pv, 3.25)
    assert (l0, l1) == (3, 4) and abs(w0 - 0.75) < 1e-15 and abs(w1 - 0.25) < 1e-15


def test_mapped_plane_indices_and_periods():
    lay, mesh, params, cent = _setup()
    n1, n2, p1, p2 = 2, 3, 2, 2
    plane = np.full((n1 + 2, n2 + 2, 3), np.nan)
    for pj in range(1, n1 + 1):
        for pi in range(1, n2 + 1):
            plane[pj, pi] = (pj, pi, 100 * pj + pi)
    inlet = dict(type=C.INLET_MAPPED, cent=cent, n1=n1, n2=n2, prds1=p1, prds2=p2, plane=plane)
    g = _ghost(lay, mesh, params, inlet)
    assert lay.my - 2 > n1 * p1 and lay.mx - 2 > n2 * p2          # the mesh reaches past both periods
    for j in range(1, lay.my - 1):
        for i in range(1, lay.mx - 1):
            jj = (j - 1) % n1 + 1 if j <= n1 * p1 else n1
            ii = (i - 1) % n2 + 1 if i <= n2 * p2 else n2
            assert tuple(g[j - 1, i - 1]) == (jj, ii, 100 * jj + ii), (j, i)


def test_top_blend_limits_and_linear_profile():
    lay, mesh, params, cent = _setup()
    rng = np.random.Generator(np.random.PCG64(5))
    ny, nx = lay.my - 2, lay.mx - 2
    H, delta = 100.0, 10.0
    cb = H - 0.5 * delta
    coords = np.sort(rng.uniform(150.0, 900.0, 10))
    a, b = rng.standard_normal(3), rng.standard_normal(3)
    top_u = [tuple(a * z + b) for z in coords]
    B = np.array([1.5, -2.0, 0.25])
    plane = np.full((3, 3, 3), np.nan)
    plane[1, 1] = B
    base = dict(type=C.INLET_MAPPED, n1=1, n2=1, prds1=ny, prds2=nx, plane=plane, merge1=1,
                avg_top_coords=tuple(coords), avg_top_u=tuple(top_u), avg_top_length=H, avg_top_delta=delta)
    # far below: the plane value exactly
    h = rng.uniform(0.0, cb - 5 * delta, (ny, nx))
    g = _ghost(lay, mesh, params, dict(base, cent=_cent_with_heights(lay, cent, h)))
    assert (g == B).all()
    # far above, inside and above the table: the linear profile, then its top value
    h = rng.uniform(coords[0], coords[-1] + 200.0, (ny, nx))
    h[0, 0] = coords[-1]
    g = _ghost(lay, mesh, params, dict(base, cent=_cent_with_heights(lay, cent, h)))
    exp = np.where((h >= coords[-1])[..., None], a * coords[-1] + b, a[None, None] * h[..., None] + b)
    assert np.abs(g - exp).max() <= 1e-12 * np.abs(exp).max()
    # at the blending height: the mean of the plane value and the profile's lowest value
    h = np.full((ny, nx), cb)
    g = _ghost(lay, mesh, params, dict(base, cent=_cent_with_heights(lay, cent, h)))
    exp = 0.5 * (B + (a * coords[0] + b))
    assert np.abs(g - exp).max() <= 1e-14 * np.abs(exp).max()
    # past the periods the ghost is the profile's top value
    g = _ghost(lay, mesh, params, dict(base, prds1=1, cent=_cent_with_heights(lay, cent, h)))
    assert (g[1:] == np.array(top_u[9])).all()


def _linear_plane(n1, n2, rng):
    A = rng.standard_normal((3, 3))                          # plane[pj, pi] = A[:, 0] pj + A[:, 1] pi + A[:, 2]
    pj, pi = np.meshgrid(np.arange(n1 + 2), np.arange(n2 + 2), indexing="ij")
    return np.stack([A[c, 0] * pj + A[c, 1] * pi + A[c, 2] for c in range(3)], axis=-1), A


def _bilinear_tables(my, mx, n1, n2, rng):
    """Stencils and weights that interpolate bilinearly at a random point (y, x) per (j, i)."""
    y = rng.uniform(0.0, n1 + 1.0 - 1e-9, (my, mx))
    x = rng.uniform(0.0, n2 + 1.0 - 1e-9, (my, mx))
    y0, x0 = np.floor(y).astype(np.int32), np.floor(x).astype(np.int32)
    fy, fx = y - y0, x - x0
    corners = [(0, 0, (1 - fy) * (1 - fx)), (0, 1, (1 - fy) * fx), (1, 0, fy * (1 - fx)), (1, 1, fy * fx)]
    order = rng.permutation(4)
    cells = np.zeros((my, mx, 4, 2), np.int32)
    w = np.zeros((my, mx, 4))
    for s, o in enumerate(order):
        dj, di, wt = corners[o]
        cells[..., s, 0] = y0 + dj
        cells[..., s, 1] = x0 + di
        w[..., s] = wt
    return cells, w, y, x


def test_interp_plane_bilinear_spline_and_shift():
    lay, mesh, params, cent = _setup()
    rng = np.random.Generator(np.random.PCG64(9))
    my, mx, n1, n2 = lay.my, lay.mx, 4, 5
    plane, A = _linear_plane(n1, n2, rng)
    cells, w, y, x = _bilinear_tables(my, mx, n1, n2, rng)
    base = dict(type=C.INLET_INTERP, cent=cent, n1=n1, n2=n2, plane=plane, cells=cells, weights=w)
    g = _ghost(lay, mesh, params, base)
    Y, X = y[1:-1, 1:-1], x[1:-1, 1:-1]
    exp = np.stack([A[c, 0] * Y + A[c, 1] * X + A[c, 2] for c in range(3)], axis=-1)
    assert np.abs(g - exp).max() <= 1e-12 * np.abs(exp).max()
    # spline: v from the _2 tables, w from the _1 tables (one-hot weights pick one plane entry each)
    pick = {}
    tabs = {}
    for s in ("_1", "_2"):
        c6 = np.stack([rng.integers(0, n1 + 2, (my, mx, 6)), rng.integers(0, n2 + 2, (my, mx, 6))], -1).astype(np.int32)
        k = rng.integers(0, 6, (my, mx))
        w6 = np.zeros((my, mx, 6))
        np.put_along_axis(w6, k[..., None], 1.0, axis=-1)
        tabs["cells" + s], tabs["weights" + s] = c6, w6
        pick[s] = np.take_along_axis(c6, k[..., None, None], axis=2)[..., 0, :]
    g = _ghost(lay, mesh, params, dict(base, spline=1, **tabs))
    for comp, s in ((1, "_2"), (2, "_1")):
        pj, pi = pick[s][1:-1, 1:-1, 0], pick[s][1:-1, 1:-1, 1]
        assert (g[..., comp] == plane[pj, pi, comp]).all()
    assert np.abs(g[..., 0] - exp[..., 0]).max() <= 1e-12 * np.abs(exp).max()
    # shift: the ghost of column i mixes the targets of columns y_ids[i] - 1 and y_ids[i]
    ids = rng.integers(1, mx, mx).astype(np.int32)
    yw = rng.uniform(0.0, 1.0, mx)
    g = _ghost(lay, mesh, params, dict(base, shift2=1, y_ids=ids, y_weights=yw))
    i = np.arange(1, mx - 1)
    Ym = (1 - yw[i]) * y[1:-1][:, ids[i] - 1] + yw[i] * y[1:-1][:, ids[i]]
    Xm = (1 - yw[i]) * x[1:-1][:, ids[i] - 1] + yw[i] * x[1:-1][:, ids[i]]
    exp = np.stack([A[c, 0] * Ym + A[c, 1] * Xm + A[c, 2] for c in range(3)], axis=-1)
    assert np.abs(g - exp).max() <= 1e-12 * np.abs(exp).max()


def test_power_law_slope_and_reference_height():
    lay, mesh, params, cent = _setup()
    rng = np.random.Generator(np.random.PCG64(11))
    ny, nx = lay.my - 2, lay.mx - 2
    uref = np.array([4.0, 3.0, 0.5])
    h = rng.uniform(1.0, 500.0, (ny, nx))
    base = dict(type=C.INLET_POWERLAW, href=90.0, uprime_rms=0.1, uref=tuple(uref))
    noise0 = np.zeros((lay.my, lay.mx), np.int32)
    g = _ghost(lay, mesh, params, dict(base, cent=_cent_with_heights(lay, cent, h), noise=noise0))
    ratio = np.linalg.norm(g, axis=-1) / np.linalg.norm(uref)
    slope = np.log(ratio) / np.log(h / 90.0)
    ok = np.abs(np.log(h / 90.0)) > 1e-3
    assert np.abs(slope[ok] - 0.107027).max() < 1e-12
    assert np.abs(g / np.linalg.norm(g, axis=-1, keepdims=True) - uref / np.linalg.norm(uref)).max() < 1e-15
    noise = rng.integers(-10000, 10000, (lay.my, lay.mx)).astype(np.int32)
    g = _ghost(lay, mesh, params, dict(base, cent=_cent_with_heights(lay, cent, np.full((ny, nx), 90.0)), noise=noise))
    exp = uref[None, None] * (1.0 + noise[1:-1, 1:-1, None] / 10000.0 * 0.1)
    assert np.abs(g - exp).max() <= 1e-15 * np.abs(exp).max()


@pytest.mark.parametrize("kind", ["mapped", "interp"])
def test_planted_mutants_fail(kind):
    """A wrong period index (i mod n2 without the n2 case) or swapped spline tables change the ghost
    plane the pins above check."""
    lay, mesh, params, cent = _setup()
    rng = np.random.Generator(np.random.PCG64(13))
    if kind == "mapped":
        n1, n2 = 2, 3
        plane = rng.standard_normal((n1 + 2, n2 + 2, 3))
        inlet = dict(type=C.INLET_MAPPED, cent=cent, n1=n1, n2=n2, prds1=2, prds2=2, plane=plane)
        g = _ghost(lay, mesh, params, inlet)
        wrong = np.array([[plane[j % n1, i % n2] for i in range(1, lay.mx - 1)] for j in range(1, lay.my - 1)])
        assert not np.allclose(g, wrong)
    else:
        n1, n2 = 4, 5
        plane, _ = _linear_plane(n1, n2, rng)
        cells, w, _, _ = _bilinear_tables(lay.my, lay.mx, n1, n2, rng)
        c6 = np.stack([rng.integers(0, n1 + 2, (lay.my, lay.mx, 6)), rng.integers(0, n2 + 2, (lay.my, lay.mx, 6))],
                      -1).astype(np.int32)
        c6b = np.stack([rng.integers(0, n1 + 2, (lay.my, lay.mx, 6)), rng.integers(0, n2 + 2, (lay.my, lay.mx, 6))],
                       -1).astype(np.int32)
        w6 = rng.uniform(0, 1, (lay.my, lay.mx, 6))
        inlet = dict(type=C.INLET_INTERP, cent=cent, n1=n1, n2=n2, plane=plane, cells=cells, weights=w, spline=1,
                     cells_1=c6, weights_1=w6, cells_2=c6b, weights_2=w6)
        g = _ghost(lay, mesh, params, inlet)
        g2 = _ghost(lay, mesh, params, dict(inlet, cells_1=c6b, cells_2=c6))
        assert not np.allclose(g[..., 1:], g2[..., 1:])
