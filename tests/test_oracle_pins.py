"""Pins of the CPU oracle that do not come from the oracle's own code.

The oracle (oracle/tosca_oracle.cpp) restates FormU, UpdateCs, UpdateNut and Buoyancy of the
reference (src/solvers/ueqn_solvers.c, src/les.c, src/sources/ueqn_sources.c).  The reference ships
no fixtures or known-answer tests for these routines, so they are checked here against:

  1. closed forms on affine meshes (a uniform, possibly sheared lattice; every metric constant).
     There the finite-volume operators reduce to textbook statements written below in numpy from
     the calculus:
       - contravariant area vectors S^a = e_b x e_c, volume e_i . (e_j x e_k), and the inverse
         Jacobian from numpy.linalg.inv;
       - central / QUICK face values, with the QUICK weights the Lagrange interpolant through the
         three upwind-biased cells (solved from a Vandermonde system);
       - viscous face flux nu_eff (grad u + grad u^T) . S with grad u from the chain rule;
       - Lilly-Germano dynamic Smagorinsky in Cartesian form: C = 1/2 <L:M> / <M:M>,
         L = <u u> - <u><u>, M = -D~^2 |S^| S^ + D^2 <|S| S>, masked Simpson test filter,
         Simpson box average, clip;
       - one-sided first-order differences at non-periodic edges = numpy.gradient(edge_order=1);
  2. invariants that hold on any mesh: free-stream preservation, formMode 1 + 2 = 0, linearity
     in scale, Cs = nu_t = 0 for uniform velocity, Galilean invariance of the dynamic
     coefficient, Buoyancy = 0 at theta = theta_ref.

A dropped term, a wrong sign, a swapped axis or a transposed tensor in the oracle moves at least
one of these by far more than the tolerances.
"""
from __future__ import annotations

import itertools

import numpy as np
import pytest
import torch

from tosca_amd import case as C
from oracle import oracle as O
from tests._util import np_mesh, to_np, relerr, random_nodes

# ------------------------------------------------------------------------------------------------
# affine meshes
# ------------------------------------------------------------------------------------------------
# lattice vectors e_i, e_j, e_k: physical displacement between neighbouring nodes along i, j, k.
# TOSCA's axes: k <-> x, j <-> z, i <-> y.
AFFINE = {
    "box": (np.array([0.0, 0.8, 0.0]), np.array([0.0, 0.0, 0.6]), np.array([1.1, 0.0, 0.0])),
    "sheared": (np.array([0.15, 0.8, 0.05]), np.array([-0.1, 0.12, 0.6]), np.array([1.1, -0.2, 0.1])),
}


def affine_nodes(n, vecs):
    """Nodes x = i e_i + j e_j + k e_k for (nx, ny, nz) cells along (k, i, j)."""
    nx, ny, nz = n
    ei, ej, ek = vecs
    k = np.arange(nx + 1)[:, None, None, None]
    j = np.arange(nz + 1)[None, :, None, None]
    i = np.arange(ny + 1)[None, None, :, None]
    return torch.as_tensor(i * ei + j * ej + k * ek, dtype=torch.float64)


def affine_geometry(vecs):
    """Area vectors, volume and inverse Jacobian of the lattice (textbook identities)."""
    ei, ej, ek = vecs
    S = np.stack([np.cross(ej, ek), np.cross(ek, ei), np.cross(ei, ej)])   # rows S^i, S^j, S^k
    vol = float(np.dot(ei, np.cross(ej, ek)))
    Jac = np.stack([ei, ej, ek], axis=1)                                    # dx/dxi, columns i, j, k
    grad_xi = np.linalg.inv(Jac)                                            # rows: grad xi_i, xi_j, xi_k
    return S, vol, grad_xi


# interior arrays are [k][j][i]; computational direction a = 0 (i), 1 (j), 2 (k) -> array axis
AX = {0: 2, 1: 1, 2: 0}


def sh(f, a, s):
    """Periodic shift: value at cell c + s along computational direction a."""
    return np.roll(f, -s, axis=AX[a])


def build(n, vecs, periodic, seed, les_model=C.LES_DSM, div_scheme=C.DIV_CENTRAL, u_const=None):
    nodes = affine_nodes(n, vecs)
    lay, mesh = C.mesh_metrics(nodes, periodic)
    rng = np.random.Generator(np.random.PCG64(seed))
    u = torch.zeros(lay.gshape + (3,), dtype=torch.float64)
    if u_const is None:
        C.interior(u, lay)[...] = torch.as_tensor(rng.standard_normal(C.interior(u, lay).shape))
    else:
        C.interior(u, lay)[...] = torch.as_tensor(u_const, dtype=torch.float64)
    bcs = []
    for a in range(3):
        bcs += [C.BC_PERIODIC if periodic[a] else C.BC_ZEROGRAD] * 2
    C.apply_vector_bcs(u, lay, periodic, bcs, [None] * 6, mesh)
    C.reset_cell_periodic(u, lay, periodic)          # edge and corner ghosts, as before FormU
    V = C.ucont_from_ucat(u, mesh, lay, periodic, bcs)
    nut = torch.zeros(lay.gshape, dtype=torch.float64)
    C.interior(nut, lay)[...] = torch.as_tensor(rng.uniform(0.0, 0.05, C.interior(nut, lay).shape))
    bcn = [C.BC_PERIODIC if periodic[a] else C.BC_ZEROGRAD for a in range(3) for _ in range(2)]
    C.apply_scalar_bcs(nut, lay, periodic, bcn, [0.0] * 6)
    params = C.default_params(periodic, bcs, bcn, div_scheme=div_scheme, les_model=les_model, nu=0.013)
    return lay, mesh, u, V, nut, params


def interior_np(t, lay):
    return to_np(C.interior(t, lay))


# ------------------------------------------------------------------------------------------------
# FormU closed form
# ------------------------------------------------------------------------------------------------
def lagrange_weights(nodes, x):
    """Weights w with sum_n w_n f(nodes_n) = p(x), p the interpolating polynomial."""
    V = np.vander(np.asarray(nodes, dtype=np.float64), increasing=True).T
    return np.linalg.solve(V, np.array([x ** p for p in range(len(nodes))]))


def formu_closed_form(u, V, nut, vecs, nu, les, mode, quick, scale, rhs_in):
    """u (K,J,I,3) interior Cartesian velocity on a periodic affine lattice, V (K,J,I,3) the
    contravariant flux at the upper face of each cell, nut (K,J,I).  Returns rhs_in + the FormU
    contribution on the upper face of each cell."""
    S, vol, gxi = affine_geometry(vecs)
    wpos = lagrange_weights([-1, 0, 1], 0.5)     # V > 0: cells c-1, c, c+1
    wneg = lagrange_weights([0, 1, 2], 0.5)      # V <= 0: cells c, c+1, c+2
    fp = np.zeros_like(u)
    for a in range(3):
        F = np.zeros_like(u)
        if mode != 2:
            Va = V[..., a:a + 1]
            if quick:
                up = wpos[0] * sh(u, a, -1) + wpos[1] * u + wpos[2] * sh(u, a, 1)
                dn = wneg[0] * u + wneg[1] * sh(u, a, 1) + wneg[2] * sh(u, a, 2)
                uf = np.where(Va > 0, up, dn)
            else:
                uf = 0.5 * (u + sh(u, a, 1))
            F += -Va * uf
        if mode != 1:
            # face-centred differences along each computational direction q
            d = []
            for q in range(3):
                if q == a:
                    d.append(sh(u, a, 1) - u)
                else:
                    ua = sh(u, a, 1)
                    d.append(0.25 * (sh(u, q, 1) + sh(ua, q, 1) - sh(u, q, -1) - sh(ua, q, -1)))
            # grad u [m][n] = sum_q d_q u_m dxi_q/dx_n
            G = sum(d[q][..., :, None] * gxi[q][None, None, None, None, :] for q in range(3))
            nue = nu + 0.5 * (nut + sh(nut, a, 1)) if les else np.full(nut.shape, nu)
            T = G + np.swapaxes(G, -1, -2)
            F += nue[..., None] * np.einsum("...mn,n->...m", T, S[a])
        fp += F - sh(F, a, -1)
    rhs = rhs_in.copy()
    for a in range(3):
        dot = fp @ S[a]
        rhs[..., a] += scale / vol * 0.5 * (dot + sh(dot, a, 1))
    return rhs


PER_ALL = [(C.PERIODIC_WRAP,) * 3, (C.PERIODIC_COPY,) * 3, (C.PERIODIC_COPY, C.PERIODIC_WRAP, C.PERIODIC_COPY)]


@pytest.mark.parametrize("geom", sorted(AFFINE))
@pytest.mark.parametrize("periodic", PER_ALL)
@pytest.mark.parametrize("scheme", [C.DIV_CENTRAL, C.DIV_QUICK])
@pytest.mark.parametrize("mode", [0, 1, 2])
def test_formu_closed_form(geom, periodic, scheme, mode):
    vecs = AFFINE[geom]
    n = (7, 6, 5)
    lay, mesh, u, V, nut, params = build(n, vecs, periodic, seed=3, div_scheme=scheme)
    S, vol, _ = affine_geometry(vecs)
    # the mesh metrics are the lattice's, everywhere the FormU stencil reads them
    for name, a in (("icsi", 0), ("jeta", 1), ("kzet", 2), ("csi", 0), ("eta", 1), ("zet", 2)):
        assert np.abs(interior_np(mesh[name], lay) - S[a]).max() <= 1e-13
    assert np.abs(interior_np(mesh["iaj"], lay) * vol - 1).max() <= 1e-13
    ui, Vi, nuti = interior_np(u, lay), interior_np(V, lay), interior_np(nut, lay)
    # the contravariant flux handed to FormU is S . (u_L + u_R) / 2
    for a in range(3):
        assert np.abs(Vi[..., a] - 0.5 * (ui + sh(ui, a, 1)) @ S[a]).max() <= 1e-13
    rng = np.random.Generator(np.random.PCG64(9))
    rhs = rng.standard_normal((lay.mz, lay.my, lay.mx, 3))
    expect = formu_closed_form(ui, Vi, nuti, vecs, params["nu"], True, mode, scheme == C.DIV_QUICK, 0.7,
                               rhs[1:-1, 1:-1, 1:-1])
    got = rhs.copy()
    O.formu(lay.info(), np_mesh(mesh), params, to_np(V), to_np(u), to_np(nut), got, 0.7, mode)
    # faces outside the interior cells are not touched
    untouched = np.ones(got.shape[:3], bool)
    untouched[1:-1, 1:-1, 1:-1] = False
    assert np.array_equal(got[untouched], rhs[untouched])
    d, e = got[1:-1, 1:-1, 1:-1] - rhs[1:-1, 1:-1, 1:-1], expect - rhs[1:-1, 1:-1, 1:-1]
    assert relerr(d, e) <= 1e-12, relerr(d, e)


def test_formu_closed_form_les_off():
    """nu_eff = nu when LES is off: nu_t must not enter the viscous flux."""
    vecs = AFFINE["sheared"]
    lay, mesh, u, V, nut, params = build((6, 5, 7), vecs, PER_ALL[0], seed=4)
    params = dict(params, les_active=0)
    rhs = np.zeros((lay.mz, lay.my, lay.mx, 3))
    expect = formu_closed_form(interior_np(u, lay), interior_np(V, lay), interior_np(nut, lay), vecs,
                               params["nu"], False, 0, False, 1.0, rhs[1:-1, 1:-1, 1:-1])
    O.formu(lay.info(), np_mesh(mesh), params, to_np(V), to_np(u), to_np(nut), rhs, 1.0, 0)
    assert relerr(rhs[1:-1, 1:-1, 1:-1], expect) <= 1e-12


def test_tgv_closed_form():
    """The Taylor-Green case of the benchmark suite (uniform periodic 2 pi box, central scheme)."""
    case = C.tgv_case(n=12, ptype=C.PERIODIC_WRAP)
    lay = case.lay
    h = 2 * np.pi / 12
    vecs = (np.array([0, h, 0.0]), np.array([0, 0, h]), np.array([h, 0, 0.0]))
    rhs = np.zeros((lay.mz, lay.my, lay.mx, 3))
    nut = np.zeros(lay.gshape)
    ui = interior_np(case.ucat, lay)
    # TGV: u = sin x cos y cos z, v = -cos x sin y cos z, w = 0 at the cell centres
    cent = interior_np(case.mesh["cent"], lay)
    x, y, z = cent[..., 0], cent[..., 1], cent[..., 2]
    assert np.abs(ui[..., 0] - np.sin(x) * np.cos(y) * np.cos(z)).max() <= 1e-14
    assert np.abs(ui[..., 1] + np.cos(x) * np.sin(y) * np.cos(z)).max() <= 1e-14
    assert np.abs(ui[..., 2]).max() == 0.0
    expect = formu_closed_form(ui, interior_np(case.ucont, lay), np.zeros(ui.shape[:3]), vecs,
                               case.params["nu"], True, 0, False, 1.0, rhs[1:-1, 1:-1, 1:-1])
    O.formu(lay.info(), np_mesh(case.mesh), case.params, to_np(case.ucont), to_np(case.ucat), nut, rhs, 1.0, 0)
    assert relerr(rhs[1:-1, 1:-1, 1:-1], expect) <= 1e-12
    # the discrete viscous operator on sin: nu * (2 - 2 cos h) / h^2 per direction (closed form);
    # on the w = 0 component of TGV the z-momentum is zero to round-off
    assert np.abs(rhs[1:-1, 1:-1, 1:-1, 1]).max() <= 1e-13


def test_discrete_laplacian_of_sine():
    """Viscous-only FormU on u = (0, 0, sin(2 pi x / L)) in a uniform periodic box: the face-dot of
    the discrete Laplacian, nu * -(2 - 2 cos(k h)) / h^2 * sin, averaged to the j face."""
    n, L = 16, 2 * np.pi
    h = L / n
    nodes = C.cartesian_nodes(np.linspace(0, L, n + 1), np.linspace(0, L, n + 1), np.linspace(0, L, n + 1), "cpu")
    per = (C.PERIODIC_WRAP,) * 3
    lay, mesh = C.mesh_metrics(nodes, per)
    u = torch.zeros(lay.gshape + (3,), dtype=torch.float64)
    xc = C.interior(mesh["cent"], lay)[..., 0]
    C.interior(u, lay)[..., 2] = torch.sin(xc)
    bcs = [C.BC_PERIODIC] * 6
    C.apply_vector_bcs(u, lay, per, bcs, [None] * 6, mesh)
    C.reset_cell_periodic(u, lay, per)
    V = C.ucont_from_ucat(u, mesh, lay, per, bcs)
    params = C.default_params(per, bcs, bcs, nu=0.1, les_active=0)
    rhs = np.zeros((lay.mz, lay.my, lay.mx, 3))
    O.formu(lay.info(), np_mesh(mesh), params, to_np(V), to_np(u), np.zeros(lay.gshape), rhs, 1.0, 2)
    lap = -(2 - 2 * np.cos(h)) / h ** 2 * np.sin(to_np(xc))       # at cell centres, w component
    # j faces carry the z component: scale * jaj * 1/2 (eta.fp_c + eta.fp_c+j), eta = h^2 z^, fp = vol * nu lap
    expect = 0.1 * h ** 2 * lap                                     # uniform in j: the j average is exact
    got = rhs[1:-1, 1:-1, 1:-1, 1]
    assert np.abs(got - expect).max() <= 1e-12 * np.abs(expect).max()
    assert np.abs(rhs[1:-1, 1:-1, 1:-1, 0]).max() <= 1e-14 and np.abs(rhs[1:-1, 1:-1, 1:-1, 2]).max() <= 1e-14


# ------------------------------------------------------------------------------------------------
# dynamic Smagorinsky closed form
# ------------------------------------------------------------------------------------------------
BIN = {3: np.array([1.0, 2.0, 1.0]) / 4.0, 5: np.array([1.0, 4.0, 6.0, 4.0, 1.0]) / 16.0}     # binomial rows
SIMP = {3: np.array([1.0, 4.0, 1.0]), 5: np.array([1.0, 4.0, 2.0, 4.0, 1.0])}              # composite Simpson


def cart_gradient(u, periodic, gxi):
    """grad u [m][n] at cell centres by the chain rule, d/dxi central in the interior and one-sided
    first order at the ends of a non-periodic direction (numpy.gradient, edge_order=1)."""
    d = []
    for q in range(3):
        ax = AX[q]
        if periodic[q]:
            d.append(0.5 * (sh(u, q, 1) - sh(u, q, -1)))
        else:
            d.append(np.gradient(u, axis=ax, edge_order=1))
    return sum(d[q][..., :, None] * gxi[q][None, None, None, None, :] for q in range(3))


def masked_filter(fields, w1, vol):
    """Separable-weight average over the interior cells only (cells outside the interior carry
    weight 0), normalised by the weights present; also returns sum(8^H * binomial * vol)."""
    H = len(w1) // 2
    shape = fields[0].shape[:3]
    pad = [(H, H)] * 3
    m = np.pad(np.ones(shape), pad)
    pf = [np.pad(f, pad + [(0, 0)] * (f.ndim - 3)) for f in fields]
    num = [np.zeros_like(f) for f in fields]
    den = np.zeros(shape)
    sv = np.zeros(shape)
    b = BIN[len(w1)]
    K, J, I = shape
    for ok, oj, oi in itertools.product(range(2 * H + 1), repeat=3):
        w = w1[ok] * w1[oj] * w1[oi]
        sl = (slice(ok, ok + K), slice(oj, oj + J), slice(oi, oi + I))
        mm = m[sl]
        den += w * mm
        sv += (2.0 ** (3 * H)) * b[ok] * b[oj] * b[oi] * mm * vol
        for t, f in zip(num, pf):
            t += w * (mm.reshape(mm.shape + (1,) * (f.ndim - 3))) * f[sl]
    return [t / den.reshape(shape + (1,) * (t.ndim - 3)) for t in num], sv


def germano_cartesian(u, Sij, Smag, vol, NP):
    """LM = L:M and MM = M:M in Cartesian form for one test filter of NP points."""
    (uf, uuf, Sf, SSf), sv = masked_filter(
        [u, u[..., :, None] * u[..., None, :], Sij, Smag[..., None, None] * Sij], SIMP[NP], vol)
    L = uuf - uf[..., :, None] * uf[..., None, :]
    Shat = np.sqrt(2 * (Sf ** 2).sum(axis=(-1, -2)))
    dt2 = sv ** (2.0 / 3.0)
    d2 = vol ** (2.0 / 3.0)
    M = -dt2[..., None, None] * Shat[..., None, None] * Sf + d2 * SSf
    return (L * M).sum(axis=(-1, -2)), (M * M).sum(axis=(-1, -2))


def box_average(f, w_own, periodic):
    """Simpson 3x3x3 average of an interior field f with weights w_own over the owned range
    (ghost ring included); periodic directions take the opposite interior cell's value."""
    K, J, I = f.shape
    g = np.zeros((K + 2, J + 2, I + 2))
    g[1:-1, 1:-1, 1:-1] = f
    for q in range(3):
        if periodic[q]:
            ax = AX[q]
            idx_lo, idx_hi, src_lo, src_hi = [slice(None)] * 3, [slice(None)] * 3, [slice(None)] * 3, [slice(None)] * 3
            idx_lo[ax], src_lo[ax] = 0, -2
            idx_hi[ax], src_hi[ax] = -1, 1
            g[tuple(idx_lo)] = g[tuple(src_lo)]
            g[tuple(idx_hi)] = g[tuple(src_hi)]
    num = np.zeros(f.shape)
    den = np.zeros(f.shape)
    for ok, oj, oi in itertools.product(range(3), repeat=3):
        s = SIMP[3][ok] * SIMP[3][oj] * SIMP[3][oi]
        sl = (slice(ok, ok + K), slice(oj, oj + J), slice(oi, oi + I))
        num += s * w_own[sl] * g[sl]
        den += s * w_own[sl]
    return num / den


def box_weights(aj_own, periodic):
    """1/aj over the owned range, zero on the ghost ring of non-periodic directions, except the
    upper j ghost, which keeps its weight (src/les.c:1911 tests the centre index j, not J)."""
    w = 1.0 / aj_own
    if not periodic[0]:
        w[:, :, 0] = 0
        w[:, :, -1] = 0
    if not periodic[2]:
        w[0] = 0
        w[-1] = 0
    if not periodic[1]:
        w[:, 0, :] = 0
    return w


def dsm_closed_form(u, vecs, periodic, aj_own, model, max_cs):
    S, vol, gxi = affine_geometry(vecs)
    G = cart_gradient(u, periodic, gxi)
    Sij = 0.5 * (G + np.swapaxes(G, -1, -2))
    Smag = np.sqrt(2 * (Sij ** 2).sum(axis=(-1, -2)))
    LM, MM = germano_cartesian(u, Sij, Smag, vol, 3)
    if model == C.LES_DSM:
        w = box_weights(aj_own, periodic)
        LMa, MMa = box_average(LM, w, periodic), box_average(MM, w, periodic)
        Cs = 0.5 * LMa / (MMa + 1e-10)
    else:
        QN, NN = germano_cartesian(u, Sij, Smag, vol, 5)
        pl = lambda f: f.mean(axis=(0, 2), keepdims=True)        # j planes (array axis 1)
        lm, mm, qn, nn = pl(LM), pl(MM), pl(QN), pl(NN)
        beta = qn * mm / (nn * lm + 1e-10)
        Cs = np.broadcast_to(0.5 * (lm / (mm + 1e-10)) / beta, LM.shape)
    Cs = np.minimum(np.maximum(Cs, 0.0), max_cs)
    nut = Cs * vol ** (2.0 / 3.0) * Smag
    return Cs, nut, Smag


def owned_np(t, lay):
    return to_np(C.owned(t, lay))


PER_DSM = [(2, 2, 2), (1, 1, 1), (2, 0, 1), (0, 0, 0), (1, 2, 0)]


@pytest.mark.parametrize("geom", sorted(AFFINE))
@pytest.mark.parametrize("periodic", PER_DSM)
@pytest.mark.parametrize("model", [C.LES_DSM, C.LES_DPASD])
def test_dynamic_smagorinsky_closed_form(geom, periodic, model):
    vecs = AFFINE[geom]
    lay, mesh, u, V, nut, params = build((8, 7, 9), vecs, periodic, seed=5, les_model=model)
    params = dict(params, max_cs=1e6)          # unclipped above: the raw dynamic coefficient
    nm = np_mesh(mesh)
    cs = np.full(lay.gshape, 7.0)
    O.update_cs(lay.info(), nm, params, to_np(u), cs)
    nt, ks = np.full(lay.gshape, 3.0), np.full(lay.gshape, 5.0)
    O.update_nut(lay.info(), nm, params, to_np(u), cs, nt, ks)
    Cs_e, nut_e, _ = dsm_closed_form(interior_np(u, lay), vecs, periodic, owned_np(mesh["aj"], lay), model, 1e6)
    ic = lambda a: a[4:-4, 4:-4, 4:-4]
    assert relerr(ic(cs), Cs_e) <= 1e-12, relerr(ic(cs), Cs_e)
    assert relerr(ic(nt), nut_e) <= 1e-12
    # k_sgs = nu_t^2 / (Ck^2 Delta_harm^2), Delta_harm^2 = 3 / sum 1/Delta_d^2, Delta_d = vol / |S^d|
    S, vol, _ = affine_geometry(vecs)
    dh2 = 3.0 / sum((np.linalg.norm(S[a]) / vol) ** 2 for a in range(3))
    assert relerr(ic(ks), nut_e ** 2 / (0.094 ** 2 * dh2)) <= 1e-12


@pytest.mark.parametrize("max_cs", [0.0, 0.01])
def test_dynamic_smagorinsky_clip(max_cs):
    vecs = AFFINE["box"]
    per = (2, 0, 2)
    lay, mesh, u, V, nut, params = build((8, 7, 9), vecs, per, seed=6, les_model=C.LES_DSM)
    params = dict(params, max_cs=max_cs)
    cs = np.zeros(lay.gshape)
    O.update_cs(lay.info(), np_mesh(mesh), params, to_np(u), cs)
    Cs_e, _, _ = dsm_closed_form(interior_np(u, lay), vecs, per, owned_np(mesh["aj"], lay), C.LES_DSM, max_cs)
    assert np.abs(cs[4:-4, 4:-4, 4:-4] - Cs_e).max() <= 1e-14
    assert cs[4:-4, 4:-4, 4:-4].min() >= 0.0 and cs[4:-4, 4:-4, 4:-4].max() <= max_cs


def test_smagorinsky_closed_form():
    """Constant Smagorinsky: Cs = 0.0289 everywhere, nu_t = Cs Delta^2 |S| (src/les.c:9-10)."""
    vecs = AFFINE["sheared"]
    per = (2, 0, 1)
    lay, mesh, u, V, nut, params = build((6, 7, 5), vecs, per, seed=7, les_model=C.LES_SMAGORINSKY)
    cs = np.zeros(lay.gshape)
    O.update_cs(lay.info(), np_mesh(mesh), params, to_np(u), cs)
    assert np.all(cs == 0.0289)
    nt = np.zeros(lay.gshape)
    O.update_nut(lay.info(), np_mesh(mesh), params, to_np(u), cs, nt, None)
    S, vol, gxi = affine_geometry(vecs)
    G = cart_gradient(interior_np(u, lay), per, gxi)
    Sij = 0.5 * (G + np.swapaxes(G, -1, -2))
    expect = 0.0289 * vol ** (2 / 3) * np.sqrt(2 * (Sij ** 2).sum(axis=(-1, -2)))
    assert relerr(nt[4:-4, 4:-4, 4:-4], expect) <= 1e-13


# ------------------------------------------------------------------------------------------------
# invariants on curvilinear meshes
# ------------------------------------------------------------------------------------------------
PER_CURV = [(0, 0, 0), (2, 0, 1), (1, 1, 1), (2, 2, 2)]


def curvilinear(periodic, seed, u_const=None, les_model=C.LES_DSM, scheme=C.DIV_QUICK):
    nodes = random_nodes((7, 6, 8), seed, amp=0.15)
    lay, mesh = C.mesh_metrics(nodes, periodic)
    rng = np.random.Generator(np.random.PCG64(100 + seed))
    u = torch.zeros(lay.gshape + (3,), dtype=torch.float64)
    if u_const is None:
        C.interior(u, lay)[...] = torch.as_tensor(rng.standard_normal(C.interior(u, lay).shape))
    else:
        u[...] = torch.as_tensor(u_const, dtype=torch.float64)
    bcs = [C.BC_PERIODIC if periodic[a] else C.BC_ZEROGRAD for a in range(3) for _ in range(2)]
    if u_const is None:
        C.apply_vector_bcs(u, lay, periodic, bcs, [None] * 6, mesh)
        C.reset_cell_periodic(u, lay, periodic)
    V = C.ucont_from_ucat(u, mesh, lay, periodic, bcs)
    params = C.default_params(periodic, bcs, bcs, div_scheme=scheme, les_model=les_model, nu=0.02)
    return lay, mesh, u, V, params


@pytest.mark.parametrize("periodic", PER_CURV)
@pytest.mark.parametrize("scheme", [C.DIV_CENTRAL, C.DIV_QUICK])
def test_free_stream_preservation(periodic, scheme):
    """Uniform velocity on a distorted mesh: the face vectors of every cell close, so the
    convective divergence vanishes to round-off and the viscous term is exactly zero."""
    c = np.array([1.3, -0.7, 0.4])
    lay, mesh, u, V, params = curvilinear(periodic, 21, u_const=c, scheme=scheme)
    nm = np_mesh(mesh)
    nut = np.full(lay.gshape, 0.01)
    scale = np.abs(to_np(V)).max() * np.abs(c).max() * np.abs(nm["iaj"]).max() * np.abs(nm["csi"]).max()
    for mode in (1, 2):
        rhs = np.zeros((lay.mz, lay.my, lay.mx, 3))
        O.formu(lay.info(), nm, params, to_np(V), to_np(u), nut, rhs, 1.0, mode)
        if mode == 1:
            assert np.abs(rhs).max() <= 1e-13 * scale
        else:
            assert np.abs(rhs).max() == 0.0


@pytest.mark.parametrize("periodic", PER_CURV)
def test_formmode_split_and_scale(periodic):
    """formMode 1 (convection) + formMode 2 (diffusion) = formMode 0, and the contribution is
    linear in scale."""
    lay, mesh, u, V, params = curvilinear(periodic, 22)
    nm = np_mesh(mesh)
    rng = np.random.Generator(np.random.PCG64(5))
    nut = rng.uniform(0, 0.05, lay.gshape)
    rhs0 = rng.standard_normal((lay.mz, lay.my, lay.mx, 3))
    out = {}
    for mode, scale in ((0, 1.0), (1, 1.0), (2, 1.0), (0, -2.5)):
        r = rhs0.copy()
        O.formu(lay.info(), nm, params, to_np(V), to_np(u), nut, r, scale, mode)
        out[(mode, scale)] = r - rhs0
    assert relerr(out[(1, 1.0)] + out[(2, 1.0)], out[(0, 1.0)]) <= 1e-13
    assert relerr(out[(0, -2.5)], -2.5 * out[(0, 1.0)]) <= 1e-14
    assert np.abs(out[(1, 1.0)]).max() > 0 and np.abs(out[(2, 1.0)]).max() > 0


@pytest.mark.parametrize("periodic", PER_CURV)
@pytest.mark.parametrize("model", [C.LES_DSM, C.LES_DPASD, C.LES_SMAGORINSKY])
def test_uniform_velocity_no_eddy_viscosity(periodic, model):
    """S = 0 for a uniform velocity, hence L = M = 0 and nu_t = 0.  The dynamic Smagorinsky
    coefficient is 0.5 * 0 / (0 + 1e-10) = 0.  The plane-averaged model divides 0 by beta = 0
    (src/les.c:2042-2043); the clip PetscMin(PetscMax(NaN, 0), maxCs) then yields maxCs, because
    both macros are plain comparisons, false for NaN."""
    lay, mesh, u, V, params = curvilinear(periodic, 23, u_const=np.array([2.0, -1.0, 0.5]), les_model=model)
    nm = np_mesh(mesh)
    cs = np.full(lay.gshape, 9.0)
    O.update_cs(lay.info(), nm, params, to_np(u), cs)
    if model == C.LES_DSM:
        assert np.all(cs[4:-4, 4:-4, 4:-4] == 0.0)
    elif model == C.LES_DPASD:
        assert np.all(cs[4:-4, 4:-4, 4:-4] == params["max_cs"])
    nt = np.full(lay.gshape, 9.0)
    O.update_nut(lay.info(), nm, params, to_np(u), cs, nt, None)
    assert np.all(nt[4:-4, 4:-4, 4:-4] == 0.0)


@pytest.mark.parametrize("model", [C.LES_DSM, C.LES_DPASD])
def test_dynamic_coefficient_galilean_invariance(model):
    """On a uniform lattice the Leonard stress L = <u u> - <u><u> and M (strain only) do not change
    when a constant velocity is added; neither does the dynamic coefficient."""
    vecs = AFFINE["sheared"]
    per = (2, 0, 1)
    lay, mesh, u, V, nut, params = build((8, 7, 9), vecs, per, seed=8, les_model=model)
    params = dict(params, max_cs=1e6)
    nm = np_mesh(mesh)
    cs1 = np.zeros(lay.gshape)
    O.update_cs(lay.info(), nm, params, to_np(u), cs1)
    u2 = to_np(u) + np.array([3.0, -2.0, 1.5])
    cs2 = np.zeros(lay.gshape)
    O.update_cs(lay.info(), nm, params, u2, cs2)
    assert relerr(cs2[4:-4, 4:-4, 4:-4], cs1[4:-4, 4:-4, 4:-4]) <= 1e-10
    assert np.abs(cs1[4:-4, 4:-4, 4:-4]).max() > 0


@pytest.mark.parametrize("periodic", PER_CURV)
def test_buoyancy_closed_form(periodic):
    """gcont^q = g . S^q_face with g = (0, 0, -9.81); bTheta^q = scale gcont^q (t_ref - theta_face) /
    t_ref on the faces of interior cells, zero elsewhere, and zero everywhere at theta = t_ref."""
    lay, mesh, u, V, params = curvilinear(periodic, 24)
    nm = np_mesh(mesh)
    rng = np.random.Generator(np.random.PCG64(3))
    theta = 300.0 + rng.standard_normal(lay.gshape)
    bt = rng.standard_normal((lay.mz, lay.my, lay.mx, 3))
    gc = np.zeros((lay.mz, lay.my, lay.mx, 3))
    O.buoyancy(lay.info(), nm, params, theta, 0.8, bt, gc)
    G3 = 3
    for a, name in enumerate(("icsi", "jeta", "kzet")):
        face = nm[name][G3:G3 + lay.mz, G3:G3 + lay.my, G3:G3 + lay.mx]
        g = -9.81 * face[..., 2]
        ic = (slice(1, -1),) * 3
        tc = theta[G3 + 1:G3 + lay.mz - 1, G3 + 1:G3 + lay.my - 1, G3 + 1:G3 + lay.mx - 1]
        sl = [slice(G3 + 1, G3 + lay.mz - 1), slice(G3 + 1, G3 + lay.my - 1), slice(G3 + 1, G3 + lay.mx - 1)]
        sl[AX[a]] = slice(sl[AX[a]].start + 1, sl[AX[a]].stop + 1)
        tn = theta[tuple(sl)]
        expect = 0.8 * g[ic] * (300.0 - 0.5 * (tc + tn)) / 300.0
        assert np.abs(bt[ic + (a,)] - expect).max() <= 1e-13 * np.abs(expect).max()
        assert np.abs(gc[ic + (a,)] - g[ic]).max() <= 1e-14 * np.abs(g).max()
    outside = np.ones(bt.shape[:3], bool)
    outside[1:-1, 1:-1, 1:-1] = False
    assert np.all(bt[outside] == 0.0)
    bt2 = np.ones_like(bt)
    O.buoyancy(lay.info(), nm, params, np.full(lay.gshape, 300.0), 1.0, bt2, None)
    assert np.all(bt2 == 0.0)
