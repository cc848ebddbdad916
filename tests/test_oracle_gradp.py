"""Pins of the oracle's GradP restatement (oracle_grad_p, src/peqn.c:3437-3773) from facts outside
its code:

  - on an affine lattice a pressure linear in the cell indices, p = a + c . (i, j, k), is linear in
    space with gradient b = sum_q c_q S^q / V (S^q the face area vectors, V the cell volume), and
    every difference the routine takes, one-sided corner rows included, is exact for it: the flux
    through face q is S^q . b at the faces it computes, zero on the right boundary face of a
    non-periodic axis and everywhere else in the vector;
  - a pressure periodic along one axis of a box lattice gives |S^q|^2 / V times the difference to
    the periodic neighbour: cell 1 beyond face m-2, read as the copied boundary cell m-1 (type 1)
    or as the wrapped ghost m+1 (type 2).
  - on a stretched orthogonal lattice (face Jacobians differing per direction) the flux of a
    pressure linear in space is still S^q . grad p between interior cells.
"""
import numpy as np
import pytest
import torch

from tosca_amd import case as C
from oracle import oracle as O
from tests._util import np_mesh
from tests.test_oracle_pins import AFFINE, affine_nodes, affine_geometry

G = 3


def _idx(lay):
    return np.meshgrid(np.arange(lay.mz + 2 * G) - G, np.arange(lay.my + 2 * G) - G, np.arange(lay.mx + 2 * G) - G,
                       indexing="ij")


def _params(periodic):
    bcs = [C.BC_PERIODIC if periodic[a // 2] else C.BC_FIXED for a in range(6)]
    return C.default_params(periodic, bcs, bcs, les_active=0)


@pytest.mark.parametrize("geom", sorted(AFFINE))
@pytest.mark.parametrize("periodic", [(0, 0, 0), (1, 0, 2), (0, 2, 0), (2, 1, 1)])
def test_linear_pressure_exact(geom, periodic):
    vecs = AFFINE[geom]
    lay, mesh_t = C.mesh_metrics(affine_nodes((6, 7, 5), vecs), periodic)
    mesh = np_mesh(mesh_t)
    S, vol, _ = affine_geometry(vecs)
    vol = abs(vol)
    c = np.array([0.7, -1.2, 0.45]) * np.array([0.0 if periodic[a] else 1.0 for a in range(3)])
    k, j, i = _idx(lay)
    p = 3.0 + c[0] * i + c[1] * j + c[2] * k
    dp = np.full((lay.mz, lay.my, lay.mx, 3), np.nan)
    O.grad_p(lay.info(), mesh, _params(periodic), p, dp)
    b = (c[0] * S[0] + c[1] * S[1] + c[2] * S[2]) / vol
    m = (lay.mx, lay.my, lay.mz)
    K, J, I = np.meshgrid(np.arange(lay.mz), np.arange(lay.my), np.arange(lay.mx), indexing="ij")
    inner = (I >= 1) & (I <= m[0] - 2) & (J >= 1) & (J <= m[1] - 2) & (K >= 1) & (K <= m[2] - 2)
    face_idx = (I, J, K)
    want = np.zeros(dp.shape)
    for q in range(3):
        on = inner & ((face_idx[q] <= m[q] - 3) | (periodic[q] != 0))
        want[..., q] = np.where(on, S[q] @ b, 0.0)
    assert np.abs(dp - want).max() <= 1e-13 * np.abs(want).max()
    assert (dp[~inner] == 0).all()


@pytest.mark.parametrize("ptype", [1, 2])
@pytest.mark.parametrize("axis", [0, 1, 2])
def test_periodic_neighbour(ptype, axis):
    vecs = AFFINE["box"]
    periodic = (ptype,) * 3
    lay, mesh_t = C.mesh_metrics(affine_nodes((6, 7, 5), vecs), periodic)
    mesh = np_mesh(mesh_t)
    S, vol, _ = affine_geometry(vecs)
    m = (lay.mx, lay.my, lay.mz)[axis]
    n = m - 2
    k, j, i = _idx(lay)
    idx = (i, j, k)[axis]
    f = lambda x: np.sin(2 * np.pi * (x - 1) / n) + 0.3 * np.cos(4 * np.pi * (x - 1) / n)
    p = f(idx.astype(float))
    if ptype == 1:      # boundary cells hold copies: cell 0 = cell m-2, cell m-1 = cell 1
        p = np.where(idx == 0, f(m - 2.0), np.where(idx == m - 1, f(1.0), p))
    else:               # type 2: the dummy cells 0 and m-1 hold anything finite; ghosts wrap
        p = np.where((idx == 0) | (idx == m - 1), 0.0, p)
        p = C.local_to_local(torch.as_tensor(p), lay, periodic).numpy()
        sl = [slice(None)] * 3
        sl[2 - axis] = m + 1 + G
        s1 = [slice(None)] * 3
        s1[2 - axis] = 1 + G
        assert np.array_equal(p[tuple(sl)], p[tuple(s1)])      # the wrap puts cell 1 at m+1
    dp = np.full((lay.mz, lay.my, lay.mx, 3), np.nan)
    O.grad_p(lay.info(), mesh, _params(periodic), p, dp)
    cells = np.arange(1, m - 1)
    nxt = np.where(cells == m - 2, 1, cells + 1)
    want = S[axis] @ S[axis] / abs(vol) * (f(nxt.astype(float)) - f(cells.astype(float)))
    inner = dp[1:-1, 1:-1, 1:-1, axis]
    got = np.moveaxis(inner, 2 - axis, -1)
    assert np.abs(got - want).max() <= 1e-13 * np.abs(want).max()
    others = [q for q in range(3) if q != axis]
    assert np.abs(dp[1:-1, 1:-1, 1:-1][..., others]).max() <= 1e-13 * np.abs(want).max()


def test_linear_pressure_stretched_box():
    """geometrically stretched, orthogonal lattice: face Jacobians differ per direction and the
    flux of a pressure linear in space is still S^q . grad p at the faces between interior cells"""
    x = np.cumsum(np.r_[0, 1.0 * 1.15 ** np.arange(6)])
    y = np.cumsum(np.r_[0, 0.8 * 1.1 ** np.arange(7)])
    z = np.cumsum(np.r_[0, 0.5 * 1.2 ** np.arange(5)])
    lay, mesh_t = C.mesh_metrics(C.cartesian_nodes(x, y, z, "cpu"), (0, 0, 0))
    mesh = np_mesh(mesh_t)
    b = np.array([0.3, -0.7, 1.1])
    p = 2.0 + mesh_t["cent"].numpy() @ b
    dp = np.full((lay.mz, lay.my, lay.mx, 3), np.nan)
    O.grad_p(lay.info(), mesh, _params((0, 0, 0)), p, dp)
    own = lambda a: a[G:G + lay.mz, G:G + lay.my, G:G + lay.mx]
    for q, name in enumerate(("icsi", "jeta", "kzet")):
        want = own(mesh[name]) @ b
        assert np.abs(dp[2:-2, 2:-2, 2:-2, q] - want[2:-2, 2:-2, 2:-2]).max() <= 1e-13 * np.abs(want).max()
