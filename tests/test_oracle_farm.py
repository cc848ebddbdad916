"""Pins of the wind-farm body force: the uniform actuator-disk geometry and force
(tosca_amd/farm.py: meshUADM, computeBladeForce), the oracle's Gaussian projection
(projectBladeForce, src/turbines/turbines_force.c:717-932) and its Cartesian-to-face conversion
(bodyForceCartesian2Contravariant, :1741-1835), each against facts outside the code:
  - the disk areas sum to pi rTip^2 and the force to 0.5 Ct Uref^2 pi rTip^2 along the rotor axis;
  - the rotation routine equals scipy's rotation matrices;
  - the projected force integrates (sum of scat x cell volume) to the truncated Gaussian's mass
    erf(n) - 2 n exp(-n^2) / sqrt(pi), n = prjNSigma, times the total force (a dropped square, a
    wrong power of eps or of pi, or a wrong cut-off fails it);
  - a rotor at the centre of a symmetric mesh gives a field mirror-symmetric about the rotor
    centre (an index or sign slip in the cell loop fails it);
  - the face conversion of a constant force F is F . (icsi, jeta, kzet) exactly on the interior
    cells and zero elsewhere, and a linear force is reproduced at the face centres."""
import math

import numpy as np
import pytest
import torch
from scipy.spatial.transform import Rotation
from scipy.special import erf

from oracle import oracle as O
from tosca_amd import case as C
from tosca_amd import farm as F
from tests._util import np_mesh


def test_rotate_matches_scipy():
    rng = np.random.Generator(np.random.PCG64(3))
    for _ in range(20):
        axis = rng.standard_normal(3)
        v = rng.standard_normal(3)
        th = rng.uniform(-4, 4)
        ref = Rotation.from_rotvec(th * axis / np.linalg.norm(axis)).apply(v)
        assert np.allclose(F.rotate(axis, v, th), ref, rtol=0, atol=1e-14 * (1 + np.abs(v).max()))


@pytest.mark.parametrize("tilt,precone", [(0.0, 0.0), (5.0, 2.5)])
def test_disk_geometry_and_force(tilt, precone):
    t = F.UadmTurbine(up_tilt=tilt, precone=precone, n_rad=7, n_azi=24, ct=0.75, uref=8.0)
    pts, da, axis = F.disk_points(t, (100.0, 200.0, 0.0))
    center, xr, yr, zr = F.rotor_frame(t, (100.0, 200.0, 0.0))
    assert pts.shape == (7 * 24, 3)
    assert math.isclose(da.sum(), math.pi * t.r_tip ** 2, rel_tol=1e-13)
    assert np.allclose(np.linalg.norm(axis), 1.0, atol=1e-12)
    # the centre: base + hTower z + overHang rotorDir
    assert np.allclose(center, [100.0 - t.over_hang, 200.0, t.h_tower], atol=1e-9)
    # every point sits on the (preconed) cone around the axis: radius from the axis in [rHub, rTip]
    rel = pts - center
    along = rel @ axis
    radial = np.linalg.norm(rel - np.outer(along, axis), axis=1)
    assert radial.min() >= t.r_hub * math.cos(math.radians(precone)) - 1e-9
    assert radial.max() <= t.r_tip + 1e-9
    if precone == 0.0:
        assert np.allclose(along, 0.0, atol=1e-9)
        assert np.allclose(sorted(set(np.round(radial, 9))), t.r_hub + np.arange(7) * (t.r_tip - t.r_hub) / 6, atol=1e-8)
    force = F.disk_force(t, da, axis)
    total = force.sum(axis=0)
    expect = 0.5 * t.ct * t.uref ** 2 * math.pi * t.r_tip ** 2
    assert np.allclose(total, expect * axis, rtol=1e-13, atol=1e-9)


def _box_case(n, h, periodic=(0, 0, 0)):
    ax = [np.arange(n + 1) * h for _ in range(3)]
    nodes = C.cartesian_nodes(ax[0], ax[1], ax[2], "cpu")
    lay, mesh = C.mesh_metrics(nodes, periodic)
    return lay, mesh


def _project(lay, mesh, farm, periodic=(0, 0, 0)):
    params = C.default_params(periodic, (C.BC_SLIP,) * 6, (C.BC_ZEROGRAD,) * 6)
    cent = mesh["cent"].numpy().copy()
    scat = np.full(lay.gshape + (3,), np.nan)
    O.project_uadm(lay.info(), params, cent, farm, scat)
    return scat, params


def test_projection_integrates_to_truncated_gaussian_mass():
    # one rotor (small disk, 5 x 8 points) in a uniform box with spacing eps / 4
    h, n = 0.25, 44
    t = F.UadmTurbine(r_tip=3.0, r_hub=0.5, h_tower=5.5, over_hang=0.0, epsilon=1.0, n_rad=5, n_azi=8,
                      ct=0.7, uref=2.0)
    farm = F.make_farm(t, [(5.5, 5.5, 0.0)])
    lay, mesh = _box_case(n, h)
    scat, _ = _project(lay, mesh, farm)
    vol = 1.0 / C.interior(mesh["aj"], lay).numpy()
    s = C.interior(torch.as_tensor(scat), lay).numpy()
    total = (s * vol[..., None]).reshape(-1, 3).sum(axis=0)
    ns = F.PRJ_NSIGMA
    mass = erf(ns) - 2.0 * ns * math.exp(-ns * ns) / math.sqrt(math.pi)
    expect = mass * farm.total_force()
    assert abs(total[0] - expect[0]) <= 2e-3 * abs(expect[0])
    assert np.abs(total[1:]).max() <= 1e-12 * abs(expect[0])
    # outside the interior cells the field is zero (no type-2 axes, so no ghost wrap)
    full = scat.copy()
    C.interior(torch.as_tensor(full), lay)[...] = 0.0
    assert np.all(full == 0.0)


def test_projection_mirror_symmetric_about_rotor_centre():
    h, n = 0.5, 24
    t = F.UadmTurbine(r_tip=3.0, r_hub=0.5, h_tower=6.0, over_hang=0.0, epsilon=1.0, n_rad=4, n_azi=10,
                      ct=0.7, uref=2.0)
    farm = F.make_farm(t, [(6.0, 6.0, 0.0)])      # rotor centre (6, 6, 6): the box centre
    lay, mesh = _box_case(n, h)
    scat, _ = _project(lay, mesh, farm)
    s = C.interior(torch.as_tensor(scat), lay).numpy()[..., 0]     # [k (x)][j (z)][i (y)]
    assert np.abs(s).max() > 0
    scale = np.abs(s).max()
    # nUnit divides by |v| + 1e-12 (inline.h:870-889): the hub sits 6e-12 below z = 6, which moves
    # the field by ~|grad| 6e-12 under the z mirror; the y mirror is exact up to rounding
    assert farm.center[0, 2] == 6.0 / (1.0 + 1e-12)
    assert np.allclose(s, s[:, :, ::-1], rtol=0, atol=1e-13 * scale)    # y -> -y
    assert np.allclose(s, s[:, ::-1, :], rtol=0, atol=5e-11 * scale)    # z -> -z
    # the rotor is thin in x: the field decays along x like the Gaussian of the distance to the disk
    prof = s.sum(axis=(1, 2))
    assert np.allclose(prof, prof[::-1], rtol=1e-12, atol=1e-13 * prof.max())


@pytest.mark.parametrize("shear", [0.0, 0.3])
def test_face_conversion_constant_and_linear(shear):
    n = (6, 5, 7)
    x = np.arange(n[0] + 1, dtype=np.float64) * 1.5
    y = np.arange(n[1] + 1, dtype=np.float64) * 0.8
    z = np.arange(n[2] + 1, dtype=np.float64) * 1.1
    nodes = C.cartesian_nodes(x, y, z, "cpu").clone()
    nodes[..., 0] += shear * nodes[..., 2]          # an affine (sheared) lattice
    lay, mesh = C.mesh_metrics(nodes, (0, 0, 0))
    params = C.default_params((0, 0, 0), (C.BC_SLIP,) * 6, (C.BC_ZEROGRAD,) * 6)
    m = np_mesh(mesh)
    Fc = np.array([1.25, -0.5, 2.0])
    scat = np.broadcast_to(Fc, lay.gshape + (3,)).copy()
    scont = np.full((lay.mz, lay.my, lay.mx, 3), np.nan)
    O.body_force_to_cont(lay.info(), m, params, scat, scont)
    I = lambda a: a[1:-1, 1:-1, 1:-1]
    G = lambda a: a[3:-3, 3:-3, 3:-3]
    exp = np.stack([G(m["icsi"]) @ Fc, G(m["jeta"]) @ Fc, G(m["kzet"]) @ Fc], axis=-1)
    assert np.allclose(I(scont), I(exp), rtol=1e-15, atol=1e-15 * np.abs(exp).max())
    rest = scont.copy()
    rest[1:-1, 1:-1, 1:-1] = 0.0
    assert np.all(rest == 0.0)
    # linear force a + B x at the cell centres: each face gets a + B x at the face centre (the
    # mean of the two centres on an affine lattice)
    B = np.array([[0.3, -0.2, 0.1], [0.05, 0.4, -0.3], [-0.1, 0.2, 0.25]])
    cent = mesh["cent"].numpy()
    scat = cent @ B.T + Fc
    O.body_force_to_cont(lay.info(), m, params, scat, scont)
    Lc = lambda k, j, i: cent[3 + k, 3 + j, 3 + i]
    for (k, j, i) in [(1, 1, 1), (3, 2, 4), (lay.mz - 2, lay.my - 2, lay.mx - 2)]:
        fi = (0.5 * (Lc(k, j, i) + Lc(k, j, i + 1))) @ B.T + Fc
        fj = (0.5 * (Lc(k, j, i) + Lc(k, j + 1, i))) @ B.T + Fc
        fk = (0.5 * (Lc(k, j, i) + Lc(k + 1, j, i))) @ B.T + Fc
        got = scont[k, j, i]
        assert math.isclose(got[0], fi @ m["icsi"][3 + k, 3 + j, 3 + i], rel_tol=1e-13, abs_tol=1e-13)
        assert math.isclose(got[1], fj @ m["jeta"][3 + k, 3 + j, 3 + i], rel_tol=1e-13, abs_tol=1e-13)
        assert math.isclose(got[2], fk @ m["kzet"][3 + k, 3 + j, 3 + i], rel_tol=1e-13, abs_tol=1e-13)


def test_shard_farm_layout():
    farm = F.shard_farm()
    assert farm.n_rotors == 16 and farm.n_points == 500
    c = farm.center
    # 7 D spacing in x and y, hub at 90 m, centred in the 5120 m square
    xs, ys = sorted(set(np.round(c[:, 0], 6))), sorted(set(np.round(c[:, 1], 6)))
    assert np.allclose(np.diff(xs), 7 * 126.0) and np.allclose(np.diff(ys), 7 * 126.0)
    assert np.allclose(c[:, 2], 90.0)
    assert math.isclose(np.mean(ys), 2560.0, rel_tol=1e-12)
    total = farm.total_force()
    expect = 16 * 0.5 * 0.75 * 8.0 ** 2 * math.pi * 63.0 ** 2
    # the rotor axis is rotorDir / (|rotorDir| + 1e-12) (nUnit), so the force is 1e-12 short
    assert math.isclose(-total[0], expect / (1.0 + 1e-12), rel_tol=1e-14)
    assert np.allclose(farm.eps, 20.0) and np.allclose(farm.sphere, 63.0 + 20.0 * F.PRJ_NSIGMA)
