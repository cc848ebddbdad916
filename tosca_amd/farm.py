"""Uniform actuator-disk rotors (the reference's uniformADM turbine model) as host-side input.

The reference sets a rotor up once (readTurbineProperties, src/turbines/turbines_io.c:3440-3535;
initUADM and meshUADM, :1820-2001), computes the force at each disk point every time step
(computeBladeForce, src/turbines/turbines_force.c:219-365) and then projects the points onto the
mesh (projectBladeForce, :717-932) and converts the projected Cartesian body force to face fluxes
(bodyForceCartesian2Contravariant, :1741-1835).  The per-rotor geometry and forces are a few
hundred points: they are prepared here, in float64 numpy, in the reference's expression order.
The two mesh-sized steps run on the GPU (tosca_project_uadm and tosca_body_force_to_cont).

Supported: the "constant" CtType with sampleType "givenVelocity" (Uref from the turbine file) --
the force of every point is then 0.5 Uref^2 dA Ct along the rotor axis.  The sampled-velocity
variants need the flow solution at upstream or disk points and are not part of this path.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import List, Sequence, Tuple

import numpy as np

# projection confidence interval in standard deviations (turbines_io.c:3469, hard-coded there)
PRJ_NSIGMA = math.sqrt(math.log(1.0 / 0.001))


@dataclass
class UadmTurbine:
    """The turbine-file parameters the uniformADM model reads (NREL5MW values by default, as in the
    reference's tests/*/turbines/NREL5MW)."""
    r_tip: float = 63.0
    r_hub: float = 1.5
    h_tower: float = 90.0
    over_hang: float = 5.01910
    precone: float = 0.0
    tower_dir: Tuple[float, float, float] = (0.0, 0.0, 1.0)
    rotor_dir: Tuple[float, float, float] = (-1.0, 0.0, 0.0)
    up_tilt: float = 0.0
    epsilon: float = 8.0
    n_rad: int = 10
    n_azi: int = 50
    ct: float = 0.783
    uref: float = 9.0


def _unit(v):
    """nUnit (src/include/inline.h:870-889): v / (|v| + 1e-12)."""
    v = np.asarray(v, dtype=np.float64)
    return v / (math.sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) + 1.e-12)


def _cross(a, b):
    return np.array([a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]])


def rotate(axis, v, theta: float) -> np.ndarray:
    """mRot (src/include/inline.h:1048-1093): v rotated by theta about axis (Rodrigues' matrix,
    summed term by term in the reference's order)."""
    a = np.asarray(axis, dtype=np.float64)
    mag = math.sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2])
    if mag < 1.0e-20:
        raise ValueError("rotation axis too small to working precision")
    ax, ay, az = a[0] / mag, a[1] / mag, a[2] / mag
    c, s = math.cos(theta), math.sin(theta)
    x, y, z = float(v[0]), float(v[1]), float(v[2])
    r = np.zeros(3)
    r[0] += (c + (1 - c) * ax * ax) * x
    r[0] += ((1 - c) * ax * ay - az * s) * y
    r[0] += ((1 - c) * ax * az + ay * s) * z
    r[1] += ((1 - c) * ax * ay + az * s) * x
    r[1] += (c + (1 - c) * ay * ay) * y
    r[1] += ((1 - c) * ay * az - ax * s) * z
    r[2] += ((1 - c) * ax * az - ay * s) * x
    r[2] += ((1 - c) * ay * az + ax * s) * y
    r[2] += (c + (1 - c) * az * az) * z
    return r


def rotor_frame(t: UadmTurbine, base) -> Tuple[np.ndarray, np.ndarray, np.ndarray, np.ndarray]:
    """Rotor centre, rotor axis and the disk's radial unit vectors (readTurbineProperties,
    turbines_io.c:3510-3535): the centre is base + hTower towerDir + overHang rotorDir; the axis
    is the up-tilted rotorDir.  Returns (centre, axis, yr_hat, zr_hat)."""
    deg2rad = math.pi / 180.0
    twr, rtr = _unit(t.tower_dir), _unit(t.rotor_dir)
    tower = t.h_tower * twr + np.asarray(base, dtype=np.float64)
    center = tower + t.over_hang * rtr
    xr = _unit(rtr)
    yr = _unit(twr)
    zr = _cross(xr, yr)
    xr = rotate(zr, xr, t.up_tilt * deg2rad)
    yr = rotate(zr, yr, t.up_tilt * deg2rad)
    return center, xr, yr, zr


def disk_points(t: UadmTurbine, base) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
    """meshUADM (turbines_io.c:1924-2001): nRad x nAzi points (radial station outer, azimuth
    inner) and their areas dA; the crown areas are stretched by (1 + hub area / swept area) so
    that they sum to pi rTip^2.  Returns (points [P,3], dA [P], rotor axis)."""
    deg2rad = math.pi / 180.0
    center, xr, yr, zr = rotor_frame(t, base)
    r_hat = rotate(zr, yr.copy(), t.precone * deg2rad)
    r_hub = t.r_hub * yr
    drval = (t.r_tip - t.r_hub) / (t.n_rad - 1)
    daval = 2 * math.pi / t.n_azi
    swept = math.pi * (t.r_tip * t.r_tip - t.r_hub * t.r_hub)
    hub_area = math.pi * t.r_hub * t.r_hub
    h_coeff = hub_area / swept
    pts, da = [], []
    r = t.r_hub
    for ri in range(t.n_rad):
        dr = drval / 2 if (ri == 0 or ri == t.n_rad - 1) else drval
        crown = math.pi * ((r + dr) * (r + dr) - r * r)
        rvec = (drval * ri) * r_hat + r_hub
        for ai in range(t.n_azi):
            da.append((1.0 + h_coeff) * crown / t.n_azi)
            p = rotate(xr, rvec, daval * ai)
            pts.append(p + center)
        r += dr
    return np.array(pts), np.array(da), xr


def disk_force(t: UadmTurbine, da: np.ndarray, axis: np.ndarray) -> np.ndarray:
    """computeBladeForce, uniformADM branch with a constant Ct and givenVelocity sampling
    (turbines_force.c:224-356): B = (0.5 Uref^2 dA Ct) along the rotor axis, per unit density."""
    b = np.empty((len(da), 3))
    for p, a in enumerate(da):
        mag = 0.5 * t.uref * t.uref * a * t.ct
        b[p] = mag * axis
    return b


@dataclass
class UadmFarm:
    """Every rotor's arrays in the layout tosca_uadm expects (float64, C order)."""
    center: np.ndarray     # [n_rotors, 3]
    sphere: np.ndarray     # [n_rotors]: rTip + eps prjNSigma (turbines_parallel.c:57-80)
    eps: np.ndarray        # [n_rotors]
    points: np.ndarray     # [n_rotors, n_points, 3]
    force: np.ndarray      # [n_rotors, n_points, 3]

    @property
    def n_rotors(self) -> int:
        return int(self.center.shape[0])

    @property
    def n_points(self) -> int:
        return int(self.points.shape[1])

    def total_force(self) -> np.ndarray:
        return self.force.reshape(-1, 3).sum(axis=0)


def make_farm(turbine: UadmTurbine, bases: Sequence[Sequence[float]]) -> UadmFarm:
    cs, sp, ep, pts, fs = [], [], [], [], []
    for base in bases:
        p, da, axis = disk_points(turbine, base)
        center, _, _, _ = rotor_frame(turbine, base)
        cs.append(center)
        sp.append(turbine.r_tip + turbine.epsilon * PRJ_NSIGMA)
        ep.append(turbine.epsilon)
        pts.append(p)
        fs.append(disk_force(turbine, da, axis))
    return UadmFarm(np.array(cs), np.array(sp), np.array(ep), np.ascontiguousarray(np.array(pts)),
                    np.ascontiguousarray(np.array(fs)))


def grid_farm(turbine: UadmTurbine, nx: int, ny: int, spacing: float, lengths: Tuple[float, float]) -> UadmFarm:
    """An nx x ny array of identical rotors, `spacing` apart in x and y, centred in an
    lengths[0] x lengths[1] domain, tower bases at z = 0."""
    x0 = 0.5 * (lengths[0] - (nx - 1) * spacing)
    y0 = 0.5 * (lengths[1] - (ny - 1) * spacing)
    bases: List[Tuple[float, float, float]] = []
    for a in range(nx):
        for b in range(ny):
            bases.append((x0 + a * spacing, y0 + b * spacing, 0.0))
    return make_farm(turbine, bases)


def shard_farm(dx: float = 10.0, lengths=(5120.0, 5120.0), uref: float = 8.0) -> UadmFarm:
    """The 512 x 512 x 192 shard's wind farm: 4 x 4 rotors of D = 126 m at hub height 90 m, 7 D
    apart, Ct = 0.75, Gaussian width eps = 2 dx, force from Uref = 8 m/s."""
    t = UadmTurbine(epsilon=2.0 * dx, ct=0.75, uref=uref)
    return grid_farm(t, 4, 4, 7.0 * 2.0 * t.r_tip, lengths)
