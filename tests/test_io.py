"""Mesh-file and field-file formats (tosca_amd.io) on the reference test cases' own mesh files
(tests/golden/meshes) and on known-answer byte strings written here from the format description."""
import struct

import numpy as np
import pytest

from tosca_amd import io
from tosca_amd import case as C
from tests._util import random_nodes

MESHES = "tests/golden/meshes"


@pytest.mark.parametrize("name,periodic,n,extent", [
    ("tgv", (2, 2, 2), (129, 129, 129), ((0, 6.2832), (0, 6.2832), (0, 6.2832))),
    ("neutral_abl", (2, 0, 2), (51, 51, 51), ((0, 1000), (0, 1000), (0, 1000))),
    ("channel395", (2, 0, 2), (65, 65, 65), ((0, 6.2832), (0, 3.1416), (0, 2))),
    ("sphere", (0, 0, 0), (91, 61, 61), ((-10, 20), (-10, 10), (-10, 10))),
])
def test_reference_mesh_files(name, periodic, n, extent):
    per, x, y, z = io.read_mesh_xyz(f"{MESHES}/{name}.xyz")
    assert per == periodic
    assert (len(x), len(y), len(z)) == n
    for a, (lo, hi) in zip((x, y, z), extent):
        assert a[0] == lo and abs(a[-1] - hi) < 1e-9 and np.all(np.diff(a) > 0)


def test_tgv_mesh_is_the_uniform_2pi_lattice():
    """the reference's TGV mesh (coordinates printed with 4 decimals) is the 128^3 box of side 2 pi"""
    _, x, y, z = io.read_mesh_xyz(f"{MESHES}/tgv.xyz")
    ref = np.linspace(0.0, 2 * np.pi, 129)
    for a in (x, y, z):
        assert np.abs(a - ref).max() < 1e-4


def test_xyz_round_trip(tmp_path):
    rng = np.random.Generator(np.random.PCG64(1))
    x, y, z = (np.cumsum(rng.uniform(0.5, 1.5, n)) for n in (5, 7, 6))
    io.write_mesh_xyz(tmp_path / "m.xyz", x, y, z, (1, 0, 2))
    per, x2, y2, z2 = io.read_mesh_xyz(tmp_path / "m.xyz")
    assert per == (1, 0, 2)
    for a, b in ((x, x2), (y, y2), (z, z2)):
        assert np.array_equal(a, b)
    per2, nodes = io.mesh_file_nodes(tmp_path / "m.xyz")
    assert tuple(nodes.shape) == (5, 6, 7, 3)
    assert np.array_equal(nodes[:, 0, 0, 0].numpy(), x) and np.array_equal(nodes[0, 0, :, 1].numpy(), y)
    assert np.array_equal(nodes[0, :, 0, 2].numpy(), z)


def test_grid_known_answer(tmp_path):
    """curvilinear file written by hand in the reference's order: header npz npx npy, then x, y, z,
    each looping i (outer), k, j (inner)"""
    npx, npy, npz = 2, 3, 4
    val = lambda k, j, i, c: 1000 * c + 100 * k + 10 * j + i
    lines = ["-jPeriodicType 1", f"{npz} {npx} {npy}"]
    for c in range(3):
        for i in range(npy):
            for k in range(npx):
                for j in range(npz):
                    lines.append(str(val(k, j, i, c)))
    (tmp_path / "m.grid").write_text("\n".join(lines) + "\n")
    per, nodes = io.read_mesh_grid(tmp_path / "m.grid")
    assert per == (0, 1, 0)
    assert nodes.shape == (npx, npz, npy, 3)
    for k in range(npx):
        for j in range(npz):
            for i in range(npy):
                for c in range(3):
                    assert nodes[k, j, i, c] == val(k, j, i, c)


def test_grid_round_trip(tmp_path):
    nodes = random_nodes((4, 5, 3), 2).numpy()
    io.write_mesh_grid(tmp_path / "m.grid", nodes, (0, 2, 0))
    per, back = io.read_mesh_grid(tmp_path / "m.grid")
    assert per == (0, 2, 0) and np.array_equal(back, nodes)


@pytest.mark.parametrize("int64", [False, True])
def test_petsc_vec_known_answer(tmp_path, int64):
    vals = [1.5, -2.0, 3.25e-7]
    fmt = ">qq" if int64 else ">ii"
    raw = struct.pack(fmt, 1211214, len(vals)) + b"".join(struct.pack(">d", v) for v in vals)
    (tmp_path / "v").write_bytes(raw)
    assert io.read_petsc_vec(tmp_path / "v").tolist() == vals
    io.write_petsc_vec(tmp_path / "w", np.array(vals), int64=int64)
    assert (tmp_path / "w").read_bytes() == raw


def test_petsc_vec_rejects_other_files(tmp_path):
    (tmp_path / "x").write_bytes(b"not a vec file at all")
    with pytest.raises(ValueError):
        io.read_petsc_vec(tmp_path / "x")


def test_fields_round_trip(tmp_path):
    lay = C.Layout(mx=6, my=5, mz=4)
    rng = np.random.Generator(np.random.PCG64(3))
    f = {"U": rng.standard_normal((4, 5, 6, 3)), "V": rng.standard_normal((4, 5, 6, 3)),
         "T": rng.standard_normal((4, 5, 6)), "nut": rng.standard_normal((4, 5, 6))}
    io.write_fields(tmp_path / "fields" / "0.0", f)
    back = io.read_fields(tmp_path / "fields" / "0.0", lay)
    assert set(back) == set(f)
    for k in f:
        assert np.array_equal(back[k], f[k])


def test_mesh_file_case_runs_the_oracle_step():
    from oracle import oracle as O
    from tests._util import np_mesh, to_np
    case = C.mesh_file_case(f"{MESHES}/neutral_abl.xyz", flow="abl")
    lay = case.lay
    assert (lay.mx, lay.my, lay.mz) == (52, 52, 52) and case.periodic == (2, 0, 2)
    out = dict(rhs=np.zeros((lay.mz, lay.my, lay.mx, 3)), cs=np.zeros(lay.gshape), nut=np.zeros(lay.gshape),
               bt=np.zeros((lay.mz, lay.my, lay.mx, 3)))
    O.rhs_les_step(lay.info(), np_mesh(case.mesh), case.params, to_np(case.ucont), to_np(case.ucat),
                   to_np(case.theta), out["rhs"], 1.0, out["cs"], out["nut"], out["bt"], 1.0)
    assert all(np.isfinite(v).all() for v in out.values())
    assert np.abs(out["rhs"]).max() > 0 and np.abs(out["cs"]).max() > 0
