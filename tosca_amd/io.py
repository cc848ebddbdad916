"""On-disk formats of TOSCA cases: mesh files and PETSc-binary field files.

  mesh.xyz  ("cartesian", src/mesh.c:58-96, 232-278): optional periodic-connectivity lines
            "-iPeriodicType n" (read by SetPeriodicConnectivity, src/boundary.c:356-612), then
            "npx npy npz", then npx lines whose first value is x(k), npy lines whose second value is
            y(i), npz lines whose third value is z(j).
  mesh.grid ("curvilinear", src/mesh.c:98-133, 279-327): the same optional lines, then
            "npz npx npy", then every node's x, then every y, then every z, each block looping i
            (outer), k, j (inner).
  fields    (src/io.c:247-420): one PETSc binary file per field (V, U, p, nv, T, cs, nut) under
            fields/<mesh>/<time>/; a Vec file is the class id 1211214 and the length as PetscInt
            (4 or 8 bytes, both big-endian), then the values as big-endian doubles, in the natural
            DMDA order [k][j][i][component], the layout of this library's global arrays.

Node arrays are (npx, npz, npy, 3) = node[k][j][i], as tosca_amd.case.mesh_metrics takes them.
"""
from __future__ import annotations

import struct
from pathlib import Path
from typing import Dict, Optional, Tuple

import numpy as np
import torch

VEC_FILE_CLASSID = 1211214
PERIODIC_KEYS = ("-iPeriodicType", "-jPeriodicType", "-kPeriodicType")


def _split_header(text: str):
    """periodic types (i, j, k) from the leading "-xPeriodicType n" lines, and the remaining tokens"""
    lines = text.splitlines()
    per = {k: 0 for k in PERIODIC_KEYS}
    n = 0
    while n < len(lines) and (not lines[n].strip() or lines[n].lstrip().startswith("-")):
        parts = lines[n].split()
        if len(parts) >= 2 and parts[0] in per:
            per[parts[0]] = int(parts[1])
        n += 1
    tokens = " ".join(lines[n:]).split()
    return (per["-iPeriodicType"], per["-jPeriodicType"], per["-kPeriodicType"]), tokens


def read_mesh_xyz(path) -> Tuple[Tuple[int, int, int], np.ndarray, np.ndarray, np.ndarray]:
    """Periodic types (i, j, k) and the 1-D node coordinates x (along k), y (along i), z (along j)."""
    periodic, tok = _split_header(Path(path).read_text())
    npx, npy, npz = (int(t) for t in tok[:3])
    v = np.array(tok[3:3 + 3 * (npx + npy + npz)], dtype=np.float64).reshape(-1, 3)
    if v.shape[0] != npx + npy + npz:
        raise ValueError(f"{path}: {v.shape[0]} coordinate lines, expected {npx + npy + npz}")
    x = v[:npx, 0].copy()
    y = v[npx:npx + npy, 1].copy()
    z = v[npx + npy:, 2].copy()
    return periodic, x, y, z


def write_mesh_xyz(path, x, y, z, periodic=(0, 0, 0)):
    lines = [f"{k} {p}" for k, p in zip(PERIODIC_KEYS, periodic) if p]
    lines.append(f"{len(x)} {len(y)} {len(z)}")
    lines += [f"{v:.17g} 0 0" for v in x] + [f"0 {v:.17g} 0" for v in y] + [f"0 0 {v:.17g}" for v in z]
    Path(path).write_text("\n".join(lines) + "\n")


def read_mesh_grid(path) -> Tuple[Tuple[int, int, int], np.ndarray]:
    """Periodic types and the node array (npx, npz, npy, 3) of a curvilinear mesh file."""
    periodic, tok = _split_header(Path(path).read_text())
    npz, npx, npy = (int(t) for t in tok[:3])
    n = npx * npy * npz
    v = np.array(tok[3:3 + 3 * n], dtype=np.float64)
    if v.size != 3 * n:
        raise ValueError(f"{path}: {v.size} values, expected {3 * n}")
    nodes = np.empty((npx, npz, npy, 3))
    for c in range(3):          # block c loops i, k, j (j fastest)
        nodes[..., c] = v[c * n:(c + 1) * n].reshape(npy, npx, npz).transpose(1, 2, 0)
    return periodic, nodes


def write_mesh_grid(path, nodes: np.ndarray, periodic=(0, 0, 0)):
    npx, npz, npy, _ = nodes.shape
    lines = [f"{k} {p}" for k, p in zip(PERIODIC_KEYS, periodic) if p]
    lines.append(f"{npz} {npx} {npy}")
    for c in range(3):
        blk = nodes[..., c].transpose(2, 0, 1).ravel()
        lines += [f"{v:.17g}" for v in blk]
    Path(path).write_text("\n".join(lines) + "\n")


def mesh_file_nodes(path, device="cpu") -> Tuple[Tuple[int, int, int], torch.Tensor]:
    """(periodic, nodes) of a mesh.xyz or mesh.grid file, nodes as a float64 tensor."""
    p = Path(path)
    if p.suffix == ".xyz":
        periodic, x, y, z = read_mesh_xyz(p)
        from tosca_amd.case import cartesian_nodes
        return periodic, cartesian_nodes(x, y, z, device)
    if p.suffix == ".grid":
        periodic, nodes = read_mesh_grid(p)
        return periodic, torch.as_tensor(nodes, device=device)
    raise ValueError(f"{path}: unknown mesh file type (expected .xyz or .grid)")


def read_petsc_vec(path, shape: Optional[Tuple[int, ...]] = None) -> np.ndarray:
    """Values of a PETSc binary Vec file (32- or 64-bit PetscInt header), optionally reshaped."""
    raw = Path(path).read_bytes()
    if len(raw) >= 8 and struct.unpack(">i", raw[:4])[0] == VEC_FILE_CLASSID:
        n, off = struct.unpack(">i", raw[4:8])[0], 8
    elif len(raw) >= 16 and struct.unpack(">q", raw[:8])[0] == VEC_FILE_CLASSID:
        n, off = struct.unpack(">q", raw[8:16])[0], 16
    else:
        raise ValueError(f"{path}: not a PETSc binary Vec file")
    if len(raw) < off + 8 * n:
        raise ValueError(f"{path}: {len(raw) - off} bytes of values, expected {8 * n}")
    v = np.frombuffer(raw, dtype=">f8", count=n, offset=off).astype(np.float64)
    return v.reshape(shape) if shape is not None else v


def write_petsc_vec(path, values: np.ndarray, int64: bool = True):
    v = np.ascontiguousarray(values, dtype=np.float64).ravel()
    head = struct.pack(">qq", VEC_FILE_CLASSID, v.size) if int64 else struct.pack(">ii", VEC_FILE_CLASSID, v.size)
    Path(path).write_bytes(head + v.astype(">f8").tobytes())


FIELD_DOF = {"V": 3, "U": 3, "p": 1, "nv": 1, "T": 1, "cs": 1, "nut": 1}


def read_fields(directory, lay) -> Dict[str, np.ndarray]:
    """The restart fields present in fields/<mesh>/<time>/ (readFields, src/io.c:247-420), as global
    arrays [mz][my][mx](, 3)."""
    out = {}
    for name, dof in FIELD_DOF.items():
        f = Path(directory) / name
        if f.exists():
            shape = (lay.mz, lay.my, lay.mx) + ((3,) if dof == 3 else ())
            out[name] = read_petsc_vec(f, shape)
    return out


def write_fields(directory, fields: Dict[str, np.ndarray], int64: bool = True):
    d = Path(directory)
    d.mkdir(parents=True, exist_ok=True)
    for name, v in fields.items():
        if name not in FIELD_DOF:
            raise ValueError(f"unknown field {name}")
        write_petsc_vec(d / name, v, int64)
