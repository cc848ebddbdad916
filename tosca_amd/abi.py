"""ctypes view of the C ABI declared in include/tosca_rhs.h.

The structures here mirror the header field for field; the loader binds libtosca_rhs.so built
in-tree by tosca_amd/build.py and fails loudly when it is missing (there is no CPU fallback).
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path
from typing import Dict, Mapping, Optional

HERE = Path(__file__).resolve().parent
# TOSCA_LIB selects another build of the library (the host-sanitizer build of the CPU tests)
LIB_PATH = Path(os.environ.get("TOSCA_LIB", str(HERE / "libtosca_rhs.so")))

TOSCA_OK, TOSCA_ERR_ARG, TOSCA_ERR_UNSUPPORTED, TOSCA_ERR_LAYOUT, TOSCA_ERR_DEVICE, TOSCA_ERR_MESH = range(6)
ERROR_NAMES = {0: "OK", 1: "ERR_ARG", 2: "ERR_UNSUPPORTED", 3: "ERR_LAYOUT", 4: "ERR_DEVICE", 5: "ERR_MESH"}

# every function the header declares, with its ctypes signature
P = C.c_void_p
I = C.c_int
D = C.c_double


class DmdaInfo(C.Structure):
    _fields_ = [(n, C.c_int64) for n in
                ("xs", "ys", "zs", "xm", "ym", "zm", "gxs", "gys", "gzs", "gxm", "gym", "gzm", "mx", "my", "mz")]


MESH_FIELDS = ("csi", "eta", "zet", "icsi", "ieta", "izet", "jcsi", "jeta", "jzet", "kcsi", "keta", "kzet",
               "aj", "iaj", "jaj", "kaj", "nvert", "meshtag")


class Mesh(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in MESH_FIELDS]


# FP32 metric copies read by FormU on the mixed path (tosca_mesh_f32)
MESH_F32_FIELDS = ("csi", "eta", "zet", "icsi", "ieta", "izet", "jcsi", "jeta", "jzet", "kcsi", "keta", "kzet",
                   "iaj", "jaj", "kaj")


class MeshF32(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in MESH_F32_FIELDS]


class Params(C.Structure):
    _fields_ = [
        ("periodic", C.c_int * 3),
        ("bc_u", C.c_int * 6),
        ("bc_nut", C.c_int * 6),
        ("nut_value", C.c_double * 6),
        ("div_scheme", C.c_int),
        ("les_active", C.c_int),
        ("les_model", C.c_int),
        ("ibm_active", C.c_int),
        ("overset_active", C.c_int),
        ("nu", C.c_double),
        ("max_cs", C.c_double),
        ("t_ref", C.c_double),
        ("tau_wall_i", C.c_void_p),
        ("tau_wall_j", C.c_void_p),
        ("wall_u", C.c_int * 4),
        ("wall_rough", C.c_double * 4),
    ]


class Inlet(C.Structure):
    _fields_ = [("type", C.c_int), ("cent", C.c_void_p), ("grnd_level", C.c_double), ("ymin", C.c_double),
                ("ly", C.c_double), ("u_tau", C.c_double), ("roughness", C.c_double), ("h_inv", C.c_double),
                ("amplitude", C.c_double), ("periods", C.c_double), ("udir", C.c_double * 3), ("uref", C.c_double * 3),
                ("fc", C.c_double), ("h_ref", C.c_double),
                ("href", C.c_double), ("uprime_rms", C.c_double), ("noise", C.c_void_p),
                ("plane", C.c_void_p), ("n1", C.c_int), ("n2", C.c_int), ("prds1", C.c_int), ("prds2", C.c_int),
                ("merge1", C.c_int), ("avg_top_length", C.c_double), ("avg_top_delta", C.c_double),
                ("avg_top_coords", C.c_double * 10), ("avg_top_u", (C.c_double * 3) * 10),
                ("spline", C.c_int), ("shift2", C.c_int), ("cells", C.c_void_p), ("weights", C.c_void_p),
                ("cells_1", C.c_void_p), ("cells_2", C.c_void_p), ("weights_1", C.c_void_p),
                ("weights_2", C.c_void_p), ("y_ids", C.c_void_p), ("y_weights", C.c_void_p)]


_INLET_PTRS = ("cent", "noise", "plane", "cells", "weights", "cells_1", "cells_2", "weights_1", "weights_2",
               "y_ids", "y_weights")


def make_inlet(d: Mapping) -> "Inlet":
    """tosca_inlet from a dict: type; the pointers (device or host, as int) cent, noise, plane, cells,
    weights, cells_1, cells_2, weights_1, weights_2, y_ids, y_weights; the scalars grnd_level, ymin,
    ly, u_tau, roughness, h_inv, amplitude, periods, fc, h_ref, href, uprime_rms, avg_top_length,
    avg_top_delta; the integers n1, n2, prds1, prds2, merge1, spline, shift2; the vectors udir,
    uref; the tables avg_top_coords (10) and avg_top_u (10 x 3)."""
    x = Inlet()
    x.type = int(d["type"])
    for n in _INLET_PTRS:
        setattr(x, n, d.get(n) or None)
    for n in ("grnd_level", "ymin", "ly", "u_tau", "roughness", "h_inv", "amplitude", "periods", "fc", "h_ref",
              "href", "uprime_rms", "avg_top_length", "avg_top_delta"):
        setattr(x, n, float(d.get(n, 0.0)))
    for n in ("n1", "n2", "prds1", "prds2", "merge1", "spline", "shift2"):
        setattr(x, n, int(d.get(n, 0)))
    x.udir[:] = [float(v) for v in d.get("udir", (1.0, 0.0, 0.0))]
    x.uref[:] = [float(v) for v in d.get("uref", (0.0, 0.0, 0.0))]
    x.avg_top_coords[:] = [float(v) for v in d.get("avg_top_coords", (0.0,) * 10)]
    top = d.get("avg_top_u", ((0.0, 0.0, 0.0),) * 10)
    for n in range(10):
        x.avg_top_u[n][:] = [float(v) for v in top[n]]
    return x


class TeqnParams(C.Structure):
    _fields_ = [("pr", C.c_double), ("wall_t", C.c_int * 4), ("q_wall_i", C.c_void_p), ("q_wall_j", C.c_void_p)]


def make_teqn_params(d: Mapping) -> TeqnParams:
    """tosca_teqn_params from a dict: pr, wall_t (4 flags), q_wall_i / q_wall_j (device or host
    pointers as ints, or None)."""
    t = TeqnParams()
    t.pr = float(d.get("pr", 0.7))
    w = tuple(d.get("wall_t", (0, 0, 0, 0)))
    for q in range(4):
        t.wall_t[q] = int(w[q])
    t.q_wall_i = d.get("q_wall_i") or None
    t.q_wall_j = d.get("q_wall_j") or None
    return t


class KrylovParams(C.Structure):
    _fields_ = [("restart", C.c_int), ("max_it", C.c_int), ("rtol", C.c_double), ("atol", C.c_double)]


class KrylovResult(C.Structure):
    _fields_ = [("iterations", C.c_int), ("residual", C.c_double), ("reason", C.c_int)]


def make_krylov_params(restart=30, max_it=1000, rtol=1e-30, atol=1e-5) -> KrylovParams:
    """kspIMEX settings; the defaults are the reference's (src/ueqn.c:217-282)"""
    k = KrylovParams()
    k.restart, k.max_it, k.rtol, k.atol = int(restart), int(max_it), float(rtol), float(atol)
    return k


class Uadm(C.Structure):
    _fields_ = [("n_rotors", C.c_int), ("n_points", C.c_int), ("nsigma", C.c_double), ("center", C.c_void_p),
                ("sphere", C.c_void_p), ("eps", C.c_void_p), ("points", C.c_void_p), ("force", C.c_void_p)]


SIGNATURES = {
    "tosca_version": (C.c_char_p, []),
    "tosca_last_error": (C.c_char_p, [P]),
    "tosca_ctx_create": (I, [C.POINTER(C.c_void_p), C.POINTER(DmdaInfo), I]),
    "tosca_ctx_destroy": (I, [P]),
    "tosca_ctx_set_stream": (I, [P, P]),
    "tosca_ctx_synchronize": (I, [P]),
    "tosca_ctx_enable_timing": (I, [P, I]),
    "tosca_ctx_timing_reset": (I, [P]),
    "tosca_ctx_timing_count": (I, [P]),
    "tosca_ctx_timing_entry": (I, [P, I, C.c_char_p, I, C.POINTER(C.c_double), C.POINTER(C.c_int64)]),
    "tosca_mesh_check": (I, [P, C.POINTER(Mesh)]),
    "tosca_formu": (I, [P, C.POINTER(Mesh), C.POINTER(Params), P, P, P, P, D, I]),
    "tosca_formu_noalloc": (I, [P, C.POINTER(Mesh), C.POINTER(Params), P, P, P, P, D, I]),
    "tosca_update_cs": (I, [P, C.POINTER(Mesh), C.POINTER(Params), P, P]),
    "tosca_update_nut": (I, [P, C.POINTER(Mesh), C.POINTER(Params), P, P, P, P]),
    "tosca_buoyancy": (I, [P, C.POINTER(Mesh), C.POINTER(Params), P, D, P, P]),
    "tosca_rhs_les_step": (I, [P, C.POINTER(Mesh), C.POINTER(Params), P, P, P, P, D, P, P, P, D]),
    "tosca_rhs_les_step_mixed": (I, [P, C.POINTER(Mesh), C.POINTER(Params), P, P, P, P, D, P, P, P, D]),
    "tosca_rhs_vjp": (I, [P, C.POINTER(Mesh), C.POINTER(Params), P, P, P, P, P, D, D, P, P, P, P]),
    "tosca_rhs_vjp_mixed": (I, [P, C.POINTER(Mesh), C.POINTER(Params), P, P, P, P, P, D, D, P, P, P, P]),
    "tosca_rhs_les_step_mixed_m32": (I, [P, C.POINTER(Mesh), C.POINTER(MeshF32), C.POINTER(Params), P, P, P, P, D, P, P,
                                         P, D]),
    "tosca_explicit_rhs_u": (I, [P, C.POINTER(Mesh), C.POINTER(Params), P, P, P, P, P, P, I, P, P, P, I]),
    "tosca_ueqn_rk3w": (I, [P, C.POINTER(Mesh), C.POINTER(Params), P, P, P, P, P, P, I, P, D, I, P, P, P, I]),
    "tosca_ueqn_rk3w_adjoint": (I, [P, C.POINTER(Mesh), C.POINTER(Params), P, P, P, P, P, P, I, P, D, I, P, P, P, P,
                                    P]),
    "tosca_update_contravariant_bcs": (I, [P, C.POINTER(Mesh), C.POINTER(Params), P, P]),
    "tosca_update_cartesian_bcs": (I, [P, C.POINTER(Mesh), C.POINTER(Params), P, P, P, P]),
    "tosca_update_temperature_bcs": (I, [P, C.POINTER(Mesh), C.POINTER(Params), P, P, P]),
    "tosca_formt": (I, [P, C.POINTER(Mesh), C.POINTER(Params), C.POINTER(TeqnParams), P, P, P, P, P, D, I]),
    "tosca_ueqn_fe": (I, [P, C.POINTER(Mesh), C.POINTER(Params), P, P, P, P, P, P, I, P, P, D, P, I]),
    "tosca_ueqn_rk4": (I, [P, C.POINTER(Mesh), C.POINTER(Params), P, P, P, P, P, P, I, P, P, D, P, P, I]),
    "tosca_imex_matvec": (I, [P, C.POINTER(Mesh), C.POINTER(Params), P, P, P, P, P, D, P]),
    "tosca_grad_p": (I, [P, C.POINTER(Mesh), C.POINTER(Params), P, P]),
    "tosca_cn_rhs": (I, [P, C.POINTER(Params), P, P, P, I, P, P, P]),
    "tosca_snes_func_eval": (I, [P, C.POINTER(Mesh), C.POINTER(Params), P, P, P, P, P, P, D, I, P]),
    "tosca_rhs_les_step_host": (I, [P, C.POINTER(Mesh), C.POINTER(Params), P, P, P, P, D, P, P, P, D, P]),
    "tosca_imex_solve": (I, [P, C.POINTER(Mesh), C.POINTER(Params), P, P, P, P, P, P, D, C.POINTER(KrylovParams),
                             C.POINTER(KrylovResult)]),
    "tosca_ueqn_rk3cn": (I, [P, C.POINTER(Mesh), C.POINTER(Params), I, P, P, P, P, P, P, I, P, P, D, P,
                             C.POINTER(KrylovParams), C.POINTER(KrylovResult)]),
    "tosca_ueqn_abcn": (I, [P, C.POINTER(Mesh), C.POINTER(Params), P, P, P, P, P, P, I, P, P, D, P, P,
                            C.POINTER(KrylovParams), C.POINTER(KrylovResult)]),
    "tosca_project_uadm": (I, [P, C.POINTER(Params), P, C.POINTER(Uadm), P]),
    "tosca_body_force_to_cont": (I, [P, C.POINTER(Mesh), C.POINTER(Params), P, P]),
    "tosca_rhs_les_step_farm": (I, [P, C.POINTER(Mesh), C.POINTER(Params), P, P, P, P, P, D, P, P, P, D]),
    "tosca_rhs_les_step_mixed_farm": (I, [P, C.POINTER(Mesh), C.POINTER(MeshF32), C.POINTER(Params), P, P, P, P, P,
                                          D, P, P, P, D]),
}

_lib: Optional[C.CDLL] = None


def load(path: Optional[Path] = None) -> C.CDLL:
    """Loads libtosca_rhs.so; raises if it has not been built."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise RuntimeError(f"{p} is missing: build the gfx950 library first (python -m tosca_amd.build)")
    lib = C.CDLL(str(p))
    for name, (res, args) in SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    if path is None:
        _lib = lib
    return lib


class ToscaError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERROR_NAMES.get(code, code)}: {msg}")
        self.code = code


def make_info(d: Mapping[str, int]) -> DmdaInfo:
    return DmdaInfo(**{k: int(v) for k, v in d.items()})


def make_params(d: Mapping) -> Params:
    p = Params()
    p.periodic[:] = list(d["periodic"])
    p.bc_u[:] = list(d["bc_u"])
    p.bc_nut[:] = list(d["bc_nut"])
    p.nut_value[:] = [float(v) for v in d.get("nut_value", (0.0,) * 6)]
    p.div_scheme = int(d["div_scheme"])
    p.les_active = int(d["les_active"])
    p.les_model = int(d["les_model"])
    p.ibm_active = int(d.get("ibm_active", 0))
    p.overset_active = int(d.get("overset_active", 0))
    p.nu = float(d["nu"])
    p.max_cs = float(d.get("max_cs", 0.5))
    p.t_ref = float(d["t_ref"])
    p.tau_wall_i = d.get("tau_wall_i") or None
    p.tau_wall_j = d.get("tau_wall_j") or None
    p.wall_u[:] = [int(v) for v in d.get("wall_u", (0, 0, 0, 0))]
    p.wall_rough[:] = [float(v) for v in d.get("wall_rough", (0.0, 0.0, 0.0, 0.0))]
    return p


def make_mesh(ptrs: Mapping[str, int]) -> Mesh:
    m = Mesh()
    for n in MESH_FIELDS:
        setattr(m, n, ptrs.get(n) or None)
    return m


def make_mesh_f32(ptrs: Mapping[str, int]) -> MeshF32:
    m = MeshF32()
    for n in MESH_F32_FIELDS:
        setattr(m, n, ptrs.get(n) or None)
    return m
