#!/usr/bin/env python3
"""Benchmark: momentum right-hand side + LES cell-updates/s on one MI355X.

One step = one evaluation of UpdateCs -> UpdateNut -> Buoyancy -> FormU (formMode 0) over every
interior cell of the workload (tosca_rhs_les_step), with all inputs resident in HBM; on the
512x512x192 shard also the wind farm's source (the VecAXPY of sourceFarmCont, a 4 x 4 actuator-disk
farm, tosca_rhs_les_step_farm).

    python bench.py [--gpus 1] [--steps 20] [--warmup 3] [--config 5] [--path fused|ref]

Prints ONE JSON line (rank 0).  Only one GPU is supported: --gpus > 1 exits with a message.
"""
from __future__ import annotations

import argparse
import subprocess
import json
import math
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "RHS+LES cell-updates/s (FP64 vs mixed) at 1 GPU, %HBM roofline, VJP/fwd time"
HBM_PEAK_GBS = 8000.0

# Algorithmic bytes per interior cell of each kernel: compulsory traffic of the arrays the kernel
# reads (once) and writes (once), 8-byte doubles, a read-modify-write output counted twice.
# Cell metrics csi/eta/zet = 72 B, face metrics i/j/k x csi/eta/zet = 216 B, face Jacobians 24 B.
KERNEL_BYTES = {
    # fused path (tosca_rhs_les_step with DSM)
    # FormU + Buoyancy: ucont, ucat, nut, face metrics, face Jacobians, cell metrics, theta,
    # rhs (+=, read and written), bTheta (written)
    "k_fused_formu": 24 + 24 + 8 + 216 + 24 + 72 + 8 + 2 * 24 + 24,      # = 448
    "k_dsm_lm": 24 + 72 + 8 + 3 * 8,                      # ucat, cell metrics, aj -> LM, MM, |S|
    "k_dsm_cs_nut": 3 * 8 + 8 + 2 * 8,                    # LM, MM, |S|, aj -> Cs, nut
    "k_nut_fused": 24 + 72 + 8 + 8 + 8,                   # ucat, cell metrics, aj, Cs -> nut
    # reference-structured path
    "k_formu_faces": 24 + 24 + 216 + 24 + 8 + 144,        # ucont, ucat, face metrics, iaj/jaj/kaj, nut -> 6 flux vectors
    "k_formu_cells": 144 + 24,                            # 6 flux vectors -> fp
    "k_formu_dot": 24 + 72 + 24 + 2 * 24,                 # fp, csi/eta/zet, iaj/jaj/kaj, rhs (+=)
    "k_strain": 24 + 72 + 8 + 80,                         # ucat, cell metrics, aj -> grad u (9) + |S|
    "k_test_filter3": 24 + 72 + 8 + 80 + 16,              # ucat, cell metrics, aj, strain -> LM, MM
    "k_box_clip": 16 + 8 + 8,                             # LM, MM, aj -> Cs
    "k_nut": 24 + 72 + 8 + 8 + 8,                         # ucat, cell metrics, aj, Cs -> nut
    "k_buoyancy": 8 + 72 + 24,                            # theta, icsi/jeta/kzet -> bTheta
    # plane-averaged dynamic model (DPASD), fused: the 5-point filter's 28 (i, j)-filtered sums
    # pass through HBM once (written by k_dpasd_ij, read by k_dpasd_k)
    "k_dpasd_ij": 24 + 72 + 8 + 28 * 8,                   # ucat, cell metrics, aj -> 28 sums
    "k_dpasd_k": 28 * 8 + 2 * 8 + 72 + 8,                 # 28 sums, LM, MM, cell metrics, aj -> plane partials
    "k_dpasd_cs_nut": 8 + 8 + 2 * 8,                      # |S|, aj -> Cs, nut
    "k_dpasd_ij_f32": 12 + 72 + 8 + 28 * 4,
    "k_dpasd_k_f32": 28 * 4 + 2 * 8 + 72 + 8,
}
# Whole step: every ABI input read once, every output written once (rhs read + written);
# nvert / meshTag are validated once per mesh (tosca_mesh_check) and not read per step.
STEP_BYTES = (24 + 24 + 216 + 72 + 24 + 8 + 8 + 24) + (24 + 8 + 8 + 24)   # = 464
# The wind-farm source of the shard workload (sourceFarmCont, a global vector) is read once more by
# FormU's kernel and by the step: + 24 B/cell each.
FARM_BYTES = 24


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", type=int, default=5)
    ap.add_argument("--path", default=os.environ.get("TOSCA_PATH", "default"))
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="budget of the CPU-oracle baseline")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-farm", action="store_true", help="shard workload without its wind-farm source")
    ap.add_argument("--no-configs", action="store_true", help="skip the step timings of the other configurations")
    return ap.parse_args()


def build_case(cfg: int, device: str):
    from tosca_amd import case as C
    spec = C.CONFIGS[cfg]
    return spec["builder"](device=device, **spec["kwargs"])


WORKLOAD_TAG = {1: "tgv64", 2: "hill256x128x96", 3: "abl384x256x128", 4: "hill192x192x128", 5: "shard512x512x192"}


def copy_peak_gbs(torch, nbytes: int = 1 << 31, reps: int = 10) -> float:
    """Device-to-device copy rate (read + write bytes / time) of torch's copy kernel over nbytes,
    the practical HBM ceiling the kernels are compared against besides the 8 TB/s peak."""
    a = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del a, b
    return 2 * nbytes / (ms * 1e-3) / 1e9


F32_METRICS = ("csi", "eta", "zet", "icsi", "ieta", "izet", "jcsi", "jeta", "jzet", "kcsi", "keta", "kzet",
               "iaj", "jaj", "kaj")


def farm_leg(torch, ctx, mesh, case, cent, lay, args):
    """The 4 x 4 actuator-disk farm of the shard: the Gaussian projection of the disks' force
    (tosca_project_uadm) and its face fluxes (tosca_body_force_to_cont), made once per time step
    in the reference (UpdateWindTurbines); the step then reads the face fluxes (sourceFarmCont)."""
    from tosca_amd import farm as F
    from tosca_amd.rhs import DeviceFarm
    lengths = case.meta["lengths"]
    dx = lengths[0] / (lay.mz - 2)
    farm = F.shard_farm(dx=dx, lengths=(lengths[0], lengths[1]))
    dfarm = DeviceFarm(farm)
    scat = torch.zeros(lay.gshape + (3,), dtype=torch.float64, device="cuda")
    scont = torch.zeros((lay.mz, lay.my, lay.mx, 3), dtype=torch.float64, device="cuda")
    e0, e1, e2 = (torch.cuda.Event(enable_timing=True) for _ in range(3))
    ctx.ProjectUADM(case.params, cent, dfarm, scat)
    ctx.BodyForceToCont(mesh, case.params, scat, scont)
    torch.cuda.synchronize()
    e0.record()
    ctx.ProjectUADM(case.params, cent, dfarm, scat)
    e1.record()
    ctx.BodyForceToCont(mesh, case.params, scat, scont)
    e2.record()
    torch.cuda.synchronize()
    proj_ms, conv_ms = e0.elapsed_time(e1), e1.elapsed_time(e2)
    print(f"# farm: {farm.n_rotors} rotors x {farm.n_points} points, projection {proj_ms:.3f} ms, "
          f"face fluxes {conv_ms:.3f} ms", file=sys.stderr, flush=True)
    del scat
    info = {"rotors": farm.n_rotors, "points_per_rotor": farm.n_points, "model": "uniformADM, Ct 0.75, Uref 8 m/s",
            "rotor_diameter_m": 126.0, "hub_height_m": 90.0, "spacing_m": 7 * 126.0, "epsilon_m": float(farm.eps[0]),
            "projection_ms": proj_ms, "face_flux_ms": conv_ms}
    return scont, info


def mixed_leg(torch, ctx, mesh, case, ucont, ucat, theta, lay, args, fp64_ms, farm=None):
    """The same step with FP32 flow fields: tosca_rhs_les_step_mixed (FP64 metrics) and
    tosca_rhs_les_step_mixed_m32 (FormU on FP32 metric copies, made once before the timing, as the
    mesh is static).  Throughput, speed-up over the FP64 step, and per-field errors of one step
    against the FP64 step on the original FP64 inputs."""
    dev = "cuda"
    V32, u32, t32 = ucont.float().contiguous(), ucat.float().contiguous(), theta.float().contiguous()
    gs, gl = lay.gshape, (lay.mz, lay.my, lay.mx, 3)
    rhs = torch.zeros(gl, dtype=torch.float64, device=dev)
    bt = torch.zeros(gl, dtype=torch.float64, device=dev)
    cs = torch.zeros(gs, dtype=torch.float64, device=dev)
    nut = torch.zeros(gs, dtype=torch.float32, device=dev)
    rhs64 = torch.zeros(gl, dtype=torch.float64, device=dev)
    bt64 = torch.zeros_like(rhs64)
    cs64 = torch.zeros(gs, dtype=torch.float64, device=dev)
    nut64 = torch.zeros(gs, dtype=torch.float64, device=dev)
    if farm is None:
        ctx.rhs_les_step(mesh, case.params, ucont, ucat, theta, rhs64, 1.0, cs64, nut64, bt64, 1.0)
    else:
        ctx.rhs_les_step_farm(mesh, case.params, ucont, ucat, theta, farm, rhs64, 1.0, cs64, nut64, bt64, 1.0)

    def mixed(m32=None):
        if farm is None and m32 is None:
            return lambda: ctx.rhs_les_step_mixed(mesh, case.params, V32, u32, t32, rhs, 1.0, cs, nut, bt, 1.0)
        if farm is None:
            return lambda: ctx.rhs_les_step_mixed_m32(mesh, m32, case.params, V32, u32, t32, rhs, 1.0, cs, nut, bt, 1.0)
        return lambda: ctx.rhs_les_step_mixed_farm(mesh, case.params, V32, u32, t32, farm, rhs, 1.0, cs, nut, bt, 1.0,
                                                   mesh32=m32)

    def measure(step):
        for _ in range(args.warmup):
            step()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(args.steps):
            step()
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / args.steps
        ctx.enable_timing(True)
        ctx.reset_timing()
        for _ in range(args.steps):
            step()
        tim = ctx.timings()
        ctx.enable_timing(False)
        rhs.zero_()
        step()
        torch.cuda.synchronize()
        err = {}
        for k, a, b in (("rhs", rhs, rhs64), ("cs", cs, cs64), ("nut", nut.double(), nut64), ("btheta", bt, bt64)):
            d = (a - b).double()
            nb = float(torch.linalg.vector_norm(b))
            err[k] = {"rel_l2": float(torch.linalg.vector_norm(d)) / max(nb, 1e-300),
                      "max_rel": float(d.abs().max()) / max(float(b.abs().max()), 1e-300),
                      "rms_rel": float(d.pow(2).mean().sqrt()) / max(float(b.pow(2).mean().sqrt()), 1e-300)}
        return ms, {k: v[0] / v[1] for k, v in tim.items()}, err

    out = {}
    ms, kms, err = measure(mixed())
    value = lay.ncells / (ms * 1e-3)
    print(f"# mixed: {ms:.3f} ms/step, {value:.3e} cells/s, x{fp64_ms / ms:.2f} vs FP64; rel L2 errors "
          + ", ".join(f"{k} {v['rel_l2']:.1e}" for k, v in err.items()), file=sys.stderr, flush=True)
    out.update({"value": value, "unit": "cell-updates/s", "ms_per_step": ms, "speedup_vs_fp64": fp64_ms / ms,
                "dtype": "f32 flow fields / f64 metrics and reductions", "errors_vs_fp64": err, "kernels_ms": kms})
    mesh32 = {k: mesh[k].float().contiguous() for k in F32_METRICS}
    ms, kms, err = measure(mixed(mesh32))
    value = lay.ncells / (ms * 1e-3)
    print(f"# mixed, FP32 metric copies: {ms:.3f} ms/step, {value:.3e} cells/s, x{fp64_ms / ms:.2f} vs FP64; rel L2 "
          "errors " + ", ".join(f"{k} {v['rel_l2']:.1e}" for k, v in err.items()), file=sys.stderr, flush=True)
    out["fp32_metrics"] = {"value": value, "unit": "cell-updates/s", "ms_per_step": ms, "speedup_vs_fp64": fp64_ms / ms,
                           "dtype": "f32 flow fields, f32 FormU metric copies / f64 LES metrics and reductions",
                           "errors_vs_fp64": err, "kernels_ms": kms}
    del mesh32
    return out


def vjp_leg(torch, ctx, mesh, case, ucont, ucat, theta, cs, lay, args, step_ms, mixed_ms=None):
    """Adjoint cost: tosca_rhs_vjp against the forward it differentiates (UpdateNut + FormU +
    Buoyancy with the step's Cs) and against the whole step (which also runs the dynamic model);
    the mixed adjoint (tosca_rhs_vjp_mixed, FP32 flow fields and packets) against the FP64 adjoint
    and the mixed step, with its difference from the FP64 adjoint at the same FP32-rounded inputs."""
    dev = "cuda"
    gs, gl = lay.gshape, (lay.mz, lay.my, lay.mx, 3)
    nut = torch.zeros(gs, dtype=torch.float64, device=dev)
    r = torch.zeros(gl, dtype=torch.float64, device=dev)
    bt = torch.zeros_like(r)
    rbar = torch.randn(gl, dtype=torch.float64, device=dev, generator=torch.Generator(device=dev).manual_seed(4))
    bars = [torch.empty_like(t) for t in (ucont, ucat, theta, cs)]

    def fwd():
        ctx.UpdateNut(mesh, case.params, ucat, cs, nut)
        ctx.FormU(mesh, case.params, ucont, ucat, nut, r, 1.0, 0)
        ctx.Buoyancy(mesh, case.params, theta, 1.0, bt)

    def vjp():
        ctx.rhs_vjp(mesh, case.params, ucont, ucat, theta, cs, rbar, 1.0, 1.0, *bars)

    def timed(fn):
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(args.steps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / args.steps

    def kernel_times(fn):
        # per-kernel device time over `steps` calls (no host gaps between the calls)
        ctx.enable_timing(True)
        ctx.reset_timing()
        for _ in range(args.steps):
            fn()
        t = ctx.timings()
        ctx.enable_timing(False)
        return t, sum(v[0] for v in t.values()) / args.steps

    fwd_ms, vjp_ms = timed(fwd), timed(vjp)
    tim, vjp_k = kernel_times(vjp)
    _, fwd_k = kernel_times(fwd)
    print(f"# vjp: {vjp_ms:.3f} ms vs forward {fwd_ms:.3f} ms (x{vjp_ms / fwd_ms:.2f}), vs step x{vjp_ms / step_ms:.2f}; "
          f"kernel time {vjp_k:.3f} vs {fwd_k:.3f} ms (x{vjp_k / fwd_k:.2f})", file=sys.stderr, flush=True)
    out = {"vjp_ms": vjp_ms, "forward_ms": fwd_ms, "vjp_over_forward": vjp_ms / fwd_ms,
           "vjp_over_step": vjp_ms / step_ms, "kernels_ms": {k: v[0] / v[1] for k, v in tim.items()},
           "vjp_kernels_ms": vjp_k, "forward_kernels_ms": fwd_k, "vjp_over_forward_kernels": vjp_k / fwd_k,
           "forward": "UpdateNut + FormU + Buoyancy (FP64) with the step's Cs"}
    # mixed adjoint
    x32 = [t.float().contiguous() for t in (ucont, ucat, theta)]
    bars32 = [torch.empty_like(t) for t in x32] + [torch.empty_like(cs)]

    def vjp32():
        ctx.rhs_vjp_mixed(mesh, case.params, *x32, cs, rbar, 1.0, 1.0, *bars32)

    m_ms = timed(vjp32)
    ctx.enable_timing(True)
    ctx.reset_timing()
    for _ in range(args.steps):
        vjp32()
    tim32 = ctx.timings()
    ctx.enable_timing(False)
    ctx.rhs_vjp(mesh, case.params, *[t.double() for t in x32], cs, rbar, 1.0, 1.0, *bars)
    vjp32()
    torch.cuda.synchronize()
    err = {}
    for k, a, b in zip(("ucont", "ucat", "theta", "cs"), bars32, bars):
        err[k] = float(torch.linalg.vector_norm(a.double() - b)) / max(float(torch.linalg.vector_norm(b)), 1e-300)
    out["mixed"] = {"vjp_ms": m_ms, "speedup_vs_fp64_vjp": vjp_ms / m_ms,
                    "vjp_over_mixed_step": (m_ms / mixed_ms) if mixed_ms else None,
                    "rel_l2_vs_fp64_vjp_same_inputs": err, "kernels_ms": {k: v[0] / v[1] for k, v in tim32.items()}}
    print(f"# vjp mixed: {m_ms:.3f} ms (x{vjp_ms / m_ms:.2f} vs FP64 adjoint"
          + (f", x{m_ms / mixed_ms:.2f} the mixed step" if mixed_ms else "") + "); rel L2 vs FP64 adjoint "
          + ", ".join(f"{k} {v:.1e}" for k, v in err.items()), file=sys.stderr, flush=True)
    return out


def dpasd_leg(torch, ctx, mesh, case, ucont, ucat, theta, lay, args, farm=None):
    """The same step with the plane-averaged dynamic model (DPASD: the 3- and 5-point test filters
    and the j-plane means) on the fused kernels, FP64 and mixed, with the per-kernel times and
    algorithmic rates of its kernels; and UpdateCs alone on the reference-structured kernels (a
    gathered 5 x 5 x 5 stencil per cell), timed over a few calls."""
    dev = "cuda"
    params = dict(case.params, les_model=5)
    gs, gl = lay.gshape, (lay.mz, lay.my, lay.mx, 3)
    rhs = torch.zeros(gl, dtype=torch.float64, device=dev)
    bt = torch.zeros_like(rhs)
    cs = torch.zeros(gs, dtype=torch.float64, device=dev)
    nut = torch.zeros(gs, dtype=torch.float64, device=dev)
    V32, u32, t32 = ucont.float().contiguous(), ucat.float().contiguous(), theta.float().contiguous()
    cs32 = torch.zeros_like(cs)
    nut32 = torch.zeros(gs, dtype=torch.float32, device=dev)

    def fp64():
        if farm is None:
            ctx.rhs_les_step(mesh, params, ucont, ucat, theta, rhs, 1.0, cs, nut, bt, 1.0)
        else:
            ctx.rhs_les_step_farm(mesh, params, ucont, ucat, theta, farm, rhs, 1.0, cs, nut, bt, 1.0)

    def mixed():
        if farm is None:
            ctx.rhs_les_step_mixed(mesh, params, V32, u32, t32, rhs, 1.0, cs32, nut32, bt, 1.0)
        else:
            ctx.rhs_les_step_mixed_farm(mesh, params, V32, u32, t32, farm, rhs, 1.0, cs32, nut32, bt, 1.0)

    def timed(fn, reps):
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        ctx.enable_timing(True)
        ctx.reset_timing()
        for _ in range(reps):
            fn()
        tim = ctx.timings()
        ctx.enable_timing(False)
        return ms, {k: v[0] / v[1] for k, v in tim.items()}

    out = {}
    for name, fn in (("fp64", fp64), ("mixed", mixed)):
        ms, kms = timed(fn, args.steps)
        rates = {k: KERNEL_BYTES[k] * lay.ncells / (v * 1e-3) / 1e9 for k, v in kms.items()
                 if k.startswith("k_dpasd") and k in KERNEL_BYTES}
        out[name] = {"ms_per_step": ms, "value": lay.ncells / (ms * 1e-3), "unit": "cell-updates/s",
                     "kernels_ms": kms, "algorithmic_GBs": rates,
                     "frac_of_hbm_peak": {k: v / HBM_PEAK_GBS for k, v in rates.items()}}
        print(f"# dpasd {name}: {ms:.3f} ms/step, {lay.ncells / (ms * 1e-3):.3e} cells/s; "
              + ", ".join(f"{k} {kms[k]:.3f} ms ({r:.0f} GB/s)" for k, r in rates.items()), file=sys.stderr, flush=True)
    out["mixed"]["speedup_vs_fp64"] = out["fp64"]["ms_per_step"] / out["mixed"]["ms_per_step"]
    d = (cs32 - cs)[3:-3, 3:-3, 3:-3]
    out["mixed"]["cs_rel_l2_vs_fp64"] = float(torch.linalg.vector_norm(d)) / max(
        float(torch.linalg.vector_norm(cs[3:-3, 3:-3, 3:-3])), 1e-300)
    # UpdateCs alone, fused and on the reference-structured kernels
    ms_f, _ = timed(lambda: ctx.UpdateCs(mesh, params, ucat, cs), 5)
    old = os.environ.get("TOSCA_PATH")
    os.environ["TOSCA_PATH"] = "ref"
    from tosca_amd.rhs import RhsContext
    rctx = RhsContext(lay.info())
    if old is None:
        os.environ.pop("TOSCA_PATH")
    else:
        os.environ["TOSCA_PATH"] = old
    rctx.mesh_check(mesh)
    cs_r = torch.zeros_like(cs)
    rctx.UpdateCs(mesh, params, ucat, cs_r)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(2):
        rctx.UpdateCs(mesh, params, ucat, cs_r)
    e1.record()
    torch.cuda.synchronize()
    ms_r = e0.elapsed_time(e1) / 2
    dr = (cs_r - cs)[3:-3, 3:-3, 3:-3]
    rel = float(torch.linalg.vector_norm(dr)) / max(float(torch.linalg.vector_norm(cs_r[3:-3, 3:-3, 3:-3])), 1e-300)
    rctx.close()
    del rctx
    out["update_cs"] = {"fused_ms": ms_f, "reference_structured_ms": ms_r, "speedup": ms_r / ms_f,
                        "cs_rel_l2_fused_vs_reference_structured": rel}
    print(f"# dpasd mixed x{out['mixed']['speedup_vs_fp64']:.2f} (Cs rel L2 {out['mixed']['cs_rel_l2_vs_fp64']:.1e}); "
          f"UpdateCs fused {ms_f:.2f} ms vs reference-structured {ms_r:.1f} ms (Cs rel L2 {rel:.1e})",
          file=sys.stderr, flush=True)
    return out


def configs_leg(torch, args, skip: int):
    """The step (FP64 and mixed, FP64 metrics) on the other benchmark configurations at their full
    sizes, each built on the GPU and timed like the main workload; and the adjoint against the
    forward it differentiates on the 192x192x128 hill, where the adjoint ratio is quoted."""
    from tosca_amd import case as C
    from tosca_amd.rhs import RhsContext
    out = {}
    for cfg in (1, 2, 3, 4):
        if cfg == skip:
            continue
        t0 = time.time()
        case = build_case(cfg, "cuda")
        lay = case.lay
        mesh = {k: v.contiguous() for k, v in case.mesh.items() if k != "cent"}
        ctx = RhsContext(lay.info())
        ctx.mesh_check(mesh)
        gs, gl = lay.gshape, (lay.mz, lay.my, lay.mx, 3)
        rhs = torch.zeros(gl, dtype=torch.float64, device="cuda")
        bt = torch.zeros_like(rhs)
        cs = torch.zeros(gs, dtype=torch.float64, device="cuda")
        nut = torch.zeros_like(cs)
        x32 = [t.float().contiguous() for t in (case.ucont, case.ucat, case.theta)]
        nut32 = torch.zeros(gs, dtype=torch.float32, device="cuda")

        def timed(fn, reps):
            for _ in range(args.warmup):
                fn()
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(reps):
                fn()
            e1.record()
            torch.cuda.synchronize()
            return e0.elapsed_time(e1) / reps

        ms = timed(lambda: ctx.rhs_les_step(mesh, case.params, case.ucont, case.ucat, case.theta, rhs, 1.0, cs, nut, bt,
                                            1.0), args.steps)
        ms32 = timed(lambda: ctx.rhs_les_step_mixed(mesh, case.params, *x32, rhs, 1.0, cs, nut32, bt, 1.0), args.steps)
        r = {"workload": workload_name(cfg, case), "cells": lay.ncells, "ms_per_step": ms,
             "value": lay.ncells / (ms * 1e-3), "unit": "cell-updates/s",
             "step_frac_of_hbm_peak": STEP_BYTES * lay.ncells / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
             "mixed_ms_per_step": ms32, "mixed_speedup": ms / ms32}
        if cfg == 4:
            rbar = torch.randn(gl, dtype=torch.float64, device="cuda", generator=torch.Generator(device="cuda").manual_seed(4))
            bars = [torch.empty_like(t) for t in (case.ucont, case.ucat, case.theta, cs)]

            def fwd():
                ctx.UpdateNut(mesh, case.params, case.ucat, cs, nut)
                ctx.FormU(mesh, case.params, case.ucont, case.ucat, nut, rhs, 1.0, 0)
                ctx.Buoyancy(mesh, case.params, case.theta, 1.0, bt)

            f_ms = timed(fwd, args.steps)
            v_ms = timed(lambda: ctx.rhs_vjp(mesh, case.params, case.ucont, case.ucat, case.theta, cs, rbar, 1.0, 1.0,
                                             *bars), args.steps)
            r.update({"vjp_ms": v_ms, "forward_ms": f_ms, "vjp_over_forward": v_ms / f_ms})
        out[WORKLOAD_TAG[cfg]] = r
        print(f"# config {cfg} ({WORKLOAD_TAG[cfg]}, built in {time.time() - t0:.1f} s): {ms:.3f} ms/step, "
              f"{r['value']:.3e} cells/s, mixed x{r['mixed_speedup']:.2f}"
              + (f", vjp/forward x{r['vjp_over_forward']:.2f}" if cfg == 4 else ""), file=sys.stderr, flush=True)
        ctx.close()
        del ctx, case, mesh, rhs, bt, cs, nut, x32, nut32
        torch.cuda.empty_cache()
    return out


def rk3w_leg(torch, ctx, mesh, case, ucont, ucat, nut, farm, lay, args, n_stages=10):
    """The checkpointed adjoint of the explicit Wray Runge-Kutta predictor over n_stages stages
    (tosca_ueqn_rk3w_adjoint: forward sweep storing each stage's state, then the backward sweep)
    against the forward predictor alone (tosca_ueqn_rk3w) over the same stages; the step's nu_t
    held, -dP, bTheta and the farm (when the workload has it) as sources."""
    dev = "cuda"
    gl = (lay.mz, lay.my, lay.mx, 3)
    g = torch.Generator(device=dev).manual_seed(13)
    dp, bt = (0.01 * torch.randn(gl, dtype=torch.float64, device=dev, generator=g) for _ in range(2))
    w = [torch.empty(gl, dtype=torch.float64, device=dev) for _ in range(3)]
    V, U = ucont.clone(), ucat.clone()
    xbar = torch.randn(gl, dtype=torch.float64, device=dev, generator=g)
    ubar = torch.empty_like(xbar)
    nbar = torch.empty_like(nut)
    dt = 0.1
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)

    def fwd():
        V.copy_(ucont)
        U.copy_(ucat)
        ctx.UeqnRK3W(mesh, case.params, V, U, nut, dp, bt, None, 0, farm, dt, n_stages, *w)

    def adj():
        V.copy_(ucont)
        U.copy_(ucat)
        ubar.copy_(xbar)
        ctx.UeqnRK3W_adjoint(mesh, case.params, V, U, nut, dp, bt, None, 0, farm, dt, n_stages, *w, ubar, nbar)

    def timed(fn, reps):
        fn()
        torch.cuda.synchronize()
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    reps = max(1, args.steps // 5)
    fwd_ms, adj_ms = timed(fwd, reps), timed(adj, reps)
    ctx.enable_timing(True)
    ctx.reset_timing()
    adj()
    tim = ctx.timings()
    ctx.enable_timing(False)
    print(f"# rk3w adjoint, {n_stages} stages: {adj_ms:.1f} ms (forward sweep + backward) vs forward {fwd_ms:.1f} ms "
          f"(x{adj_ms / fwd_ms:.2f})", file=sys.stderr, flush=True)
    del dp, bt, w, V, U, xbar, ubar, nbar
    return {"n_stages": n_stages, "dt": dt, "forward_ms": fwd_ms, "adjoint_ms": adj_ms,
            "adjoint_over_forward": adj_ms / fwd_ms, "kernels_ms": {k: v[0] for k, v in tim.items()},
            "function": "final Ucont of the explicit Wray RK3 predictor as a function of the initial Ucont and "
                        "nu_t; every stage's state checkpointed in HBM, both sweeps timed"}


# ExplicitRhsU algorithmic bytes per cell: contravariant-to-Cartesian (ucont 24 + cell metrics 72
# in, ucat 24 out), Rhs assembly (dP, bTheta, bTheta_o, farm 96 in, Rhs 24 out), FormU without
# Buoyancy (ucont 24, ucat 24, nu_t 8, face metrics 216, face Jacobians 24, cell metrics 72, Rhs 48
# read + write); the boundary-patch resets touch planes only
EXPLICIT_BYTES = (24 + 72 + 24) + (96 + 24) + (24 + 24 + 8 + 216 + 24 + 72 + 48)   # = 656


def explicit_leg(torch, ctx, mesh, case, ucont, ucat, nut, lay, args):
    """ExplicitRhsU (tosca_explicit_rhs_u) with every source term on: the caller of the path.
    It resets ucont and recomputes ucat in place, so it runs on copies of the step's fields."""
    dev = "cuda"
    gl = (lay.mz, lay.my, lay.mx, 3)
    g = torch.Generator(device=dev).manual_seed(11)
    dp, bt, bto, farm, uo = (torch.randn(gl, dtype=torch.float64, device=dev, generator=g) for _ in range(5))
    V, U = ucont.clone(), ucat.clone()
    rhs = torch.empty(gl, dtype=torch.float64, device=dev)

    def run():
        ctx.ExplicitRhsU(mesh, case.params, V, U, nut, dp, bt, bto, 1, farm, uo, rhs, 0)

    for _ in range(args.warmup):
        run()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(args.steps):
        run()
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / args.steps
    ctx.enable_timing(True)
    ctx.reset_timing()
    for _ in range(args.steps):
        run()
    tim = ctx.timings()
    ctx.enable_timing(False)
    value = lay.ncells / (ms * 1e-3)
    gbs = EXPLICIT_BYTES * value / 1e9
    print(f"# explicit_rhs_u: {ms:.3f} ms, {value:.3e} cells/s, {gbs:.0f} GB/s algorithmic", file=sys.stderr, flush=True)
    # the four-stage Runge-Kutta predictor built on it (UeqnRK4), one call per step
    utmp = torch.empty(gl, dtype=torch.float64, device=dev)
    for _ in range(max(1, args.warmup // 2)):
        ctx.UeqnRK4(mesh, case.params, V, U, nut, dp, bt, bto, 1, farm, uo, 1e-3, rhs, utmp, 0)
    torch.cuda.synchronize()
    e0.record()
    n_rk = max(1, args.steps // 4)
    for _ in range(n_rk):
        ctx.UeqnRK4(mesh, case.params, V, U, nut, dp, bt, bto, 1, farm, uo, 1e-3, rhs, utmp, 0)
    e1.record()
    torch.cuda.synchronize()
    rk_ms = e0.elapsed_time(e1) / n_rk
    print(f"# ueqn_rk4: {rk_ms:.3f} ms per predictor step ({rk_ms / ms:.2f} x ExplicitRhsU)", file=sys.stderr,
          flush=True)
    # the IMEX Runge-Kutta predictor (UeqnRK3CN_W) with the device GMRES at the reference's
    # kspIMEX settings (restart 30, atol 1e-5), dt chosen for a CFL number near 0.5 on the shard
    dt_imex = 0.25
    ctx.UeqnRK3CN(mesh, case.params, 0, V, U, nut, dp, bt, bto, 1, farm, uo, dt_imex, rhs)
    torch.cuda.synchronize()
    e0.record()
    its, res, reason = ctx.UeqnRK3CN(mesh, case.params, 0, V, U, nut, dp, bt, bto, 1, farm, uo, dt_imex, rhs)
    e1.record()
    torch.cuda.synchronize()
    imex_ms = e0.elapsed_time(e1)
    print(f"# ueqn_rk3cn (RK3WCN): {imex_ms:.1f} ms per predictor step, last stage GMRES {its} iterations, "
          f"residual {res:.2e}", file=sys.stderr, flush=True)
    # the Crank-Nicolson residual the matrix-free Newton solve evaluates (UeqnSNES's constant
    # right-hand side once per step, SNESFuncEval per Newton and finite-difference evaluation)
    bU, F = torch.empty(gl, dtype=torch.float64, device=dev), torch.empty(gl, dtype=torch.float64, device=dev)
    ctx.CNRhs(case.params, dp, bt, bto, 1, rhs, farm, bU)
    u = uo + 1e-3 * rhs

    def cn_time(fn):
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()
        e0.record()
        for _ in range(args.steps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / args.steps

    cn_ms = cn_time(lambda: ctx.CNRhs(case.params, dp, bt, bto, 1, rhs, farm, bU))
    fe_ms = cn_time(lambda: ctx.SNESFuncEval(mesh, case.params, u, V, U, nut, bU, uo, 1e-3, 1, F))
    print(f"# cn_rhs: {cn_ms:.3f} ms; snes_func_eval: {fe_ms:.3f} ms ({fe_ms / ms:.2f} x ExplicitRhsU)", file=sys.stderr,
          flush=True)
    del dp, bt, bto, farm, uo, V, U, rhs, utmp, bU, F, u
    return {"ms": ms, "value": value, "unit": "cell-updates/s", "algorithmic_bytes_per_cell": EXPLICIT_BYTES,
            "achieved": gbs, "peak": HBM_PEAK_GBS, "frac": gbs / HBM_PEAK_GBS,
            "kernels_ms": {k: v[0] / v[1] for k, v in tim.items()},
            "terms": "-dP + 3/2 bTheta - 1/2 bTheta_o + farm + FormU(formMode 0), synthetic sources",
            "ueqn_rk4_ms": rk_ms,
            "ueqn_rk3cn_w": {"ms": imex_ms, "dt": dt_imex, "last_stage_gmres_iterations": its, "residual": res,
                             "reason": reason},
            "cn_rhs_ms": cn_ms, "snes_func_eval_ms": fe_ms}


def boundary_leg(torch, ctx, mesh, case, ucont, ucat, theta, lay, args):
    """The ghost fills before the path on the shard (periodic in x and y, walls at the bottom and
    top): UpdateContravariantBCs, UpdateCartesianBCs with the bottom patch as noSlip, as a Shumann
    and as a Cabot wall function (smooth: the integration table is built in the warm-up; rough),
    and UpdateTemperatureBCs.  They touch the boundary planes only, so they are latency-bound; the
    times are what they add to a step."""
    from tosca_amd import case as C
    dev = "cuda"
    V, U, T = ucont.clone(), ucat.clone(), theta.clone()
    ustar = torch.zeros(lay.gshape, dtype=torch.float64, device=dev)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)

    def timed(fn):
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()
        e0.record()
        for _ in range(args.steps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / args.steps

    p = case.params
    bc = list(p["bc_u"])
    out = {"contravariant_ms": timed(lambda: ctx.UpdateContravariantBCs(mesh, p, V, U))}
    out["cartesian_noslip_slip_ms"] = timed(lambda: ctx.UpdateCartesianBCs(mesh, p, U))
    wf = tuple(C.BCU_WALLFUNCTION if q == 2 else bc[q] for q in range(6))
    for name, extra in (("shumann", dict(wall_u=(0, 0, -3, 0))),
                        ("cabot_smooth", dict(wall_u=(0, 0, -1, 0), wall_rough=(0.0,) * 4)),
                        ("cabot_rough", dict(wall_u=(0, 0, -1, 0), wall_rough=(0.0, 0.0, 0.1, 0.0)))):
        pw = dict(p, bc_u=wf, **extra)
        out[f"cartesian_{name}_ms"] = timed(lambda: ctx.UpdateCartesianBCs(mesh, pw, U, None, ustar))
    bct = [C.BCT_PERIODIC, C.BCT_PERIODIC, C.BCT_ZEROGRADIENT, C.BCT_FIXEDGRADIENT, C.BCT_PERIODIC, C.BCT_PERIODIC]
    tval = [0.0, 0.0, 0.0, 0.003, 0.0, 0.0]
    out["temperature_ms"] = timed(lambda: ctx.UpdateTemperatureBCs(mesh, p, bct, tval, T))
    out["ghost_points"] = {"i": 2 * (lay.my - 2) * (lay.mz - 2), "j": 2 * (lay.mx - 2) * (lay.mz - 2),
                           "k": 2 * (lay.mx - 2) * (lay.my - 2)}
    print("# ghost fills: " + ", ".join(f"{k[:-3]} {v:.3f} ms" for k, v in out.items() if k.endswith("_ms")),
          file=sys.stderr, flush=True)
    del V, U, T, ustar
    return out


# FormT, per interior cell: theta 8, ucont 24, nine face vectors 216, three face Jacobians 24,
# aj 8, the closure's diffusivity (k_t or nu_t) 8, rhs written 8
FORMT_BYTES = 8 + 24 + 216 + 24 + 8 + 8 + 8   # = 296
GRADP_BYTES = 8 + 216 + 24 + 24                # = 272


def formt_leg(torch, ctx, mesh, case, ucont, theta, nut, lay, args):
    """FormT (tosca_formt) on the step's fields, with the two closures it distinguishes under LES:
    the subgrid diffusivity k_t (here synthetic, 3 nu_t) and nu_t over the stability-corrected
    turbulent Prandtl number."""
    dev = "cuda"
    kt = (3.0 * nut).contiguous()
    rhs = torch.empty((lay.mz, lay.my, lay.mx), dtype=torch.float64, device=dev)
    tp = dict(pr=0.7, wall_t=(0, 0, 0, 0), q_wall_i=None, q_wall_j=None)
    out = {}
    for name, model in (("k_t", 2), ("prandtl", 0)):
        params = dict(case.params, les_model=model)

        def run():
            ctx.FormT(mesh, params, tp, ucont, theta, nut, kt, rhs, 1.0, 0)

        for _ in range(args.warmup):
            run()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(args.steps):
            run()
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / args.steps
        ctx.enable_timing(True)
        ctx.reset_timing()
        for _ in range(args.steps):
            run()
        tim = ctx.timings()
        ctx.enable_timing(False)
        value = lay.ncells / (ms * 1e-3)
        gbs = FORMT_BYTES * value / 1e9
        print(f"# formt ({name}): {ms:.3f} ms, {value:.3e} cells/s, {gbs:.0f} GB/s algorithmic", file=sys.stderr,
              flush=True)
        out[name] = {"ms": ms, "value": value, "unit": "cell-updates/s", "achieved": gbs, "peak": HBM_PEAK_GBS,
                     "frac": gbs / HBM_PEAK_GBS, "kernels_ms": {k: v[0] / v[1] for k, v in tim.items()}}
    out["algorithmic_bytes_per_cell"] = FORMT_BYTES
    # GradP on the same mesh, theta standing in for the pressure: p 8, nine face vectors 216, three
    # face Jacobians 24, dP written 24 B/cell
    dp = torch.empty((lay.mz, lay.my, lay.mx, 3), dtype=torch.float64, device=dev)
    for _ in range(args.warmup):
        ctx.GradP(mesh, case.params, theta, dp)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(args.steps):
        ctx.GradP(mesh, case.params, theta, dp)
    e1.record()
    torch.cuda.synchronize()
    gp_ms = e0.elapsed_time(e1) / args.steps
    gp_gbs = GRADP_BYTES * lay.ncells / (gp_ms * 1e-3) / 1e9
    print(f"# grad_p: {gp_ms:.3f} ms, {gp_gbs:.0f} GB/s algorithmic", file=sys.stderr, flush=True)
    out["grad_p"] = {"ms": gp_ms, "algorithmic_bytes_per_cell": GRADP_BYTES, "achieved": gp_gbs,
                     "frac": gp_gbs / HBM_PEAK_GBS}
    del kt, rhs, dp
    return out


def host_leg(torch, ctx, mesh, case, ucont, ucat, theta, lay, args, step_ms):
    """The step with host-resident flow fields (tosca_rhs_les_step_host): pinned host copies of
    ucont, ucat, theta, rhs, cs, nut and bTheta cross the host link around the device step.  This is
    the PCIe-inclusive rate; the headline value keeps every input resident in HBM."""
    gl = (lay.mz, lay.my, lay.mx, 3)
    pin = lambda t: t.cpu().pin_memory()
    hV, hU, hT = pin(ucont), pin(ucat), pin(theta)
    hR = torch.zeros(gl, dtype=torch.float64).pin_memory()
    hB = torch.zeros(gl, dtype=torch.float64).pin_memory()
    hC = torch.zeros(lay.gshape, dtype=torch.float64).pin_memory()
    hN = torch.zeros(lay.gshape, dtype=torch.float64).pin_memory()
    ctx.rhs_les_step_host(mesh, case.params, hV, hU, hT, hR, 1.0, hC, hN, hB, 1.0)
    n = 3
    cp = st = 0.0
    t0 = time.perf_counter()
    for _ in range(n):
        c, s_ = ctx.rhs_les_step_host(mesh, case.params, hV, hU, hT, hR, 1.0, hC, hN, hB, 1.0)
        cp += c
        st += s_
    wall = (time.perf_counter() - t0) / n * 1e3
    nbytes = sum(t.numel() * 8 for t in (hV, hU, hT, hR, hC, hN)) + sum(t.numel() * 8 for t in (hR, hC, hN, hB))
    gbs = nbytes / (cp / n * 1e-3) / 1e9
    print(f"# host-staged step: {wall:.1f} ms per step ({cp / n:.1f} ms copies at {gbs:.0f} GB/s, {st / n:.2f} ms step),"
          f" {lay.ncells / (wall * 1e-3):.3e} cells/s", file=sys.stderr, flush=True)
    del hV, hU, hT, hR, hB, hC, hN
    return {"ms": wall, "value": lay.ncells / (wall * 1e-3), "unit": "cell-updates/s", "copy_ms": cp / n,
            "step_ms": st / n, "bytes_copied": nbytes, "copy_GBs": gbs, "vs_resident": step_ms / wall}


def workload_name(cfg: int, case) -> str:
    lay = case.lay
    n = f"{lay.mz - 2}x{lay.mx - 2}x{lay.my - 2}"
    names = {1: "TGV periodic box", 2: "cosine-hill curvilinear", 3: "stretched ABL box",
             4: "cosine-hill curvilinear", 5: "ABL shard (dx=dy=10 m, stretched z)"}
    return f"{names[cfg]} {n} cells (x,y,z)"


def _cpu_model() -> str:
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _native_oracle() -> str:
    """Builds the oracle with -march=native on this host (for the timing only) and selects it;
    returns a description of the build used."""
    try:
        subprocess.run(["make", "-s", "-C", str(ROOT / "oracle"), "native"], check=True, timeout=240,
                       capture_output=True)
        os.environ["TOSCA_ORACLE_LIB"] = str(ROOT / "oracle" / "liboracle_native.so")
        return "g++ -O3 -march=native -ffp-contract=off -fopenmp, built on this host"
    except Exception as e:      # no compiler or a failed build: the portable build is timed instead
        return f"g++ -O3 -march=x86-64-v3 -ffp-contract=off -fopenmp (native build failed: {type(e).__name__})"


def cpu_baseline(cfg_case_kwargs, seconds: float):
    """Times the CPU oracle (the reference-structured C++ restatement of the same passes) on the
    host: the benchmark's workload on a bounded sample of the same case builder and physics (all
    of this process's CPU share, OpenMP over k slabs, and one thread), plus one step each of the
    64^3 Taylor-Green box and the 256x128x96 hill at their full sizes and of the stretched ABL box
    at an eighth of its size."""
    import numpy as np
    build = _native_oracle()
    from tosca_amd import case as C
    from oracle import oracle as O

    ncpu = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    # the host's share for this GPU: OMP_NUM_THREADS where the environment sets it
    ncores = max(1, min(ncpu, int(os.environ.get("OMP_NUM_THREADS", ncpu) or ncpu)))

    def prepare(builder, kw):
        case = builder(device="cpu", **kw)
        lay = case.lay
        mesh = {k: v.numpy().copy() for k, v in case.mesh.items() if k != "cent"}
        f = (case.ucont.numpy().copy(), case.ucat.numpy().copy(), case.theta.numpy().copy())
        out = dict(rhs=np.zeros((lay.mz, lay.my, lay.mx, 3)), cs=np.zeros(lay.gshape), nut=np.zeros(lay.gshape),
                   btheta=np.zeros((lay.mz, lay.my, lay.mx, 3)))
        return case, lay, mesh, f, out

    def run(prep, threads, budget, max_reps=50):
        case, lay, mesh, f, out = prep
        O.set_threads(threads)
        step = lambda: O.rhs_les_step(lay.info(), mesh, case.params, *f, out["rhs"], 1.0, out["cs"], out["nut"],
                                      out["btheta"], 1.0)
        step()                                                        # warm-up
        t0, reps = time.perf_counter(), 0
        while True:
            step()
            reps += 1
            el = time.perf_counter() - t0
            if el >= budget or reps >= max_reps:
                return lay.ncells * reps / el, reps

    builder, kw = cfg_case_kwargs
    kw = dict(kw)
    if "n" in kw:
        kw.update(n=48)
    else:
        kw.update(nx=96, ny=96, nz=48)
    prep = prepare(builder, kw)
    lay = prep[1]
    v_all, r_all = run(prep, ncores, 0.6 * seconds)
    v_one, r_one = run(prep, 1, 0.4 * seconds)
    sample = (f"CPU oracle (reference-structured C++ restatement) on the same case builder and physics at "
              f"{lay.mz - 2}x{lay.mx - 2}x{lay.my - 2} = {lay.ncells} cells; {r_all} steps on {ncores} threads, "
              f"{r_one} steps on 1 thread")
    del prep
    configs = {}
    for name, cfg, extra in (("tgv 64^3", 1, {}), ("hill 256x128x96", 2, {}),
                             ("abl 192x128x64 (an eighth of 384x256x128, 200 Fourier modes)", 3,
                              dict(nx=192, ny=128, nz=64, n_modes=200))):
        spec = C.CONFIGS[cfg]
        p = prepare(spec["builder"], dict(spec["kwargs"], **extra))
        v, reps = run(p, ncores, 0.0, max_reps=1)
        configs[name] = {"value": v, "unit": "cell-updates/s", "cells": p[1].ncells, "threads": ncores}
        del p
    return {
        "value": v_all, "unit": "cell-updates/s", "cores": ncores, "kind": "port",
        "value_1core": v_one, "sample": sample, "host": {"cpu_model": _cpu_model(), "nproc": ncpu}, "build": build,
        "configs": configs,
    }


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              f"(the multi-GPU sharding is designed, not built or measured)", file=sys.stderr)
        sys.exit(2)
    import torch
    from tosca_amd import case as C
    from tosca_amd.rhs import RhsContext

    if not torch.cuda.is_available():
        print("bench.py: no HIP device", file=sys.stderr)
        sys.exit(1)
    if args.path != "default":
        os.environ["TOSCA_PATH"] = args.path
    dev = "cuda"
    t0 = time.time()
    case = build_case(args.config, dev)
    lay = case.lay
    mesh = {k: v.contiguous() for k, v in case.mesh.items() if k != "cent"}
    ucont, ucat, theta = case.ucont.contiguous(), case.ucat.contiguous(), case.theta.contiguous()
    cent = case.mesh.pop("cent").contiguous()
    gs, gl = lay.gshape, (lay.mz, lay.my, lay.mx, 3)
    rhs = torch.zeros(gl, dtype=torch.float64, device=dev)
    cs = torch.zeros(gs, dtype=torch.float64, device=dev)
    nut = torch.zeros(gs, dtype=torch.float64, device=dev)
    btheta = torch.zeros(gl, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    setup_s = time.time() - t0
    print(f"# case built in {setup_s:.1f} s: {lay} cells={lay.ncells}", file=sys.stderr, flush=True)

    ctx = RhsContext(lay.info())
    ctx.mesh_check(mesh)
    # the shard workload carries the wind farm: its face fluxes are made once, outside the timing,
    # and the step adds them (VecAXPY of sourceFarmCont) in FormU's pass
    farm, farm_info = None, None
    if args.config == 5 and not args.no_farm:
        farm, farm_info = farm_leg(torch, ctx, mesh, case, cent, lay, args)
    del cent

    def step():
        if farm is None:
            ctx.rhs_les_step(mesh, case.params, ucont, ucat, theta, rhs, 1.0, cs, nut, btheta, 1.0)
        else:
            ctx.rhs_les_step_farm(mesh, case.params, ucont, ucat, theta, farm, rhs, 1.0, cs, nut, btheta, 1.0)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()

    # timed region: exactly K steps between two synchronisations
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(args.steps):
        step()
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / args.steps
    value = lay.ncells / (ms * 1e-3)
    print(f"# {ms:.3f} ms/step, {value:.3e} cells/s", file=sys.stderr, flush=True)

    # per-kernel HIP-event timing on the launch stream (a second pass over K steps)
    ctx.enable_timing(True)
    ctx.reset_timing()
    for _ in range(args.steps):
        step()
    tim = ctx.timings()
    ctx.enable_timing(False)
    per = {k: (v[0] / v[1], v[1] // args.steps) for k, v in tim.items()}
    for k, (avg, n) in sorted(per.items(), key=lambda kv: -kv[1][0] * kv[1][1]):
        print(f"#   {k:24s} {avg:9.4f} ms x {n}/step", file=sys.stderr, flush=True)
    dom = max(per, key=lambda k: per[k][0] * per[k][1])
    dom_ms = per[dom][0]
    dom_bytes = KERNEL_BYTES.get(dom)
    if farm is not None and dom_bytes is not None and dom.startswith("k_fused_formu"):
        dom_bytes += FARM_BYTES
    step_bytes = STEP_BYTES + (FARM_BYTES if farm is not None else 0)
    achieved = dom_bytes * lay.ncells / (dom_ms * 1e-3) / 1e9 if dom_bytes else None
    traffic = None
    # HBM bytes per launch measured with rocprofv3 (FETCH_SIZE + WRITE_SIZE, calibrated on a copy)
    tfile = ROOT / "profiles" / f"traffic_{dom}_{WORKLOAD_TAG[args.config]}.txt"
    if tfile.exists():
        for ln in tfile.read_text().splitlines():
            if ln.startswith("bytes_per_launch:"):
                traffic = float(ln.split(":", 1)[1])
    copy_gbs = copy_peak_gbs(torch)
    print(f"# device copy: {copy_gbs:.0f} GB/s", file=sys.stderr, flush=True)
    roof = {"kernel": dom, "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": (achieved / HBM_PEAK_GBS) if achieved else None, "traffic": traffic,
            "algorithmic_bytes_per_cell": dom_bytes, "cells_per_launch": lay.ncells, "avg_launch_ms": dom_ms,
            "copy_peak": copy_gbs, "frac_of_copy": (achieved / copy_gbs) if achieved else None}
    step_gbs = step_bytes * value / 1e9

    line = {
        "metric": METRIC, "value": value, "unit": "cell-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": workload_name(args.config, case), "cells": lay.ncells,
                   "les_model": {0: "smagorinsky", 2: "dynamicSmagorinsky", 5: "dynamicPASD"}[case.params["les_model"]],
                   "div_scheme": {0: "central", 1: "quick"}[case.params["div_scheme"]],
                   "path": os.environ.get("TOSCA_PATH", "default"), "parallelism": "single GPU",
                   "wind_farm": farm_info},
        "roofline": roof,
        "step_roofline": {"bound": "hbm", "algorithmic_bytes_per_cell": step_bytes, "achieved": step_gbs,
                          "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": step_gbs / HBM_PEAK_GBS,
                          "frac_of_copy": step_gbs / copy_gbs},
        "kernels_ms": {k: v[0] for k, v in per.items()},
    }
    line["mixed"] = mixed_leg(torch, ctx, mesh, case, ucont, ucat, theta, lay, args, ms, farm)
    line["dpasd"] = dpasd_leg(torch, ctx, mesh, case, ucont, ucat, theta, lay, args, farm)
    line["vjp"] = vjp_leg(torch, ctx, mesh, case, ucont, ucat, theta, cs, lay, args, ms, line["mixed"]["ms_per_step"])
    line["explicit_rhs_u"] = explicit_leg(torch, ctx, mesh, case, ucont, ucat, nut, lay, args)
    line["rk3w_adjoint"] = rk3w_leg(torch, ctx, mesh, case, ucont, ucat, nut, farm, lay, args)
    line["formt"] = formt_leg(torch, ctx, mesh, case, ucont, theta, nut, lay, args)
    line["ghost_fills"] = boundary_leg(torch, ctx, mesh, case, ucont, ucat, theta, lay, args)
    line["host_staged"] = host_leg(torch, ctx, mesh, case, ucont, ucat, theta, lay, args, ms)
    if not args.no_configs:
        line["other_configs"] = configs_leg(torch, args, args.config)
    if not args.no_cpu:
        spec = C.CONFIGS[args.config]
        line["cpu_baseline"] = cpu_baseline((spec["builder"], spec["kwargs"]), args.cpu_seconds)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
