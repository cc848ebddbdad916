"""Host-side mirror of the reference's momentum-RHS operator interface over the C ABI.

    ctx = RhsContext(case.lay.info())
    ctx.UpdateCs(mesh, params, ucat, cs)                          # src/les.c:165
    ctx.UpdateNut(mesh, params, ucat, cs, nut)                    # src/les.c:2119
    ctx.Buoyancy(mesh, params, theta, 1.0, btheta)                # src/sources/ueqn_sources.c:2726
    ctx.FormU(mesh, params, ucont, ucat, nut, rhs, 1.0, 0)        # src/solvers/ueqn_solvers.c:6

Arrays are torch CUDA float64 tensors in the DMDA local layout (see include/tosca_rhs.h); the
calls queue work on torch's current stream and raise ToscaError on a non-zero return code.
"""
from __future__ import annotations

import ctypes as C
from typing import Dict, Mapping, Optional, Tuple

import torch

from . import abi


def _ptr(t: Optional[torch.Tensor], dtype=torch.float64) -> Optional[int]:
    if t is None:
        return None
    if not isinstance(t, torch.Tensor):
        raise TypeError("expected a torch tensor")
    if t.dtype != dtype or not t.is_cuda or not t.is_contiguous():
        raise ValueError(f"arrays must be contiguous {dtype} CUDA tensors")
    return t.data_ptr()


def _p32(t: Optional[torch.Tensor]) -> Optional[int]:
    return _ptr(t, torch.float32)


def mesh_struct(mesh: Mapping[str, torch.Tensor]) -> abi.Mesh:
    return abi.make_mesh({n: _ptr(mesh[n]) for n in abi.MESH_FIELDS if n in mesh})


_CHECK_KEY = ("nvert", "meshtag", "aj", "csi", "eta", "zet")


class RhsContext:
    """One tosca_ctx (workspace + stream) for one DMDA layout on one GPU."""

    def __init__(self, info: Mapping[str, int], device: int = 0, lib=None):
        self.lib = lib or abi.load()
        self.info = abi.make_info(info)
        h = C.c_void_p()
        rc = self.lib.tosca_ctx_create(C.byref(h), C.byref(self.info), int(device))
        if rc != abi.TOSCA_OK:
            raise abi.ToscaError(rc, self.lib.tosca_last_error(None).decode())
        self.h = h
        self.device = device
        self._checked = set()

    def close(self):
        if getattr(self, "h", None):
            self.lib.tosca_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int):
        if rc != abi.TOSCA_OK:
            raise abi.ToscaError(rc, self.lib.tosca_last_error(self.h).decode())

    def _mesh(self, mesh) -> abi.Mesh:
        """The mesh struct; a mesh not yet seen by this context is validated first, once, as
        SetMeshMetrics validates it once in the reference (src/mesh.c:1306-1346)."""
        key = tuple(_ptr(mesh[n]) if n in mesh else None for n in _CHECK_KEY)
        if key not in self._checked and None not in key:
            self.mesh_check(mesh)
        return mesh_struct(mesh)

    def _stream(self):
        self._check(self.lib.tosca_ctx_set_stream(self.h, C.c_void_p(torch.cuda.current_stream().cuda_stream)))

    def synchronize(self):
        self._check(self.lib.tosca_ctx_synchronize(self.h))

    # ---- the four reference operators ---------------------------------------------------------
    def FormU(self, mesh, params, ucont, ucat, nut, rhs, scale: float = 1.0, form_mode: int = 0):
        self._stream()
        m, p = self._mesh(mesh), abi.make_params(params)
        self._check(self.lib.tosca_formu(self.h, C.byref(m), C.byref(p), _ptr(ucont), _ptr(ucat), _ptr(nut),
                                         _ptr(rhs), float(scale), int(form_mode)))

    def FormU_noalloc(self, mesh, params, ucont, ucat, nut, rhs, scale: float = 1.0, form_mode: int = 0):
        """FormU through tosca_formu_noalloc: the same result as FormU, with the scratch in the
        library's fixed per-device array instead of the context's workspace (no allocation)."""
        self._stream()
        m, p = self._mesh(mesh), abi.make_params(params)
        self._check(self.lib.tosca_formu_noalloc(self.h, C.byref(m), C.byref(p), _ptr(ucont), _ptr(ucat), _ptr(nut),
                                                 _ptr(rhs), float(scale), int(form_mode)))

    def UpdateCs(self, mesh, params, ucat, cs):
        self._stream()
        m, p = self._mesh(mesh), abi.make_params(params)
        self._check(self.lib.tosca_update_cs(self.h, C.byref(m), C.byref(p), _ptr(ucat), _ptr(cs)))

    def UpdateNut(self, mesh, params, ucat, cs, nut, ksgs=None):
        self._stream()
        m, p = self._mesh(mesh), abi.make_params(params)
        self._check(self.lib.tosca_update_nut(self.h, C.byref(m), C.byref(p), _ptr(ucat), _ptr(cs), _ptr(nut),
                                              _ptr(ksgs)))

    def Buoyancy(self, mesh, params, theta, scale, btheta, gcont=None):
        self._stream()
        m, p = self._mesh(mesh), abi.make_params(params)
        self._check(self.lib.tosca_buoyancy(self.h, C.byref(m), C.byref(p), _ptr(theta), float(scale),
                                            _ptr(btheta), _ptr(gcont)))

    def rhs_les_step(self, mesh, params, ucont, ucat, theta, rhs, scale, cs, nut, btheta, buoy_scale=1.0):
        self._stream()
        m, p = self._mesh(mesh), abi.make_params(params)
        self._check(self.lib.tosca_rhs_les_step(self.h, C.byref(m), C.byref(p), _ptr(ucont), _ptr(ucat), _ptr(theta),
                                                _ptr(rhs), float(scale), _ptr(cs), _ptr(nut), _ptr(btheta),
                                                float(buoy_scale)))

    def rhs_les_step_mixed(self, mesh, params, ucont, ucat, theta, rhs, scale, cs, nut, btheta, buoy_scale=1.0):
        """Mixed-precision step: ucont, ucat, theta, nut float32; mesh, rhs, cs, btheta float64."""
        self._stream()
        m, p = self._mesh(mesh), abi.make_params(params)
        self._check(self.lib.tosca_rhs_les_step_mixed(self.h, C.byref(m), C.byref(p), _p32(ucont), _p32(ucat),
                                                      _p32(theta), _ptr(rhs), float(scale), _ptr(cs), _p32(nut),
                                                      _ptr(btheta), float(buoy_scale)))

    def rhs_les_step_mixed_m32(self, mesh, mesh32, params, ucont, ucat, theta, rhs, scale, cs, nut, btheta,
                               buoy_scale=1.0):
        """Mixed-precision step whose FormU reads FP32 metric copies (mesh32: float32 tensors of the
        metrics in abi.MESH_F32_FIELDS); the LES part reads the float64 mesh."""
        self._stream()
        m, p = self._mesh(mesh), abi.make_params(params)
        m32 = abi.make_mesh_f32({n: _p32(mesh32[n]) for n in abi.MESH_F32_FIELDS})
        self._check(self.lib.tosca_rhs_les_step_mixed_m32(self.h, C.byref(m), C.byref(m32), C.byref(p), _p32(ucont),
                                                          _p32(ucat), _p32(theta), _ptr(rhs), float(scale), _ptr(cs),
                                                          _p32(nut), _ptr(btheta), float(buoy_scale)))

    def rhs_vjp(self, mesh, params, ucont, ucat, theta, cs, rbar, scale, buoy_scale, ucont_bar, ucat_bar, theta_bar,
                cs_bar):
        """J^T rbar of R = FormU(ucont, ucat, UpdateNut(ucat, cs)) + Buoyancy(theta) into the four
        *_bar arrays (overwritten)."""
        self._stream()
        m, p = self._mesh(mesh), abi.make_params(params)
        self._check(self.lib.tosca_rhs_vjp(self.h, C.byref(m), C.byref(p), _ptr(ucont), _ptr(ucat), _ptr(theta),
                                           _ptr(cs), _ptr(rbar), float(scale), float(buoy_scale), _ptr(ucont_bar),
                                           _ptr(ucat_bar), _ptr(theta_bar), _ptr(cs_bar)))

    def rhs_vjp_mixed(self, mesh, params, ucont, ucat, theta, cs, rbar, scale, buoy_scale, ucont_bar, ucat_bar,
                      theta_bar, cs_bar):
        """rhs_vjp for the mixed step's FP32 flow fields: ucont, ucat, theta and ucont_bar, ucat_bar,
        theta_bar FP32; cs, rbar and cs_bar FP64."""
        self._stream()
        m, p = self._mesh(mesh), abi.make_params(params)
        self._check(self.lib.tosca_rhs_vjp_mixed(self.h, C.byref(m), C.byref(p), _p32(ucont), _p32(ucat), _p32(theta),
                                                 _ptr(cs), _ptr(rbar), float(scale), float(buoy_scale),
                                                 _p32(ucont_bar), _p32(ucat_bar), _p32(theta_bar), _ptr(cs_bar)))

    def ExplicitRhsU(self, mesh, params, ucont, ucat, nut, dp, btheta, btheta_o, ab2, farm, ucont_o, rhs,
                     form_mode: int = 0):
        """ExplicitRhsU(ueqn, formMode) (src/solvers/ueqn_solvers.c:1276-1407): no-penetration and
        periodic resets of ucont (in place), ucat from ucont (in place), then rhs (overwritten) =
        -dp + bTheta terms + farm + FormU, with the non-resolved faces reset.  btheta / farm None:
        the term is off; ab2: the second-order bTheta extrapolation (clock->it > itStart)."""
        self._stream()
        m, p = self._mesh(mesh), abi.make_params(params)
        self._check(self.lib.tosca_explicit_rhs_u(self.h, C.byref(m), C.byref(p), _ptr(ucont), _ptr(ucat), _ptr(nut),
                                                  _ptr(dp), _ptr(btheta), _ptr(btheta_o), int(ab2), _ptr(farm),
                                                  _ptr(ucont_o), _ptr(rhs), int(form_mode)))

    # ---- boundary-condition ghost fills (src/boundary.c) ---------------------------------------
    def UpdateContravariantBCs(self, mesh, params, ucont, ucat):
        """UpdateContravariantBCs(ueqn): boundary-face fluxes of ucont from ucat, in place."""
        self._stream()
        m, p = self._mesh(mesh), abi.make_params(params)
        self._check(self.lib.tosca_update_contravariant_bcs(self.h, C.byref(m), C.byref(p), _ptr(ucont), _ptr(ucat)))

    def UpdateCartesianBCs(self, mesh, params, ucat, u_fixed=None, ustar=None, inlet=None):
        """UpdateCartesianBCs(ueqn): ghost cells of ucat, in place; u_fixed: 6 x 3 fixedValue vectors;
        ustar: local scalar array (lUstar) the Cabot wall-function patches write; inlet: dict of the
        kLeft inlet-function profile (abi.make_inlet keys; its arrays CUDA tensors: cent float64,
        plane / weights* / y_weights float64, noise / cells* / y_ids int32)."""
        self._stream()
        m, p = self._mesh(mesh), abi.make_params(params)
        uf = None
        if u_fixed is not None:
            uf = (C.c_double * 18)(*[float(x) for row in u_fixed for x in row])
        il = None
        if inlet is not None:
            ints = ("noise", "cells", "cells_1", "cells_2", "y_ids")
            il = abi.make_inlet(dict(inlet, **{k: _ptr(inlet.get(k), torch.int32 if k in ints else torch.float64)
                                               for k in abi._INLET_PTRS}))
        self._check(self.lib.tosca_update_cartesian_bcs(self.h, C.byref(m), C.byref(p), _ptr(ucat),
                                                        C.cast(uf, C.c_void_p) if uf is not None else None,
                                                        _ptr(ustar), C.cast(C.pointer(il), C.c_void_p) if il is not None else None))

    def UpdateTemperatureBCs(self, mesh, params, bc_t, t_value, theta):
        """UpdateTemperatureBCs(teqn): ghost cells of theta, in place; bc_t: 6 TOSCA_BCT_* kinds."""
        self._stream()
        m, p = self._mesh(mesh), abi.make_params(params)
        kinds = (C.c_int * 6)(*[int(b) for b in bc_t])
        vals = (C.c_double * 6)(*[float(v) for v in t_value])
        self._check(self.lib.tosca_update_temperature_bcs(self.h, C.byref(m), C.byref(p), C.cast(kinds, C.c_void_p),
                                                          C.cast(vals, C.c_void_p), _ptr(theta)))

    # ---- instrumentation -----------------------------------------------------------------------
    def GradP(self, mesh, params, p, dp):
        """GradP (src/peqn.c:3437-3773): dp (global vector) from the local pressure p."""
        m, pa = self._mesh(mesh), abi.make_params(params)
        self._check(self.lib.tosca_grad_p(self.h, C.byref(m), C.byref(pa), _ptr(p), _ptr(dp)))

    def UeqnFE(self, mesh, params, ucont, ucat, nut, dp, btheta, btheta_o, ab2, farm, ucont_o, dt, rhs,
               form_mode: int = 0):
        """UeqnFE (src/solvers/ueqn_solvers.c:1411-1441): ucont <- ucont_o + dt ExplicitRhsU."""
        m, p = self._mesh(mesh), abi.make_params(params)
        self._check(self.lib.tosca_ueqn_fe(self.h, C.byref(m), C.byref(p), _ptr(ucont), _ptr(ucat), _ptr(nut),
                                           _ptr(dp), _ptr(btheta), _ptr(btheta_o), int(ab2), _ptr(farm),
                                           _ptr(ucont_o), float(dt), _ptr(rhs), int(form_mode)))

    def UeqnRK4(self, mesh, params, ucont, ucat, nut, dp, btheta, btheta_o, ab2, farm, ucont_o, dt, rhs, utmp,
                form_mode: int = 0):
        """UeqnRK4 (src/solvers/ueqn_solvers.c:1446-1504): four ExplicitRhsU stages."""
        m, p = self._mesh(mesh), abi.make_params(params)
        self._check(self.lib.tosca_ueqn_rk4(self.h, C.byref(m), C.byref(p), _ptr(ucont), _ptr(ucat), _ptr(nut),
                                            _ptr(dp), _ptr(btheta), _ptr(btheta_o), int(ab2), _ptr(farm),
                                            _ptr(ucont_o), float(dt), _ptr(rhs), _ptr(utmp), int(form_mode)))

    def UeqnRK3W(self, mesh, params, ucont, ucat, nut, dp, btheta, btheta_o, ab2, farm, dt, n_stages, ucont_o, rhs,
                 rhs_o, form_mode: int = 0):
        """The explicit Wray Runge-Kutta predictor (UeqnRK3CN_W's scheme, src/solvers/ueqn_solvers.c:
        1508-1583, with the whole right-hand side explicit): n_stages ExplicitRhsU stages, a time
        step every third; ucont_o, rhs, rhs_o are global work vectors."""
        self._stream()
        m, p = self._mesh(mesh), abi.make_params(params)
        self._check(self.lib.tosca_ueqn_rk3w(self.h, C.byref(m), C.byref(p), _ptr(ucont), _ptr(ucat), _ptr(nut),
                                             _ptr(dp), _ptr(btheta), _ptr(btheta_o), int(ab2), _ptr(farm), float(dt),
                                             int(n_stages), _ptr(ucont_o), _ptr(rhs), _ptr(rhs_o), int(form_mode)))

    def UeqnRK3W_adjoint(self, mesh, params, ucont, ucat, nut, dp, btheta, btheta_o, ab2, farm, dt, n_stages, ucont_o,
                         rhs, rhs_o, ucont_bar, nut_bar):
        """Checkpointed adjoint of UeqnRK3W (form mode 0): the forward sweep as UeqnRK3W, then
        ucont_bar (global) goes from the cotangent of the final Ucont to that of the initial one,
        and nut_bar (local) gets the cotangent of nu_t."""
        self._stream()
        m, p = self._mesh(mesh), abi.make_params(params)
        self._check(self.lib.tosca_ueqn_rk3w_adjoint(self.h, C.byref(m), C.byref(p), _ptr(ucont), _ptr(ucat),
                                                     _ptr(nut), _ptr(dp), _ptr(btheta), _ptr(btheta_o), int(ab2),
                                                     _ptr(farm), float(dt), int(n_stages), _ptr(ucont_o), _ptr(rhs),
                                                     _ptr(rhs_o), _ptr(ucont_bar), _ptr(nut_bar)))

    def IMEXMatVec(self, mesh, params, v, ucont, ucat, nut, ucont_o, coef, av):
        """IMEXMatVec (src/solvers/ueqn_solvers.c:1739-1801): av = v - coef Visc(v)."""
        m, p = self._mesh(mesh), abi.make_params(params)
        self._check(self.lib.tosca_imex_matvec(self.h, C.byref(m), C.byref(p), _ptr(v), _ptr(ucont), _ptr(ucat),
                                               _ptr(nut), _ptr(ucont_o), float(coef), _ptr(av)))

    def IMEXSolve(self, mesh, params, b, x, ucont, ucat, nut, ucont_o, coef, **krylov):
        """KSPSolve(kspIMEX, b, x) with IMEXMatVec: restarted GMRES on the device; x is updated in
        place.  krylov: restart, max_it, rtol, atol (the reference's defaults otherwise).
        Returns (iterations, residual, reason)."""
        m, p = self._mesh(mesh), abi.make_params(params)
        kp, kr = abi.make_krylov_params(**krylov), abi.KrylovResult()
        self._check(self.lib.tosca_imex_solve(self.h, C.byref(m), C.byref(p), _ptr(b), _ptr(x), _ptr(ucont), _ptr(ucat),
                                              _ptr(nut), _ptr(ucont_o), float(coef), C.byref(kp), C.byref(kr)))
        return kr.iterations, kr.residual, kr.reason

    def UeqnRK3CN(self, mesh, params, scheme, ucont, ucat, nut, dp, btheta, btheta_o, ab2, farm, ucont_o, dt, rhs,
                  **krylov):
        """UeqnRK3CN_W (scheme 0) / UeqnRK3CN_SO (scheme 1, src/solvers/ueqn_solvers.c:1508-1661);
        returns the last stage's (iterations, residual, reason)."""
        m, p = self._mesh(mesh), abi.make_params(params)
        kp, kr = abi.make_krylov_params(**krylov), abi.KrylovResult()
        self._check(self.lib.tosca_ueqn_rk3cn(self.h, C.byref(m), C.byref(p), int(scheme), _ptr(ucont), _ptr(ucat),
                                              _ptr(nut), _ptr(dp), _ptr(btheta), _ptr(btheta_o), int(ab2), _ptr(farm),
                                              _ptr(ucont_o), float(dt), _ptr(rhs), C.byref(kp), C.byref(kr)))
        return kr.iterations, kr.residual, kr.reason

    def UeqnABCN(self, mesh, params, ucont, ucat, nut, dp, btheta, btheta_o, ab2, farm, ucont_o, dt, rhs, conv_o,
                 **krylov):
        """UeqnABCN (src/solvers/ueqn_solvers.c:1665-1737); conv_o carries RhsConv_o between steps.
        Returns (iterations, residual, reason) of the solve."""
        m, p = self._mesh(mesh), abi.make_params(params)
        kp, kr = abi.make_krylov_params(**krylov), abi.KrylovResult()
        self._check(self.lib.tosca_ueqn_abcn(self.h, C.byref(m), C.byref(p), _ptr(ucont), _ptr(ucat), _ptr(nut),
                                             _ptr(dp), _ptr(btheta), _ptr(btheta_o), int(ab2), _ptr(farm),
                                             _ptr(ucont_o), float(dt), _ptr(rhs), _ptr(conv_o), C.byref(kp),
                                             C.byref(kr)))
        return kr.iterations, kr.residual, kr.reason

    def rhs_les_step_host(self, mesh, params, ucont, ucat, theta, rhs, scale, cs, nut, btheta, buoy_scale=1.0):
        """The step with host-resident flow fields (contiguous float64 CPU tensors, preferably pinned);
        the mesh stays on the device.  Returns (copy ms, step ms)."""
        def hp(t):
            if t.dtype != torch.float64 or t.is_cuda or not t.is_contiguous():
                raise ValueError("host arrays must be contiguous float64 CPU tensors")
            return t.data_ptr()
        m, p = self._mesh(mesh), abi.make_params(params)
        times = (C.c_double * 2)()
        self._check(self.lib.tosca_rhs_les_step_host(self.h, C.byref(m), C.byref(p), hp(ucont), hp(ucat), hp(theta),
                                                     hp(rhs), float(scale), hp(cs), hp(nut), hp(btheta),
                                                     float(buoy_scale), C.cast(times, C.c_void_p)))
        return times[0], times[1]

    def CNRhs(self, params, dp, btheta, btheta_o, ab2, rhs_o, farm, bU):
        """UeqnSNES's constant right-hand side bU (src/solvers/ueqn_solvers.c:1099-1156)."""
        p = abi.make_params(params)
        self._check(self.lib.tosca_cn_rhs(self.h, C.byref(p), _ptr(dp), _ptr(btheta), _ptr(btheta_o), int(ab2),
                                          _ptr(rhs_o), _ptr(farm), _ptr(bU)))

    def SNESFuncEval(self, mesh, params, u, ucont, ucat, nut, bU, ucont_o, dt, ab2, F):
        """SNESFuncEval (src/solvers/ueqn_solvers.c:1185-1272): F = the Crank-Nicolson residual at u."""
        m, p = self._mesh(mesh), abi.make_params(params)
        self._check(self.lib.tosca_snes_func_eval(self.h, C.byref(m), C.byref(p), _ptr(u), _ptr(ucont), _ptr(ucat),
                                                  _ptr(nut), _ptr(bU), _ptr(ucont_o), float(dt), int(ab2), _ptr(F)))

    def FormT(self, mesh, params, tparams, ucont, theta, nut, kt, rhs, scale: float = 1.0, form_mode: int = 0):
        """FormT (src/solvers/teqn_solvers.c:6-592): rhs (global scalar) at interior cells.
        tparams: dict(pr=..., wall_t=(iLeft, iRight, jLeft, jRight) thetaWallFunction flags,
        q_wall_i=tensor (3, mz, my) or None, q_wall_j=tensor (3, mz, mx) or None)."""
        m, p = self._mesh(mesh), abi.make_params(params)
        tp = abi.make_teqn_params(dict(tparams, q_wall_i=_ptr(tparams.get("q_wall_i")),
                                       q_wall_j=_ptr(tparams.get("q_wall_j"))))
        self._check(self.lib.tosca_formt(self.h, C.byref(m), C.byref(p), C.byref(tp), _ptr(ucont), _ptr(theta),
                                         _ptr(nut), _ptr(kt), _ptr(rhs), float(scale), int(form_mode)))

    # ---- wind-farm body force (src/turbines/turbines_force.c) ---------------------------------
    def ProjectUADM(self, params, cent, farm: "DeviceFarm", scat):
        """projectBladeForce, uniformADM branch (turbines_force.c:717-932): scat (local AoS) from the
        rotors of farm, cent = mesh->lCent (local AoS)."""
        self._stream()
        p = abi.make_params(params)
        self._check(self.lib.tosca_project_uadm(self.h, C.byref(p), _ptr(cent), C.byref(farm.struct), _ptr(scat)))

    def BodyForceToCont(self, mesh, params, scat, scont):
        """bodyForceCartesian2Contravariant (turbines_force.c:1741-1835): scont (global AoS)."""
        self._stream()
        m, p = self._mesh(mesh), abi.make_params(params)
        self._check(self.lib.tosca_body_force_to_cont(self.h, C.byref(m), C.byref(p), _ptr(scat), _ptr(scont)))

    def rhs_les_step_farm(self, mesh, params, ucont, ucat, theta, farm, rhs, scale, cs, nut, btheta,
                          buoy_scale=1.0):
        """rhs += farm (sourceFarmCont, global AoS), then rhs_les_step, in one pass."""
        self._stream()
        m, p = self._mesh(mesh), abi.make_params(params)
        self._check(self.lib.tosca_rhs_les_step_farm(self.h, C.byref(m), C.byref(p), _ptr(ucont), _ptr(ucat),
                                                     _ptr(theta), _ptr(farm), _ptr(rhs), float(scale), _ptr(cs),
                                                     _ptr(nut), _ptr(btheta), float(buoy_scale)))

    def rhs_les_step_mixed_farm(self, mesh, params, ucont, ucat, theta, farm, rhs, scale, cs, nut, btheta,
                                buoy_scale=1.0, mesh32=None):
        """rhs_les_step_mixed (or _m32 when mesh32, the float32 metric copies, is given) with the
        float64 farm source added first."""
        self._stream()
        m, p = self._mesh(mesh), abi.make_params(params)
        m32 = None
        if mesh32 is not None:
            m32 = C.byref(abi.make_mesh_f32({n: _p32(mesh32[n]) for n in abi.MESH_F32_FIELDS}))
        self._check(self.lib.tosca_rhs_les_step_mixed_farm(self.h, C.byref(m), m32, C.byref(p), _p32(ucont),
                                                           _p32(ucat), _p32(theta), _ptr(farm), _ptr(rhs), float(scale),
                                                           _ptr(cs), _p32(nut), _ptr(btheta), float(buoy_scale)))

    def enable_timing(self, on: bool = True):
        self._check(self.lib.tosca_ctx_enable_timing(self.h, int(on)))

    def reset_timing(self):
        self._check(self.lib.tosca_ctx_timing_reset(self.h))

    def timings(self) -> Dict[str, Tuple[float, int]]:
        """{kernel name: (total event milliseconds, launches)} since the last reset."""
        n = self.lib.tosca_ctx_timing_count(self.h)
        if n < 0:
            self._check(abi.TOSCA_ERR_ARG)
        out = {}
        buf = C.create_string_buffer(128)
        ms, cnt = C.c_double(), C.c_int64()
        for i in range(n):
            self._check(self.lib.tosca_ctx_timing_entry(self.h, i, buf, 128, C.byref(ms), C.byref(cnt)))
            out[buf.value.decode()] = (ms.value, cnt.value)
        return out

    def mesh_check(self, mesh):
        """tosca_mesh_check: validates the mesh (nvert, meshTag, volumes, areas) and registers it with
        the context; every operator taking this mesh requires it."""
        self._stream()
        m = mesh_struct(mesh)
        self._check(self.lib.tosca_mesh_check(self.h, C.byref(m)))
        self._checked.add(tuple(_ptr(mesh[n]) for n in _CHECK_KEY))


class DeviceFarm:
    """A tosca_amd.farm.UadmFarm on the device: float64 CUDA tensors and the tosca_uadm struct that
    points at them (kept alive together)."""

    def __init__(self, farm, device="cuda"):
        from .farm import PRJ_NSIGMA
        self.t = {k: torch.as_tensor(getattr(farm, k), dtype=torch.float64).contiguous().to(device)
                  for k in ("center", "sphere", "eps", "points", "force")}
        u = abi.Uadm()
        u.n_rotors, u.n_points, u.nsigma = int(farm.n_rotors), int(farm.n_points), float(PRJ_NSIGMA)
        for k, v in self.t.items():
            setattr(u, k, v.data_ptr() if v.numel() else None)
        self.struct = u
