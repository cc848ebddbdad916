// kernels.h -- host-side launch entry points of the gfx950 kernels (internal to the library).
#pragma once

#include "common.h"
#include "timing.h"

namespace tosca {

// ghost handling (ref_kernels.hip)
template <class T> void launch_wrap(T *a, int ncomp, const Geo &g, hipStream_t s);
void launch_reset_cell_periodic(double *a, int ncomp, const Geo &g, hipStream_t s);
void launch_reset_face_periodic(double *a, int ncomp, const Geo &g, hipStream_t s);
void launch_fill(double *a, long n, double v, hipStream_t s);
template <class T> void launch_nut_bcs(T *nut, const tosca_params *p, const Geo &g, hipStream_t s);

// reference-structured path (ref_kernels.hip); `ws` is a workspace of >= 22 * nloc doubles
void ref_formu(const Geo &g, const tosca_mesh *m, const tosca_params *p, const double *ucont, const double *ucat,
               const double *nut, double *rhs, double scale, int mode, double *ws, hipStream_t s);
void ref_update_cs(const Geo &g, const tosca_mesh *m, const tosca_params *p, const double *ucat, double *cs,
                   double *ws, hipStream_t s);
void ref_update_nut(const Geo &g, const tosca_mesh *m, const tosca_params *p, const double *ucat, const double *cs,
                    double *nut, double *ksgs, hipStream_t s);
void ref_buoyancy(const Geo &g, const tosca_mesh *m, const tosca_params *p, const double *theta, double scale,
                  double *btheta, double *gcont, hipStream_t s);

// ExplicitRhsU steps around FormU (explicit_rhs.hip)
void launch_no_penetration(double *ucont, const tosca_params *p, const Geo &g, hipStream_t s);
void launch_contra_to_cart(const double *ucont, const tosca_mesh *m, double *ucat, const Geo &g, hipStream_t s);
void launch_rhs_init(double *rhs, const double *dp, const double *bt, const double *bto, int ab2, const double *farm,
                     const Geo &g, hipStream_t s);
void launch_non_resolved(double *rhs, const double *ucont_o, const tosca_params *p, const Geo &g, hipStream_t s);
void launch_fe_update(double *rhs, const double *ucont_o, double dt, double *ucont, const Geo &g, hipStream_t s);
void launch_waxpy_local(double *ucont, double al, const double *rhs, const double *ucont_o, const Geo &g, hipStream_t s);
void launch_axpy(double *y, double a, const double *x, long n, hipStream_t s);
void launch_rk_stage(double *utmp, double a, const double *rhs, double *ucont, double al, const double *ucont_o,
                     int use_acc, const Geo &g, hipStream_t s);
void launch_snes_finish(double *F, const double *u, const double *ucont_o, double dt, const tosca_params *p,
                        const Geo &g, hipStream_t s);
void launch_cn_rhs(double *bU, const double *dp, const double *bt, const double *bto, const double *rhs_o,
                   const double *farm, double wbt, long n, hipStream_t s);
void launch_global_to_local(double *loc, const double *glob, const Geo &g, hipStream_t s);

// explicit Wray Runge-Kutta predictor and its adjoint's transposed steps (rk3w.hip)
void launch_rk3w_update(double *ucont, const double *rhs, double *rhs_o, double a, double b, const Geo &g,
                        hipStream_t s);
void launch_rk3w_adj_combine(const double *xbar, double *carry, double *rbar, double a, double b, const Geo &g,
                             hipStream_t s);
void launch_reset_cell_periodic_adj(double *a, int ncomp, const Geo &g, hipStream_t s);
void launch_c2c_adj(const double *ubar, const tosca_mesh *m, double *qbar, double *vbar, const Geo &g, hipStream_t s);
void launch_add_owned(double *xbar, const double *vbar, const Geo &g, hipStream_t s);
void launch_face_resets_adj(double *xbar, const double *x, const tosca_params *p, const Geo &g, hipStream_t s);

// Krylov vector kernels (krylov.hip)
void launch_mdot(const double *V, long stride, int nv, const double *w, long n, double *partial, double *out,
                 hipStream_t s);
size_t mdot_partial_doubles(int nvmax);
void launch_maxpy(double *w, const double *V, long stride, int nv, const double *h, double sign, long n, hipStream_t s);
void launch_scale_to(double *y, const double *x, double a, long n, hipStream_t s);
void launch_residual(double *r, const double *b, const double *t, long n, hipStream_t s);
void launch_local_to_global(double *glob, const double *loc, const Geo &g, hipStream_t s);

// boundary-condition ghost fills (bc_fill.hip)
void launch_contravariant_bcs(double *ucont, const double *ucat, const tosca_mesh *m, const tosca_params *p,
                              const Geo &g, hipStream_t s);
// the Cabot wall law's integration table: y+ in [0, 1e7] at steps of 2 (wallfunctions.c:296-360)
constexpr long CABOT_INTERVAL = 2, CABOT_MAX_YP = 10000000, CABOT_NYP = CABOT_MAX_YP / CABOT_INTERVAL;
size_t cabot_table_doubles();       // the table (CABOT_NYP + 1) and the chunk totals of its prefix sum
void launch_cabot_table(double *tab, hipStream_t s);
void launch_cartesian_bcs(double *ucat, const double *u_fixed, double *ustar, const double *cabot_tab,
                          const tosca_inlet *inlet, const tosca_mesh *m, const tosca_params *p,
                          const Geo &g, hipStream_t s);
void launch_temperature_bcs(double *t, const int *bc_t, const double *t_value, const tosca_mesh *m, const Geo &g,
                            hipStream_t s);

// GradP (gradp.hip): dp (global vector) from the local pressure p
void launch_grad_p(const double *p, const tosca_mesh *m, double *dp, const Geo &g, hipStream_t s);

// FormT (formt.hip): reference-structured face and cell kernels; ws holds 6 local doubles per point
void ref_formt(const Geo &g, const tosca_mesh *m, const tosca_params *p, const tosca_teqn_params *tp,
               const double *ucont, const double *theta, const double *nut, const double *kt, double *rhs,
               double scale, int mode, double *ws, hipStream_t s);

// wind-farm body force (farm.hip): projectBladeForce's uniformADM branch into the local scat, and
// bodyForceCartesian2Contravariant into the global scont
void launch_uadm_project(const double *cent, const tosca_uadm *F, double *scat, const Geo &g, hipStream_t s);
void launch_body_force_to_cont(const double *scat, const tosca_mesh *m, double *scont, const Geo &g, hipStream_t s);

// mesh validation (api.hip): counts of offending points, written to `out` (device, 4 ints)
void launch_mesh_check(const Geo &g, const tosca_mesh *m, int *out, hipStream_t s);

}  // namespace tosca
