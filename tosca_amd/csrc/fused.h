// fused.h -- argument blocks and launchers of the fused gfx950 path (fused_formu.hip,
// fused_les.hip).  The fused path serves tosca_formu, tosca_update_cs (DSM and DPASD),
// tosca_update_nut and tosca_rhs_les_step; the standalone Buoyancy keeps the
// reference-structured kernel.
#pragma once

#include "kernels.h"

namespace tosca {

// T is the type of the flow fields (ucont, ucat, nut, theta): double on the FP64 path, float on
// the mixed path.  Metrics, Jacobians, the wall stresses and the outputs are always double.
// flow type T; metric storage type MT (double, or the FP32 metric copies of the mixed path)
template <class T, class MT = double>
struct FormUArgsT {
    const T *ucont, *ucat, *nut;
    const MT *csi, *eta, *zet;
    const MT *icsi, *ieta, *izet, *jcsi, *jeta, *jzet, *kcsi, *keta, *kzet;
    const MT *iaj, *jaj, *kaj;
    const double *tauI, *tauJ;
    double *rhs;
    const double *farm;    // wind-farm source added to rhs before the accumulation (nullptr: none)
    double scale;
    int mode, quick, les;
    double nu;
    int bc_u[6];
    // fused Buoyancy (btheta == nullptr: not evaluated)
    const T *theta;
    double *btheta;
    double bscale, tRef;
    // 1: the cell-centred csi, eta, zet of every interior cell equal 0.5 a + 0.5 b of the face
    // bases icsi, jeta, kzet on its two faces of that direction (how SetMeshMetrics forms them,
    // mesh.c:925-960, checked bit for bit by tosca_mesh_check); the face dot then forms them from
    // the face data it already holds instead of loading them
    int derive_cc;
};
using FormUArgs = FormUArgsT<double>;

// FormU (+ Buoyancy when A.btheta != nullptr); ws holds formu_ws_doubles(g) doubles for the
// boundary-face fluxes and periodic-seam fp.  Instantiated for double and float flow fields.
template <class T, class MT = double>
void fused_formu(const Geo &g, const FormUArgsT<T, MT> &A, double *ws, hipStream_t s);
size_t formu_ws_doubles(const Geo &g);

// Dynamic Smagorinsky: LM, MM and |S| of every interior cell into ws (3 local scalar fields),
// then the box average + clip into cs (interior) and, when nut != nullptr, nut = Cs Delta^2 |S|
// (interior).  Ghost handling is left to the caller.  T: type of ucat and nut.  m32 (FP32 flow
// fields only): FP32 copies of csi, eta, zet the filter kernel reads instead of the mesh's.
template <class T>
void fused_dsm(const Geo &g, const tosca_mesh *m, const T *ucat, double maxCs, double *cs, T *nut, double *ws,
               hipStream_t s, const tosca_mesh_f32 *m32 = nullptr);

// Plane-averaged dynamic model (DPASD): LM, MM, |S| (3-point filter) and the 5-point filter's
// QN, NN summed over every j plane into Cs(j), written to cs (interior) with, when nut != nullptr,
// nut = Cs Delta^2 |S| (interior).  ws holds dpasd_ws_doubles(g) doubles.  T: type of ucat and nut.
template <class T>
void fused_dpasd(const Geo &g, const tosca_mesh *m, const T *ucat, double maxCs, double *cs, T *nut, double *ws,
                 hipStream_t s);
size_t dpasd_ws_doubles(const Geo &g);

// nut = Cs Delta^2 |S| at interior cells from a given Cs (Smagorinsky, or any stored Cs)
template <class T>
void fused_nut(const Geo &g, const tosca_mesh *m, const T *ucat, const double *cs, T *nut, T *ksgs, hipStream_t s);

// zero the shell of an nx*ny*nz array (ncomp values per point) outside its interior cells,
// the interior being [1+lo, n-2-hi] on each axis in array coordinates
template <class T>
void zero_shell(T *a, int ncomp, int nx, int ny, int nz, int lo, int hi, hipStream_t s);

// y += x (VecAXPY with a = 1) on the points of an nx*ny*nz vector array outside [1, n-2] on every
// axis: the part of the wind-farm AXPY the fused FormU kernel does not cover
void shell_add(double *y, const double *x, int nx, int ny, int nz, hipStream_t s);

// zero every point of a local scalar array outside the interior cells
template <class T>
void zero_outside_interior(T *a, const Geo &g, hipStream_t s);

// nut ghost cells: UpdateNutBCs subset + ghost wrap
template <class T>
void nut_bcs(T *nut, const tosca_params *p, const Geo &g, hipStream_t s);

// Adjoint of R = FormU(ucont, ucat, UpdateNut(ucat, cs)) + Buoyancy(theta): accumulates the
// cotangents of the four inputs (overwritten) from rbar; ws holds >= 52 * nloc doubles.
void vjp_rhs(const Geo &g, const tosca_mesh *mesh, const tosca_params *p, const double *ucont, const double *ucat,
             const double *theta, const double *cs, const double *rbar, double scale, double buoy_scale,
             double *ucont_bar, double *ucat_bar, double *theta_bar, double *cs_bar, double *ws, hipStream_t s);

// The same adjoint with FP32 flow fields and FP32 cotangent packets (the mixed step's function):
// the forward intermediates in FP32 as the mixed step computes them, the cotangent arithmetic in
// FP64, the packets and intermediates stored in FP32, ucont_bar / ucat_bar / theta_bar FP32.
void vjp_rhs_mixed(const Geo &g, const tosca_mesh *mesh, const tosca_params *p, const float *ucont, const float *ucat,
                   const double *cs, const double *rbar, double scale, double buoy_scale, float *ucont_bar,
                   float *ucat_bar, float *theta_bar, double *cs_bar, double *ws, hipStream_t s);

// Adjoint of FormU(ucont, ucat, nu_t; scale, formMode 0) with nu_t an input: ucont_bar and ucat_bar
// overwritten, the nu_t cotangent added to nut_bar (LES active); ws holds >= 52 * nloc doubles.
void vjp_formu_nut(const Geo &g, const tosca_mesh *mesh, const tosca_params *p, const double *ucont, const double *ucat,
                   const double *nut, const double *rbar, double scale, double *ucont_bar, double *ucat_bar,
                   double *nut_bar, double *ws, hipStream_t s);

}  // namespace tosca
