// rk3w.hip -- the explicit Wray Runge-Kutta momentum predictor and the transposed steps of its
// checkpointed adjoint, on gfx950.  Compiled with FP contraction off: the forward stage update is
// the reference's VecCopy / VecAXPY sequence evaluated per element in the same order.
//
// Forward stage i (the low-storage scheme of UeqnRK3CN_W, src/solvers/ueqn_solvers.c:1508-1583,
// beta / delta of :1518-1520, the whole right-hand side explicit):
//   k_rk3w_update        Ucont = (Ucont + dt beta_i Rhs) + (-1.0 dt delta_i) Rhs_o into the owned
//                        range of the local ucont, and Rhs_o = Rhs (:1533-1541, :1563-1569)
// Adjoint (the transpose of each forward step, in reverse order; no reference counterpart):
//   k_rk3w_adj_combine   cotangent of the stage's Rhs from the two stage updates that read it
//   k_reset_axis_adj     transpose of resetCellPeriodicFluxes' boundary copies (inline.h:5496-5514)
//   k_wrap_axis_adj      transpose of the periodic ghost scatter (DMLocalToLocal)
//   k_c2c_adj_solve      transpose of contravariantToCartesian's 3x3 solve (src/ueqn.c:571-613)
//   k_c2c_adj_gather     transpose of its face averages (inline.h:4307-4310), added to ucont_bar
//   k_add_owned          the owned range of a local cotangent added to a global one
//   k_face_periodic_adj  transpose of resetFacePeriodicFluxesVector's face-0 copy (inline.h:5357-5451)
//   k_no_penetration_adj transpose of resetNoPenetrationFluxes (inline.h:5262-5355): the zeroed
//                        fluxes carry no cotangent; a zeroGradient flux only where it was not clipped
#include "kernels.h"

namespace tosca {

namespace {

struct Kinds { int bc[6]; };

__device__ __forceinline__ long owned_local(const Geo &g, long q) {
    const int i = (int)(q % g.mx);
    const long r = q / g.mx;
    const int j = (int)(r % g.my), k = (int)(r / g.my);
    return g.L(k, j, i);
}

__device__ __forceinline__ long gidx(const Geo &g, int k, int j, int i) { return ((long)k * g.my + j) * g.mx + i; }

// one thread per owned double
__global__ void k_rk3w_update(double *ucont, const double *rhs, double *rhs_o, double a, double b, Geo g, long n) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    double *u = ucont + 3 * owned_local(g, t / 3) + t % 3;
    const double r = rhs[t];
    double v = *u;
    v = v + a * r;
    v = v + b * rhs_o[t];
    *u = v;
    rhs_o[t] = r;
}

// rbar = a xbar + carry (the cotangent of this stage's Rhs: from the next state, a = dt beta_i, and
// from the following stage's Rhs_o term, carry); carry = b xbar (b = -1.0 dt delta_i: the cotangent of
// the previous stage's Rhs through this stage's Rhs_o term)
__global__ void k_rk3w_adj_combine(const double *xbar, double *carry, double *rbar, double a, double b, long n) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const double x = xbar[t];
    rbar[t] = a * x + carry[t];
    carry[t] = b * x;
}

// transpose of a copy a[dst] = a[src]: abar[src] += abar[dst]; abar[dst] = 0.  Within one launch
// the destinations are distinct, the sources are distinct and no source is a destination, so the
// threads touch disjoint points.
__device__ __forceinline__ void copy_adj(double *a, int ncomp, long cd, long cs) {
    for (int q = 0; q < ncomp; q++) {
        a[ncomp * cs + q] = a[ncomp * cs + q] + a[ncomp * cd + q];
        a[ncomp * cd + q] = 0.0;
    }
}

// transpose of k_reset_axis: boundary cells 0 / m-1 took cells m-2 / 1 over the owned range of the
// other two axes
__global__ void k_reset_axis_adj(double *a, int ncomp, Geo g, int axis) {
    const int m[3] = {g.mx, g.my, g.mz};
    const int t1 = (axis == 0) ? 1 : 0, t2 = (axis == 2) ? 1 : 2;
    const int a1 = blockIdx.x * blockDim.x + threadIdx.x, a2 = blockIdx.y, side = blockIdx.z;
    if (a1 >= m[t1] || a2 >= m[t2]) return;
    int d[3], s[3];
    d[axis] = side ? m[axis] - 1 : 0;
    s[axis] = side ? 1 : m[axis] - 2;
    d[t1] = s[t1] = a1;
    d[t2] = s[t2] = a2;
    copy_adj(a, ncomp, g.L(d[2], d[1], d[0]), g.L(s[2], s[1], s[0]));
}

// transpose of k_wrap_axis: ghost layers took the wrapped owned values, over the transverse ranges
// that include the ghosts of the other type-2 axes.  One launch per side (layer0 = 0: the ghosts
// below the axis, GW: above), so that the sources of one launch are distinct even on an axis of
// fewer than 2 GW points.
__global__ void k_wrap_axis_adj(double *a, int ncomp, Geo g, int axis, int layer0) {
    const int m[3] = {g.mx, g.my, g.mz};
    const int t1 = (axis == 0) ? 1 : 0, t2 = (axis == 2) ? 1 : 2;
    const int lo1 = (g.per[t1] == TOSCA_PERIODIC_WRAP) ? -GW : 0, n1 = m[t1] + ((g.per[t1] == TOSCA_PERIODIC_WRAP) ? 2 * GW : 0);
    const int lo2 = (g.per[t2] == TOSCA_PERIODIC_WRAP) ? -GW : 0, n2 = m[t2] + ((g.per[t2] == TOSCA_PERIODIC_WRAP) ? 2 * GW : 0);
    const int a1 = blockIdx.x * blockDim.x + threadIdx.x, a2 = blockIdx.y, sl = layer0 + blockIdx.z;
    if (a1 >= n1 || a2 >= n2) return;
    const int n = m[axis];
    const int dst = (sl < GW) ? (sl - GW) : (n + sl - GW);
    const int src = (sl < GW) ? (n + sl - GW) : (sl - GW);
    int d[3], s[3];
    d[axis] = dst; s[axis] = src;
    d[t1] = s[t1] = lo1 + a1;
    d[t2] = s[t2] = lo2 + a2;
    copy_adj(a, ncomp, g.L(d[2], d[1], d[0]), g.L(s[2], s[1], s[0]));
}

// ucat = M^-1 q at an interior cell, M the rows csi, eta, zet: qbar = M^-T ucat_bar, by Cramer's
// rule on M^T; written for the interior cells (every other point of qbar stays 0)
__global__ void k_c2c_adj_solve(const double *ubar, const double *csi, const double *eta, const double *zet,
                                double *qbar, Geo g) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1, j = blockIdx.y * blockDim.y + threadIdx.y + 1,
              k = blockIdx.z + 1;
    if (i > g.mx - 2 || j > g.my - 2 || k > g.mz - 2) return;
    const long c = g.L(k, j, i);
    const V3 cs = ld3(csi, c), et = ld3(eta, c), ze = ld3(zet, c);
    // T = M^T: columns csi, eta, zet
    const double t00 = cs.x, t01 = et.x, t02 = ze.x, t10 = cs.y, t11 = et.y, t12 = ze.y, t20 = cs.z, t21 = et.z,
                 t22 = ze.z;
    const V3 r = ld3(ubar, c);
    const double det = t00 * (t11 * t22 - t12 * t21) - t01 * (t10 * t22 - t12 * t20) + t02 * (t10 * t21 - t11 * t20);
    const double d0 = r.x * (t11 * t22 - t12 * t21) - r.y * (t01 * t22 - t02 * t21) + r.z * (t01 * t12 - t02 * t11);
    const double d1 = -r.x * (t10 * t22 - t12 * t20) + r.y * (t00 * t22 - t02 * t20) - r.z * (t00 * t12 - t02 * t10);
    const double d2 = r.x * (t10 * t21 - t11 * t20) - r.y * (t00 * t21 - t01 * t20) + r.z * (t00 * t11 - t01 * t10);
    st3(qbar, c, V3{d0 / det, d1 / det, d2 / det});
}

// q_a of cell c averages the faces c - e_a and c of component a, so face p collects half of qbar_a
// of the cells p and p + e_a (zero outside the interior cells); one thread per owned point
__global__ void k_c2c_adj_gather(const double *qbar, double *vbar, Geo g) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, k = blockIdx.z;
    if (i > g.mx - 1) return;
    const long c = g.L(k, j, i);
    vbar[3 * c + 0] = vbar[3 * c + 0] + 0.5 * (qbar[3 * c + 0] + qbar[3 * (long)g.L(k, j, i + 1) + 0]);
    vbar[3 * c + 1] = vbar[3 * c + 1] + 0.5 * (qbar[3 * c + 1] + qbar[3 * (long)g.L(k, j + 1, i) + 1]);
    vbar[3 * c + 2] = vbar[3 * c + 2] + 0.5 * (qbar[3 * c + 2] + qbar[3 * (long)g.L(k + 1, j, i) + 2]);
}

// xbar (global) += the owned range of vbar (local); one thread per owned double
__global__ void k_add_owned(double *xbar, const double *vbar, Geo g, long n) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    xbar[t] = xbar[t] + vbar[3 * owned_local(g, t / 3) + t % 3];
}

// transpose of the face-0 copy of a periodic axis on a global vector: plane 0 took plane m-2 over
// the whole owned range of the other two axes (all three components)
__global__ void k_face_periodic_adj(double *xbar, Geo g, int axis) {
    const int m[3] = {g.mx, g.my, g.mz};
    const int t1 = (axis == 0) ? 1 : 0, t2 = (axis == 2) ? 1 : 2;
    const int a1 = blockIdx.x * blockDim.x + threadIdx.x, a2 = blockIdx.y;
    if (a1 >= m[t1] || a2 >= m[t2]) return;
    int d[3], s[3];
    d[axis] = 0;
    s[axis] = m[axis] - 2;
    d[t1] = s[t1] = a1;
    d[t2] = s[t2] = a2;
    const long cd = gidx(g, d[2], d[1], d[0]), cs = gidx(g, s[2], s[1], s[0]);
    for (int q = 0; q < 3; q++) {
        xbar[3 * cs + q] = xbar[3 * cs + q] + xbar[3 * cd + q];
        xbar[3 * cd + q] = 0.0;
    }
}

// transpose of k_no_penetration on a global vector, with x the flux before the reset: one thread
// per face of the six patches (the same faces and rules as the forward kernel)
__global__ void k_no_penetration_adj(double *xbar, const double *x, Geo g, Kinds P) {
    const int q = blockIdx.z, a = q >> 1, side = q & 1;
    const int m[3] = {g.mx, g.my, g.mz};
    const int t1 = (a == 0) ? 1 : 0, t2 = (a == 2) ? 1 : 2;
    const int x1 = blockIdx.x * blockDim.x + threadIdx.x, x2 = blockIdx.y;
    if (x1 >= m[t1] - 1 || x2 >= m[t2] - 1) return;
    const int b = P.bc[q];
    if (b != TOSCA_BCU_NOSLIP && b != TOSCA_BCU_WALLFUNCTION && b != TOSCA_BCU_SLIP && b != TOSCA_BCU_ZEROGRADIENT) return;
    int idx[3];
    idx[a] = side ? m[a] - 2 : 0;
    idx[t1] = x1;
    idx[t2] = x2;
    const long o = 3 * gidx(g, idx[2], idx[1], idx[0]) + a;
    if (b == TOSCA_BCU_ZEROGRADIENT) {
        const double v = x[o];
        if ((side == 0 && v > 0.0) || (side == 1 && v < 0.0)) xbar[o] = 0.0;
    } else {
        xbar[o] = 0.0;
    }
}

Kinds kinds_of(const tosca_params *p) {
    Kinds K;
    for (int q = 0; q < 6; q++) K.bc[q] = p->bc_u[q];
    return K;
}

}  // namespace

void launch_rk3w_update(double *ucont, const double *rhs, double *rhs_o, double a, double b, const Geo &g,
                        hipStream_t s) {
    const long n = 3L * g.mx * g.my * g.mz;
    TOSCA_LAUNCH("k_rk3w_update", k_rk3w_update, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, ucont, rhs, rhs_o,
                 a, b, g, n);
}

void launch_rk3w_adj_combine(const double *xbar, double *carry, double *rbar, double a, double b, const Geo &g,
                             hipStream_t s) {
    const long n = 3L * g.mx * g.my * g.mz;
    TOSCA_LAUNCH("k_rk3w_adj_combine", k_rk3w_adj_combine, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, xbar,
                 carry, rbar, a, b, n);
}

// transpose of launch_reset_cell_periodic (k_reset_axis on each periodic axis, then the wraps):
// the wraps' transposes in reverse axis order, then the resets'
void launch_reset_cell_periodic_adj(double *a, int ncomp, const Geo &g, hipStream_t s) {
    const int m[3] = {g.mx, g.my, g.mz};
    for (int axis = 2; axis >= 0; axis--) {
        if (g.per[axis] != TOSCA_PERIODIC_WRAP) continue;
        const int t1 = (axis == 0) ? 1 : 0, t2 = (axis == 2) ? 1 : 2;
        const int n1 = m[t1] + ((g.per[t1] == TOSCA_PERIODIC_WRAP) ? 2 * GW : 0);
        const int n2 = m[t2] + ((g.per[t2] == TOSCA_PERIODIC_WRAP) ? 2 * GW : 0);
        for (int layer0 = GW; layer0 >= 0; layer0 -= GW)
            TOSCA_LAUNCH("k_wrap_axis_adj", k_wrap_axis_adj, dim3((n1 + 127) / 128, n2, GW), dim3(128), 0, s, a, ncomp,
                         g, axis, layer0);
    }
    for (int axis = 2; axis >= 0; axis--) {
        if (g.per[axis] == TOSCA_PERIODIC_NONE) continue;
        const int t1 = (axis == 0) ? 1 : 0, t2 = (axis == 2) ? 1 : 2;
        TOSCA_LAUNCH("k_reset_axis_adj", k_reset_axis_adj, dim3((m[t1] + 127) / 128, m[t2], 2), dim3(128), 0, s, a, ncomp,
                     g, axis);
    }
}

void launch_c2c_adj(const double *ubar, const tosca_mesh *m, double *qbar, double *vbar, const Geo &g, hipStream_t s) {
    launch_fill(qbar, 3L * g.nloc(), 0.0, s);
    TOSCA_LAUNCH("k_c2c_adj_solve", k_c2c_adj_solve, dim3((g.mx - 2 + 63) / 64, (g.my - 2 + 3) / 4, g.mz - 2),
                 dim3(64, 4), 0, s, ubar, m->csi, m->eta, m->zet, qbar, g);
    TOSCA_LAUNCH("k_c2c_adj_gather", k_c2c_adj_gather, dim3((g.mx + 255) / 256, g.my, g.mz), dim3(256), 0, s, qbar, vbar,
                 g);
}

void launch_add_owned(double *xbar, const double *vbar, const Geo &g, hipStream_t s) {
    const long n = 3L * g.mx * g.my * g.mz;
    TOSCA_LAUNCH("k_add_owned", k_add_owned, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, xbar, vbar, g, n);
}

// transpose of the ExplicitRhsU face resets y = Pf(Pn(x)) on a global cotangent: Pf's copies in
// reverse axis order, then Pn's mask (x: the state before the resets)
void launch_face_resets_adj(double *xbar, const double *x, const tosca_params *p, const Geo &g, hipStream_t s) {
    const int m[3] = {g.mx, g.my, g.mz};
    for (int axis = 2; axis >= 0; axis--) {
        if (g.per[axis] == TOSCA_PERIODIC_NONE) continue;
        const int t1 = (axis == 0) ? 1 : 0, t2 = (axis == 2) ? 1 : 2;
        TOSCA_LAUNCH("k_face_periodic_adj", k_face_periodic_adj, dim3((m[t1] + 127) / 128, m[t2]), dim3(128), 0, s, xbar,
                     g, axis);
    }
    const int n = std::max(g.mx, std::max(g.my, g.mz));
    TOSCA_LAUNCH("k_no_penetration_adj", k_no_penetration_adj, dim3((n + 127) / 128, n, 6), dim3(128), 0, s, xbar, x, g,
                 kinds_of(p));
}

}  // namespace tosca
