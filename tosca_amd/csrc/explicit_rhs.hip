// explicit_rhs.hip -- the steps of ExplicitRhsU (src/solvers/ueqn_solvers.c:1276-1407) around
// FormU, on gfx950.  Compiled with FP contraction off: every step is a copy, a selection or a
// short expression evaluated exactly as the reference writes it.
//
//   resetNoPenetrationFluxes   src/include/inline.h:5262-5355   -> k_no_penetration
//   contravariantToCartesian   src/ueqn.c:528-635               -> k_contra_to_cart
//   Rhs = -dP + bTheta terms + farm (VecSet / VecAXPY order)     -> k_rhs_init
//   resetNonResolvedCellFaces  src/include/inline.h:1674-1838   -> k_non_resolved
// and the vector updates of the explicit predictors UeqnFE / UeqnRK4 and of IMEXMatVec
// (ueqn_solvers.c:1411-1504, 1739-1801): k_fe_update, k_waxpy_local, k_axpy, k_global_to_local.
// The periodic resets of ucont and ucat are the reference-structured ghost kernels
// (launch_reset_face_periodic, launch_reset_cell_periodic).
#include "kernels.h"

namespace tosca {

namespace {

struct PatchKinds { int bc[6]; };   // boundaryU kinds: iLeft iRight jLeft jRight kLeft kRight

// the two axes across axis a, in the order the ghost kernels use
__device__ __forceinline__ void across(int a, int &t1, int &t2) {
    t1 = (a == 0) ? 1 : 0;
    t2 = (a == 2) ? 1 : 2;
}

// One thread per face of the six boundary patches.  The reference loops k < mz-1, j < my-1,
// i < mx-1 and tests each patch; patch q = 2a + side sits at index 0 (side 0) or m_a - 2
// (side 1) of axis a and changes only component a, so the six patches never write the same
// double.  noSlip, velocityWallFunction and slip: no normal flux; zeroGradient: the inflowing
// normal flux is removed (> 0 on a left patch, < 0 on a right patch).
__global__ void k_no_penetration(double *ucont, Geo g, PatchKinds P) {
    const int q = blockIdx.z, a = q >> 1, side = q & 1;
    const int m[3] = {g.mx, g.my, g.mz};
    int t1, t2;
    across(a, t1, t2);
    const int x1 = blockIdx.x * blockDim.x + threadIdx.x, x2 = blockIdx.y;
    if (x1 >= m[t1] - 1 || x2 >= m[t2] - 1) return;
    const int b = P.bc[q];
    if (b != TOSCA_BCU_NOSLIP && b != TOSCA_BCU_WALLFUNCTION && b != TOSCA_BCU_SLIP && b != TOSCA_BCU_ZEROGRADIENT) return;
    int idx[3];
    idx[a] = side ? m[a] - 2 : 0;
    idx[t1] = x1;
    idx[t2] = x2;
    double *u = ucont + 3 * (long)g.L(idx[2], idx[1], idx[0]) + a;
    const double v = *u;
    if (b == TOSCA_BCU_ZEROGRADIENT) {
        if ((side == 0 && v > 0.0) || (side == 1 && v < 0.0)) *u = 0.0;
    } else {
        *u = 0.0;
    }
}

// Cartesian velocity of an interior cell: the contravariant fluxes of its opposite faces
// averaged (central, inline.h:4307-4310) and the 3x3 system of the cell basis solved by
// Cramer's rule in the reference's expression order (src/ueqn.c:571-613).
__global__ void k_contra_to_cart(const double *ucont, const double *csi, const double *eta, const double *zet,
                                 double *ucat, Geo g) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1, j = blockIdx.y * blockDim.y + threadIdx.y + 1,
              k = blockIdx.z + 1;
    if (i > g.mx - 2 || j > g.my - 2 || k > g.mz - 2) return;
    const long c = g.L(k, j, i);
    const V3 cs = ld3(csi, c), et = ld3(eta, c), ze = ld3(zet, c);
    const double m00 = cs.x, m01 = cs.y, m02 = cs.z, m10 = et.x, m11 = et.y, m12 = et.z, m20 = ze.x, m21 = ze.y,
                 m22 = ze.z;
    const double q0 = 0.5 * (ucont[3 * (long)g.L(k, j, i - 1)] + ucont[3 * c]);
    const double q1 = 0.5 * (ucont[3 * (long)g.L(k, j - 1, i) + 1] + ucont[3 * c + 1]);
    const double q2 = 0.5 * (ucont[3 * (long)g.L(k - 1, j, i) + 2] + ucont[3 * c + 2]);
    const double det = m00 * (m11 * m22 - m12 * m21) - m01 * (m10 * m22 - m12 * m20) + m02 * (m10 * m21 - m11 * m20);
    const double det0 = q0 * (m11 * m22 - m12 * m21) - q1 * (m01 * m22 - m02 * m21) + q2 * (m01 * m12 - m02 * m11);
    const double det1 = -q0 * (m10 * m22 - m12 * m20) + q1 * (m00 * m22 - m02 * m20) - q2 * (m00 * m12 - m02 * m10);
    const double det2 = q0 * (m10 * m21 - m11 * m20) - q1 * (m00 * m21 - m01 * m20) + q2 * (m00 * m11 - m01 * m10);
    ucat[3 * c] = det0 / det;
    ucat[3 * c + 1] = det1 / det;
    ucat[3 * c + 2] = det2 / det;
}

// Rhs = 0; Rhs += -1 dP; [Rhs += 3/2 bTheta; Rhs += -1/2 bTheta_o | Rhs += bTheta]; [Rhs += farm]
// (ueqn_solvers.c:1304-1344), one VecAXPY (y = y + a x) after the other, over every owned double
__global__ void k_rhs_init(double *rhs, const double *dp, const double *bt, const double *bto, int ab2,
                           const double *farm, long n) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    double r = 0.0;
    r = r + -1.0 * dp[t];
    if (bt) {
        if (ab2) {
            r = r + (3.0 / 2.0) * bt[t];
            r = r + (-1.0 / 2.0) * bto[t];
        } else {
            r = r + 1.0 * bt[t];
        }
    }
    if (farm) r = r + 1.0 * farm[t];
    rhs[t] = r;
}

// resetNonResolvedCellFaces on a global vector: per axis a the planes 0, m-1 and (non-periodic
// axes) m-2, each over the whole owned range of the other two axes, component a only.  Plane 0
// keeps its value on a zeroGradient patch or where the old flux leaves the domain (ucont_o < 0),
// plane m-2 likewise (ucont_o > 0); plane m-1 is an unused face and is zeroed.
__global__ void k_non_resolved(double *v, const double *ucont_o, Geo g, PatchKinds P) {
    const int a = blockIdx.z / 3, kind = blockIdx.z % 3;      // kind: 0 -> index 0, 1 -> m-1, 2 -> m-2
    const int m[3] = {g.mx, g.my, g.mz};
    int t1, t2;
    across(a, t1, t2);
    const int x1 = blockIdx.x * blockDim.x + threadIdx.x, x2 = blockIdx.y;
    if (x1 >= m[t1] || x2 >= m[t2]) return;
    if (kind == 2 && g.per[a] != TOSCA_PERIODIC_NONE) return;
    int idx[3];
    idx[a] = (kind == 0) ? 0 : (kind == 1) ? m[a] - 1 : m[a] - 2;
    idx[t1] = x1;
    idx[t2] = x2;
    const long gi = 3 * (((long)idx[2] * g.my + idx[1]) * g.mx + idx[0]) + a;
    const double uo = ucont_o[gi];
    bool keep = false;
    if (kind == 0) keep = P.bc[2 * a] == TOSCA_BCU_ZEROGRADIENT || uo < 0;
    else if (kind == 2) keep = P.bc[2 * a + 1] == TOSCA_BCU_ZEROGRADIENT || uo > 0;
    if (!keep) v[gi] = 0.0;
}

PatchKinds kinds(const tosca_params *p) {
    PatchKinds P;
    for (int q = 0; q < 6; q++) P.bc[q] = p->bc_u[q];
    return P;
}

int maxdim(const Geo &g) { return std::max(g.mx, std::max(g.my, g.mz)); }

// local index of owned point q (global order [k][j][i])
__device__ __forceinline__ long owned_local(const Geo &g, long q) {
    const int i = (int)(q % g.mx);
    const long r = q / g.mx;
    const int j = (int)(r % g.my), k = (int)(r / g.my);
    return g.L(k, j, i);
}

// UeqnFE (ueqn_solvers.c:1427-1433): VecScale(Rhs, dt); VecAXPY(Rhs, 1, Ucont_o); VecCopy(Rhs, Ucont),
// the copy going into the owned range of the local array; one thread per double
__global__ void k_fe_update(double *rhs, const double *uo, double dt, double *ucont, Geo g, long n) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    double r = dt * rhs[t];
    r = r + 1.0 * uo[t];
    rhs[t] = r;
    ucont[3 * owned_local(g, t / 3) + t % 3] = r;
}

// VecWAXPY(Ucont, al, Rhs, Ucont_o) into the owned range of the local array (ueqn_solvers.c:1484)
__global__ void k_waxpy_local(double *ucont, double al, const double *rhs, const double *uo, Geo g, long n) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    ucont[3 * owned_local(g, t / 3) + t % 3] = al * rhs[t] + uo[t];
}

// One pass over a Runge-Kutta stage's Rhs (ueqn_solvers.c:1480-1497): the accumulator update
// utmp = utmp + a Rhs (VecAXPY) and the owned range of the next stage's local ucont, either
// al Rhs + Ucont_o (VecWAXPY, use_acc == 0) or the updated accumulator (the final copy,
// use_acc == 1); each value is formed by the same expression as in the separate passes
__global__ void k_rk_stage(double *utmp, double a, const double *rhs, double *ucont, double al, const double *uo,
                           int use_acc, Geo g, long n) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const double r = rhs[t];
    const double acc = utmp[t] + a * r;
    utmp[t] = acc;
    ucont[3 * owned_local(g, t / 3) + t % 3] = use_acc ? acc : al * r + uo[t];
}

// VecAXPY(y, a, x): y = y + a x
__global__ void k_axpy(double *y, double a, const double *x, long n) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    y[t] = y[t] + a * x[t];
}

// UeqnSNES's constant right-hand side in one pass: bU = -dP (+ 3/2 bTheta - 1/2 bTheta_o, or
// + bTheta) (+ 1/2 Rhs_o) (+ farm), each term added to the running value in the order of the
// reference's VecScale / VecAXPY sequence (ueqn_solvers.c:1102-1156), so per element the result
// equals the one of the separate passes; null pointers skip their terms
__global__ void k_cn_rhs(double *bU, const double *dp, const double *bt, const double *bto, const double *rhs_o,
                         const double *farm, double wbt, long n) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    double y = dp[t] * -1.0;
    if (bt) y = y + wbt * bt[t];
    if (bto) y = y + -1.0 / 2.0 * bto[t];
    if (rhs_o) y = y + 0.5 * rhs_o[t];
    if (farm) y = y + 1.0 * farm[t];
    bU[t] = y;
}

// The end of SNESFuncEval in one pass (ueqn_solvers.c:1264-1271): F = dt F, the non-resolved
// face reset of k_non_resolved (same rule, evaluated per element), then F - u + ucont_o, in the
// order of the reference's VecScale / reset / VecAXPY sequence
__global__ void k_snes_finish(double *F, const double *u, const double *ucont_o, double dt, Geo g, PatchKinds P, long n) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int a = (int)(t % 3);
    const long q = t / 3;
    const int i = (int)(q % g.mx);
    const long r = q / g.mx;
    const int j = (int)(r % g.my), k = (int)(r / g.my);
    const int idx = (a == 0) ? i : (a == 1) ? j : k, m = (a == 0) ? g.mx : (a == 1) ? g.my : g.mz;
    double y = F[t] * dt;
    const double uo = ucont_o[t];
    bool zero = false;
    if (idx == 0) zero = !(P.bc[2 * a] == TOSCA_BCU_ZEROGRADIENT || uo < 0);
    else if (idx == m - 1) zero = true;
    else if (idx == m - 2 && g.per[a] == TOSCA_PERIODIC_NONE) zero = !(P.bc[2 * a + 1] == TOSCA_BCU_ZEROGRADIENT || uo > 0);
    if (zero) y = 0.0;
    y = y + -1.0 * u[t];
    y = y + 1.0 * uo;
    F[t] = y;
}

// the owned-range copy of DMGlobalToLocal for a vector field
__global__ void k_global_to_local(double *loc, const double *glob, Geo g, long n) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    loc[3 * owned_local(g, t / 3) + t % 3] = glob[t];
}

}  // namespace

void launch_no_penetration(double *ucont, const tosca_params *p, const Geo &g, hipStream_t s) {
    const int n = maxdim(g);
    TOSCA_LAUNCH("k_no_penetration", k_no_penetration, dim3((n + 127) / 128, n, 6), dim3(128), 0, s, ucont, g, kinds(p));
}

void launch_contra_to_cart(const double *ucont, const tosca_mesh *m, double *ucat, const Geo &g, hipStream_t s) {
    TOSCA_LAUNCH("k_contra_to_cart", k_contra_to_cart, dim3((g.mx - 2 + 63) / 64, (g.my - 2 + 3) / 4, g.mz - 2),
                 dim3(64, 4), 0, s, ucont, m->csi, m->eta, m->zet, ucat, g);
}

void launch_rhs_init(double *rhs, const double *dp, const double *bt, const double *bto, int ab2, const double *farm,
                     const Geo &g, hipStream_t s) {
    const long n = 3L * g.mx * g.my * g.mz;
    TOSCA_LAUNCH("k_rhs_init", k_rhs_init, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, rhs, dp, bt, bto, ab2, farm, n);
}

void launch_fe_update(double *rhs, const double *ucont_o, double dt, double *ucont, const Geo &g, hipStream_t s) {
    const long n = 3L * g.mx * g.my * g.mz;
    TOSCA_LAUNCH("k_fe_update", k_fe_update, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, rhs, ucont_o, dt, ucont, g, n);
}

void launch_waxpy_local(double *ucont, double al, const double *rhs, const double *ucont_o, const Geo &g, hipStream_t s) {
    const long n = 3L * g.mx * g.my * g.mz;
    TOSCA_LAUNCH("k_waxpy_local", k_waxpy_local, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, ucont, al, rhs, ucont_o, g, n);
}

void launch_rk_stage(double *utmp, double a, const double *rhs, double *ucont, double al, const double *ucont_o,
                     int use_acc, const Geo &g, hipStream_t s) {
    const long n = 3L * g.mx * g.my * g.mz;
    TOSCA_LAUNCH("k_rk_stage", k_rk_stage, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, utmp, a, rhs, ucont, al,
                 ucont_o, use_acc, g, n);
}

void launch_axpy(double *y, double a, const double *x, long n, hipStream_t s) {
    TOSCA_LAUNCH("k_axpy", k_axpy, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, y, a, x, n);
}

void launch_cn_rhs(double *bU, const double *dp, const double *bt, const double *bto, const double *rhs_o,
                   const double *farm, double wbt, long n, hipStream_t s) {
    TOSCA_LAUNCH("k_cn_rhs", k_cn_rhs, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, bU, dp, bt, bto, rhs_o, farm,
                 wbt, n);
}

void launch_global_to_local(double *loc, const double *glob, const Geo &g, hipStream_t s) {
    const long n = 3L * g.mx * g.my * g.mz;
    TOSCA_LAUNCH("k_global_to_local", k_global_to_local, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, loc, glob, g, n);
}

void launch_snes_finish(double *F, const double *u, const double *ucont_o, double dt, const tosca_params *p,
                        const Geo &g, hipStream_t s) {
    const long n = 3L * g.mx * g.my * g.mz;
    TOSCA_LAUNCH("k_snes_finish", k_snes_finish, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, F, u, ucont_o, dt, g,
                 kinds(p), n);
}

void launch_non_resolved(double *rhs, const double *ucont_o, const tosca_params *p, const Geo &g, hipStream_t s) {
    const int n = maxdim(g);
    TOSCA_LAUNCH("k_non_resolved", k_non_resolved, dim3((n + 127) / 128, n, 9), dim3(128), 0, s, rhs, ucont_o, g, kinds(p));
}

}  // namespace tosca
