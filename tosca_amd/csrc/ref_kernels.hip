// ref_kernels.hip -- the reference-structured gfx950 path: one thread per face or cell, the same
// passes and expression order as the reference routines, compiled with FP contraction off so its
// results track the CPU oracle to the last bits.  It serves every entry point and is the baseline
// the fused kernels (fused_kernels.hip) are checked and timed against.
//
//   FormU       src/solvers/ueqn_solvers.c:6-1083   -> k_formu_faces, k_formu_cells, k_formu_dot
//   UpdateCs    src/les.c:165-2115                  -> k_strain, k_test_filter<3|5>, k_box_clip,
//                                                      k_plane_partial/k_plane_final/k_dpasd_cs
//   UpdateNut   src/les.c:2119-2578 + boundary.c:2630-2929 -> k_nut, k_nut_bc
//   Buoyancy    src/sources/ueqn_sources.c:2726-2880 -> k_buoyancy
//   ghost scatters (PETSc DMLocalToLocal), resetCellPeriodicFluxes and
//   resetFacePeriodicFluxesVector -> k_wrap_axis, k_reset_axis
#include "kernels.h"

namespace tosca {

// ---------------------------------------------------------------------------------------------
// ghost handling
// ---------------------------------------------------------------------------------------------
// DMLocalToLocal along one DM_BOUNDARY_PERIODIC axis: ghost slabs take the wrapped owned values.
// The transverse ranges cover the ghosts of periodic-wrap axes and the owned range otherwise.
template <class T>
__global__ void k_wrap_axis(T *a, int ncomp, Geo g, int axis) {
    const int m[3] = {g.mx, g.my, g.mz};
    int t1 = (axis == 0) ? 1 : 0, t2 = (axis == 2) ? 1 : 2;
    int lo1 = (g.per[t1] == TOSCA_PERIODIC_WRAP) ? -GW : 0, n1 = m[t1] + ((g.per[t1] == TOSCA_PERIODIC_WRAP) ? 2 * GW : 0);
    int lo2 = (g.per[t2] == TOSCA_PERIODIC_WRAP) ? -GW : 0, n2 = m[t2] + ((g.per[t2] == TOSCA_PERIODIC_WRAP) ? 2 * GW : 0);
    int a1 = blockIdx.x * blockDim.x + threadIdx.x;
    int a2 = blockIdx.y;
    int s = blockIdx.z;                   // 0..2*GW-1: which ghost layer
    if (a1 >= n1 || a2 >= n2) return;
    int n = m[axis];
    int dst = (s < GW) ? (s - GW) : (n + s - GW);       // -3..-1, n..n+2
    int src = (s < GW) ? (n + s - GW) : (s - GW);       // n-3..n-1, 0..2
    int idx_d[3], idx_s[3];
    idx_d[axis] = dst; idx_s[axis] = src;
    idx_d[t1] = idx_s[t1] = lo1 + a1;
    idx_d[t2] = idx_s[t2] = lo2 + a2;
    int cd = g.L(idx_d[2], idx_d[1], idx_d[0]), cs = g.L(idx_s[2], idx_s[1], idx_s[0]);
    for (int q = 0; q < ncomp; q++) a[ncomp * cd + q] = a[ncomp * cs + q];
}

// resetCellPeriodicFluxes along one periodic axis: boundary ghost cells 0 / m-1 take cells
// m-2 / 1 over the whole owned range of the other two axes (src/include/inline.h:5496-5514).
__global__ void k_reset_axis(double *a, int ncomp, Geo g, int axis) {
    const int m[3] = {g.mx, g.my, g.mz};
    int t1 = (axis == 0) ? 1 : 0, t2 = (axis == 2) ? 1 : 2;
    int a1 = blockIdx.x * blockDim.x + threadIdx.x, a2 = blockIdx.y, side = blockIdx.z;
    if (a1 >= m[t1] || a2 >= m[t2]) return;
    int idx_d[3], idx_s[3];
    idx_d[axis] = side ? m[axis] - 1 : 0;
    idx_s[axis] = side ? 1 : m[axis] - 2;
    idx_d[t1] = idx_s[t1] = a1;
    idx_d[t2] = idx_s[t2] = a2;
    int cd = g.L(idx_d[2], idx_d[1], idx_d[0]), cs = g.L(idx_s[2], idx_s[1], idx_s[0]);
    for (int q = 0; q < ncomp; q++) a[ncomp * cd + q] = a[ncomp * cs + q];
}

__global__ void k_fill(double *a, long n, double v) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

// ---------------------------------------------------------------------------------------------
// FormU
// ---------------------------------------------------------------------------------------------
struct FaceArgs {
    const double *ucont, *ucat, *nut;
    const double *icsi, *ieta, *izet, *jcsi, *jeta, *jzet, *kcsi, *keta, *kzet;
    const double *iaj, *jaj, *kaj;
    const double *tauI, *tauJ;
    double *div[3], *visc[3];
    int bc_u[6];
    int quick, les, mode;
    double nu;
};

// Face pass (ueqn_solvers.c:173-808): each thread owns DMDA point (k,j,i) and its three faces.
__global__ void k_formu_faces(FaceArgs A, Geo g) {
    int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y * blockDim.y + threadIdx.y, k = blockIdx.z;
    if (i >= g.mx - 1 || j >= g.my - 1 || k >= g.mz - 1) return;
    const int c = g.L(k, j, i);
    const double *u = A.ucat;
    // i face
    if (j != 0 && k != 0) {
        if (A.mode != 2) {
            double V = A.ucont[3 * c];
            V3 f;
            if (A.quick) { int L, R; face2cell4(g.per[0], i, g.mx, L, R);
                f = quick(ld3(u, g.L(k, j, L)), ld3(u, c), ld3(u, g.L(k, j, i + 1)), ld3(u, g.L(k, j, R)), V); }
            else { V3 a = ld3(u, c), b = ld3(u, g.L(k, j, i + 1)); f = V3{0.5 * (a.x + b.x), 0.5 * (a.y + b.y), 0.5 * (a.z + b.z)}; }
            st3(A.div[0], c, V3{-V * f.x, -V * f.y, -V * f.z});
        }
        if (A.mode != 1) {
            Grad gr = du_i(g, u, i, j, k);
            double nuEff;
            V3 base{0, 0, 0};
            if (A.les) {
                double nt = 0.5 * (A.nut[c] + A.nut[g.L(k, j, i + 1)]);
                if ((A.bc_u[0] == TOSCA_BCU_WALLFUNCTION && i == 0) || (A.bc_u[1] == TOSCA_BCU_WALLFUNCTION && i == g.mx - 2)) {
                    int t = k * g.my + j, P = g.mz * g.my;
                    base = V3{-A.tauI[t], -A.tauI[P + t], -A.tauI[2 * P + t]};
                    nuEff = A.nu;
                } else if ((A.bc_u[0] == TOSCA_BCU_SLIP && i == 0) || (A.bc_u[1] == TOSCA_BCU_SLIP && i == g.mx - 2)) nuEff = 0.0;
                else nuEff = A.nu + nt;
            } else nuEff = A.nu;
            V3 v = visc_flux(gr, ld3(A.icsi, c), ld3(A.ieta, c), ld3(A.izet, c), 0, A.iaj[c], nuEff);
            st3(A.visc[0], c, V3{base.x + v.x, base.y + v.y, base.z + v.z});
        }
    }
    // j face
    if (i != 0 && k != 0) {
        if (A.mode != 2) {
            double V = A.ucont[3 * c + 1];
            V3 f;
            if (A.quick) { int L, R; face2cell4(g.per[1], j, g.my, L, R);
                f = quick(ld3(u, g.L(k, L, i)), ld3(u, c), ld3(u, g.L(k, j + 1, i)), ld3(u, g.L(k, R, i)), V); }
            else { V3 a = ld3(u, c), b = ld3(u, g.L(k, j + 1, i)); f = V3{0.5 * (a.x + b.x), 0.5 * (a.y + b.y), 0.5 * (a.z + b.z)}; }
            st3(A.div[1], c, V3{-V * f.x, -V * f.y, -V * f.z});
        }
        if (A.mode != 1) {
            Grad gr = du_j(g, u, i, j, k);
            double nuEff;
            V3 base{0, 0, 0};
            if (A.les) {
                double nt = 0.5 * (A.nut[c] + A.nut[g.L(k, j + 1, i)]);
                if ((A.bc_u[2] == TOSCA_BCU_WALLFUNCTION && j == 0) || (A.bc_u[3] == TOSCA_BCU_WALLFUNCTION && j == g.my - 2)) {
                    int t = k * g.mx + i, P = g.mz * g.mx;
                    base = V3{-A.tauJ[t], -A.tauJ[P + t], -A.tauJ[2 * P + t]};
                    nuEff = A.nu;
                } else if ((A.bc_u[2] == TOSCA_BCU_SLIP && j == 0) || (A.bc_u[3] == TOSCA_BCU_SLIP && j == g.my - 2)) nuEff = 0.0;
                else nuEff = A.nu + nt;
            } else nuEff = A.nu;
            V3 v = visc_flux(gr, ld3(A.jcsi, c), ld3(A.jeta, c), ld3(A.jzet, c), 1, A.jaj[c], nuEff);
            st3(A.visc[1], c, V3{base.x + v.x, base.y + v.y, base.z + v.z});
        }
    }
    // k face
    if (i != 0 && j != 0) {
        if (A.mode != 2) {
            double V = A.ucont[3 * c + 2];
            V3 f;
            if (A.quick) { int L, R; face2cell4(g.per[2], k, g.mz, L, R);
                f = quick(ld3(u, g.L(L, j, i)), ld3(u, c), ld3(u, g.L(k + 1, j, i)), ld3(u, g.L(R, j, i)), V); }
            else { V3 a = ld3(u, c), b = ld3(u, g.L(k + 1, j, i)); f = V3{0.5 * (a.x + b.x), 0.5 * (a.y + b.y), 0.5 * (a.z + b.z)}; }
            st3(A.div[2], c, V3{-V * f.x, -V * f.y, -V * f.z});
        }
        if (A.mode != 1) {
            Grad gr = du_k(g, u, i, j, k);
            double nuEff;
            if (A.les) {
                double nt = 0.5 * (A.nut[c] + A.nut[g.L(k + 1, j, i)]);
                if ((A.bc_u[4] == TOSCA_BCU_SLIP && k == 0) || (A.bc_u[5] == TOSCA_BCU_SLIP && k == g.mz - 2)) nuEff = 0.0;
                else nuEff = A.nu + nt;
            } else nuEff = A.nu;
            st3(A.visc[2], c, visc_flux(gr, ld3(A.kcsi, c), ld3(A.keta, c), ld3(A.kzet, c), 2, A.kaj[c], nuEff));
        }
    }
}

// Face index seen by the cell pass after resetFacePeriodicFluxesVector: on a periodic axis the
// face 0 holds the value of face m-2.
__device__ __forceinline__ int face_lo(int f, int ptype, int m) { return (f == 0 && ptype != TOSCA_PERIODIC_NONE) ? m - 2 : f; }

// Cell pass (ueqn_solvers.c:871-948): fp = sum of face-flux differences.
struct CellArgs { const double *div[3], *visc[3]; };

__global__ void k_formu_cells(CellArgs A, double *fp, int mode, Geo g) {
    const double *const *div = A.div, *const *visc = A.visc;
    int i = blockIdx.x * blockDim.x + threadIdx.x + 1, j = blockIdx.y * blockDim.y + threadIdx.y + 1, k = blockIdx.z + 1;
    if (i >= g.mx - 1 || j >= g.my - 1 || k >= g.mz - 1) return;
    const int c = g.L(k, j, i);
    const int ci = g.L(k, j, face_lo(i - 1, g.per[0], g.mx));
    const int cj = g.L(k, face_lo(j - 1, g.per[1], g.my), i);
    const int ck = g.L(face_lo(k - 1, g.per[2], g.mz), j, i);
    double f[3] = {0.0, 0.0, 0.0};
    if (mode != 2) {
        for (int q = 0; q < 3; q++)
            f[q] += (div[0][3 * c + q] - div[0][3 * ci + q] + div[1][3 * c + q] - div[1][3 * cj + q] + div[2][3 * c + q] - div[2][3 * ck + q]);
    }
    if (mode != 1) {
        for (int q = 0; q < 3; q++)
            f[q] += (visc[0][3 * c + q] - visc[0][3 * ci + q] + visc[1][3 * c + q] - visc[1][3 * cj + q] + visc[2][3 * c + q] - visc[2][3 * ck + q]);
    }
    fp[3 * c] = f[0]; fp[3 * c + 1] = f[1]; fp[3 * c + 2] = f[2];
}

// fp index on the + side of the last interior cell (ghost m-1): a copy of cell m-2 on a
// non-periodic axis (ueqn_solvers.c:967-990), of cell 1 on a periodic one (resetCellPeriodicFluxes).
__device__ __forceinline__ int cell_hi(int c, int ptype, int m) { return (c == m - 1) ? ((ptype == TOSCA_PERIODIC_NONE) ? m - 2 : 1) : c; }

// Face-dot pass (ueqn_solvers.c:1006-1035).
__global__ void k_formu_dot(const double *fp, const double *csi, const double *eta, const double *zet,
                            const double *iaj, const double *jaj, const double *kaj, double *rhs, double scale, Geo g) {
    int i = blockIdx.x * blockDim.x + threadIdx.x + 1, j = blockIdx.y * blockDim.y + threadIdx.y + 1, k = blockIdx.z + 1;
    if (i >= g.mx - 1 || j >= g.my - 1 || k >= g.mz - 1) return;
    const int c = g.L(k, j, i);
    const int ci = g.L(k, j, i + 1), cj = g.L(k, j + 1, i), ck = g.L(k + 1, j, i);
    const int fi = g.L(k, j, cell_hi(i + 1, g.per[0], g.mx));
    const int fj = g.L(k, cell_hi(j + 1, g.per[1], g.my), i);
    const int fk = g.L(cell_hi(k + 1, g.per[2], g.mz), j, i);
    V3 f0 = ld3(fp, c);
    const int r = g.G(k, j, i);
    double v1, v2;
    V3 a = ld3(csi, c), b = ld3(csi, ci), fb = ld3(fp, fi);
    v1 = a.x * f0.x + a.y * f0.y + a.z * f0.z;
    v2 = b.x * fb.x + b.y * fb.y + b.z * fb.z;
    rhs[3 * r] += scale * iaj[c] * (0.5 * (v1 + v2));
    a = ld3(eta, c); b = ld3(eta, cj); fb = ld3(fp, fj);
    v1 = a.x * f0.x + a.y * f0.y + a.z * f0.z;
    v2 = b.x * fb.x + b.y * fb.y + b.z * fb.z;
    rhs[3 * r + 1] += scale * jaj[c] * (0.5 * (v1 + v2));
    a = ld3(zet, c); b = ld3(zet, ck); fb = ld3(fp, fk);
    v1 = a.x * f0.x + a.y * f0.y + a.z * f0.z;
    v2 = b.x * fb.x + b.y * fb.y + b.z * fb.z;
    rhs[3 * r + 2] += scale * kaj[c] * (0.5 * (v1 + v2));
}

// ---------------------------------------------------------------------------------------------
// UpdateCs
// ---------------------------------------------------------------------------------------------
// Strain pass (les.c:277-374): velocity gradient rows (lSx, lSy, lSz) and |S| at interior cells.
__global__ void k_strain(const double *ucat, const double *csi, const double *eta, const double *zet,
                         const double *aj, double *S, Geo g) {
    int i = blockIdx.x * blockDim.x + threadIdx.x + 1, j = blockIdx.y * blockDim.y + threadIdx.y + 1, k = blockIdx.z + 1;
    if (i >= g.mx - 1 || j >= g.my - 1 || k >= g.mz - 1) return;
    const int c = g.L(k, j, i);
    Grad gr = du_center(g, ucat, i, j, k);
    Cart d = du_dxyz(ld3(csi, c), ld3(eta, c), ld3(zet, c), aj[c], gr);
    double *s = S + 10 * (long)c;
    s[0] = d.du_dx; s[1] = d.du_dy; s[2] = d.du_dz;
    s[3] = d.dv_dx; s[4] = d.dv_dy; s[5] = d.dv_dz;
    s[6] = d.dw_dx; s[7] = d.dw_dy; s[8] = d.dw_dz;
    s[9] = strain_mag(d);
}

// Masked Simpson test filter + Germano at interior cells (les.c:466-801 for NP=3, :820-1164 for
// NP=5).  The 34 filtered quantities are accumulated point by point in the reference's
// [R][Q][P] order; masked points (DMDA boundary ghosts) contribute nothing.
template <int NP>
__global__ void k_test_filter(const double *ucat, const double *csi, const double *eta, const double *zet,
                              const double *aj, const double *S, double *num_out, double *den_out, Geo g) {
    int i = blockIdx.x * blockDim.x + threadIdx.x + 1, j = blockIdx.y * blockDim.y + threadIdx.y + 1, k = blockIdx.z + 1;
    if (i >= g.mx - 1 || j >= g.my - 1 || k >= g.mz - 1) return;
    constexpr int H = NP / 2;
    // sum_vol in the reference's p (i), q (j), r (k) loop order
    double sum_vol = 0;
    for (int p = -H; p <= H; p++)
        for (int q = -H; q <= H; q++)
            for (int r = -H; r <= H; r++) {
                int K = k + r, J = j + q, I = i + p;
                bool in = (I > 0 && I < g.mx - 1 && J > 0 && J < g.my - 1 && K > 0 && K < g.mz - 1);
                if (in) {
                    double coef, scl;
                    if (NP == 3) { const double b3[3] = {0.25, 0.5, 0.25}; coef = b3[r + H] * b3[q + H] * b3[p + H]; scl = 8.0; }
                    else { const double b5[5] = {1.0 / 16, 4.0 / 16, 6.0 / 16, 4.0 / 16, 1.0 / 16}; coef = b5[r + H] * b5[q + H] * b5[p + H]; scl = 64.0; }
                    sum_vol += 1. / aj[g.L(K, J, I)] * scl * coef;
                }
            }
    double wsum = 0, acc[33];
    for (int n = 0; n < 33; n++) acc[n] = 0;
    for (int R = 0; R < NP; R++)
        for (int Q = 0; Q < NP; Q++)
            for (int P = 0; P < NP; P++) {
                int K = k + R - H, J = j + Q - H, I = i + P - H;
                bool in = (I > 0 && I < g.mx - 1 && J > 0 && J < g.my - 1 && K > 0 && K < g.mz - 1);
                double sw;
                if (NP == 3) sw = (R == 1 ? 4. : 1.) * (Q == 1 ? 4. : 1.) * (P == 1 ? 4. : 1.);
                else {
                    auto s5 = [](int t) { return (t == 0 || t == 4) ? 1. : ((t % 2 == 0) ? 2. : 4.); };
                    sw = 1.0 * s5(R) * s5(Q) * s5(P);
                }
                if (!in) continue;
                const int c = g.L(K, J, I);
                wsum += sw * 1.0;
                V3 uu = ld3(ucat, c), cc = ld3(csi, c), ee = ld3(eta, c), zz = ld3(zet, c);
                double U = uu.x * cc.x + uu.y * cc.y + uu.z * cc.z;
                double V = uu.x * ee.x + uu.y * ee.y + uu.z * ee.z;
                double W = uu.x * zz.x + uu.y * zz.y + uu.z * zz.z;
                const double *s = S + 10 * (long)c;
                double Sxx = 0.5 * (s[0] + s[0]), Sxy = 0.5 * (s[1] + s[3]), Sxz = 0.5 * (s[2] + s[6]);
                double Syy = 0.5 * (s[4] + s[4]), Syz = 0.5 * (s[5] + s[7]), Szz = 0.5 * (s[8] + s[8]);
                double Sm = s[9];
                double v[33] = {uu.x, uu.y, uu.z, U, V, W,
                                U * uu.x, U * uu.y, U * uu.z, V * uu.x, V * uu.y, V * uu.z, W * uu.x, W * uu.y, W * uu.z,
                                Sxx, Sxy, Sxz, Sxy, Syy, Syz, Sxz, Syz, Szz,
                                Sxx * Sm, Sxy * Sm, Sxz * Sm, Sxy * Sm, Syy * Sm, Syz * Sm, Sxz * Sm, Syz * Sm, Szz * Sm};
                for (int n = 0; n < 33; n++) acc[n] += sw * 1.0 * v[n];
            }
    const int c0 = g.L(k, j, i);
    double filter = pow(1. / aj[c0], 1. / 3.);
    double test_filter = pow(sum_vol, 1. / 3.);
    double F[33];
    for (int n = 0; n < 33; n++) F[n] = acc[n] / wsum;
    double L[3][3], Sh[3][3], SSh[3][3];
    L[0][0] = F[6] - F[3] * F[0]; L[0][1] = F[7] - F[3] * F[1]; L[0][2] = F[8] - F[3] * F[2];
    L[1][0] = F[9] - F[4] * F[0]; L[1][1] = F[10] - F[4] * F[1]; L[1][2] = F[11] - F[4] * F[2];
    L[2][0] = F[12] - F[5] * F[0]; L[2][1] = F[13] - F[5] * F[1]; L[2][2] = F[14] - F[5] * F[2];
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) { Sh[a][b] = F[15 + 3 * a + b]; SSh[a][b] = F[24 + 3 * a + b]; }
    double nm, dn;
    germano(L, Sh, SSh, filter, test_filter, ld3(csi, c0), ld3(eta, c0), ld3(zet, c0), nm, dn);
    num_out[c0] = nm;
    den_out[c0] = dn;
}

// Index remap of the DSM box average (les.c:1885-1931) along one axis; returns false when the
// neighbour's weight is set to zero.  `ghost` is set when the remapped point is a DMDA boundary
// ghost whose LM/MM were never written (value 0 in the reference).
__device__ __forceinline__ bool box_remap(int &I, int centre, int m, int ptype, bool jquirk, bool &ghost) {
    if (ptype == TOSCA_PERIODIC_COPY || ptype == TOSCA_PERIODIC_WRAP) {
        if (I == 0) I = m - 2;
        else if (I == m - 1) I = 1;
        return true;
    }
    if (jquirk) {                       // `J==0 || j==my-1` with the centre j (les.c:1911)
        if (I == 0 || centre == m - 1) return false;
        if (I == m - 1) ghost = true;
        return true;
    }
    return !(I == 0 || I == m - 1);
}

// Box average + clip (les.c:1779-1977) for DSM.
__global__ void k_box_clip(const double *LM, const double *MM, const double *aj, double *Cs, double maxCs, Geo g) {
    int i = blockIdx.x * blockDim.x + threadIdx.x + 1, j = blockIdx.y * blockDim.y + threadIdx.y + 1, k = blockIdx.z + 1;
    if (i >= g.mx - 1 || j >= g.my - 1 || k >= g.mz - 1) return;
    double wsum = 0, lsum = 0, msum = 0;
    for (int R = 0; R < 3; R++)
        for (int Q = 0; Q < 3; Q++)
            for (int P = 0; P < 3; P++) {
                int K = k + R - 1, J = j + Q - 1, I = i + P - 1;
                double w = 1. / aj[g.L(K, J, I)];
                bool ghost = false;
                bool ok = box_remap(I, i, g.mx, g.per[0], false, ghost);
                ok = box_remap(J, j, g.my, g.per[1], true, ghost) && ok;
                ok = box_remap(K, k, g.mz, g.per[2], false, ghost) && ok;
                if (!ok) w = 0;
                double sw = (R == 1 ? 4. : 1.) * (Q == 1 ? 4. : 1.) * (P == 1 ? 4. : 1.);
                double l = 0, mm = 0;
                if (!ghost && w != 0) { int c = g.L(K, J, I); l = LM[c]; mm = MM[c]; }
                wsum += sw * w;
                lsum += sw * w * l;
                msum += sw * w * mm;
            }
    double LM_avg = lsum / wsum, MM_avg = msum / wsum;
    double C = 0.5 * LM_avg / (MM_avg + 1e-10);
    Cs[g.L(k, j, i)] = pmin(pmax(C, 0.0), maxCs);
}

// DPASD j-plane sums (les.c:1979-2027): block (jplane, chunk) sums LM, MM, QN, NN over a
// contiguous chunk of the plane's (k, i) cells.  Each thread sums its strided cells, each wave
// reduces its 64 lanes with a butterfly of cross-lane exchanges (no LDS, no barrier), and the
// four wave sums are added in wave order after one barrier; the chunk partials are combined in a
// fixed order by k_dpasd_cs, so the plane means are bitwise reproducible from run to run.
constexpr int PLANE_CHUNKS = 16;
constexpr int PLANE_THREADS = 256;
__global__ void __launch_bounds__(PLANE_THREADS) k_plane_partial(const double *LM, const double *MM, const double *QN,
                                                                 const double *NN, double *part, Geo g) {
    const int j = blockIdx.x + 1, chunk = blockIdx.y;
    const int nk = g.mz - 2, ni = g.mx - 2, n = nk * ni;
    const int per = (n + PLANE_CHUNKS - 1) / PLANE_CHUNKS;
    const int lo = chunk * per, hi = min(n, lo + per);
    double s[4] = {0, 0, 0, 0};
    for (int t = lo + threadIdx.x; t < hi; t += blockDim.x) {
        int k = t / ni + 1, i = t % ni + 1;
        int c = g.L(k, j, i);
        s[0] += LM[c]; s[1] += MM[c]; s[2] += QN[c]; s[3] += NN[c];
    }
    __shared__ double red[4][PLANE_THREADS / 64];
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    for (int q = 0; q < 4; q++) {
        const double w = wave_sum(s[q]);
        if (lane == 0) red[q][wave] = w;
    }
    __syncthreads();
    if (threadIdx.x == 0)
        for (int q = 0; q < 4; q++) {
            double t = 0.0;
            for (int w = 0; w < PLANE_THREADS / 64; w++) t += red[q][w];
            part[((long)j * PLANE_CHUNKS + chunk) * 4 + q] = t;
        }
}

__global__ void k_dpasd_cs(const double *part, double *Cs, double maxCs, Geo g) {
    int i = blockIdx.x * blockDim.x + threadIdx.x + 1, j = blockIdx.y * blockDim.y + threadIdx.y + 1, k = blockIdx.z + 1;
    if (i >= g.mx - 1 || j >= g.my - 1 || k >= g.mz - 1) return;
    double s[4] = {0, 0, 0, 0};
    for (int c = 0; c < PLANE_CHUNKS; c++)
        for (int q = 0; q < 4; q++) s[q] += part[((long)j * PLANE_CHUNKS + c) * 4 + q];
    const double cnt = (double)((g.mz - 2) * (g.mx - 2));
    double J_LM = s[0] / cnt, J_MM = s[1] / cnt, J_QN = s[2] / cnt, J_NN = s[3] / cnt;
    double beta = (J_QN * J_MM) / (J_NN * J_LM + 1e-10);
    double v = 0.5 * (J_LM / (J_MM + 1e-10)) / beta;
    Cs[g.L(k, j, i)] = pmin(pmax(v, 0.0), maxCs);
}

// ---------------------------------------------------------------------------------------------
// UpdateNut
// ---------------------------------------------------------------------------------------------
__global__ void k_nut(const double *ucat, const double *csi, const double *eta, const double *zet,
                      const double *aj, const double *Cs, double *nut, double *ksgs, Geo g) {
    int i = blockIdx.x * blockDim.x + threadIdx.x + 1, j = blockIdx.y * blockDim.y + threadIdx.y + 1, k = blockIdx.z + 1;
    if (i >= g.mx - 1 || j >= g.my - 1 || k >= g.mz - 1) return;
    const int c = g.L(k, j, i);
    V3 cc = ld3(csi, c), ee = ld3(eta, c), zz = ld3(zet, c);
    Grad gr = du_center(g, ucat, i, j, k);
    Cart d = du_dxyz(cc, ee, zz, aj[c], gr);
    double S = strain_mag(d);
    double filter = pow(1. / aj[c], 1. / 3.);
    double nt = Cs[c] * pow(filter, 2.0) * S;
    nut[c] = nt;
    if (ksgs) {
        double Ck = 0.094, V = 1.0 / aj[c];
        double A_x = sqrt(zz.x * zz.x + zz.y * zz.y + zz.z * zz.z);
        double A_y = sqrt(cc.x * cc.x + cc.y * cc.y + cc.z * cc.z);
        double A_z = sqrt(ee.x * ee.x + ee.y * ee.y + ee.z * ee.z);
        double dx = V / A_x, dy = V / A_y, dz = V / A_z;
        double inv = 1.0 / (dx * dx) + 1.0 / (dy * dy) + 1.0 / (dz * dz);
        double d2 = 3.0 / inv;
        ksgs[c] = pow(nt, 2.0) / (pow(Ck, 2.0) * d2);
    }
}

struct NutBC { int bc[6]; double val[6]; };

// UpdateNutBCs (boundary.c:2812-2917) along one axis: ghost cells next to the interior cells of
// both patches of `axis`.  Reads interior values only.
template <class T>
__global__ void k_nut_bc(T *nut, NutBC B, Geo g, int axis) {
    const int m[3] = {g.mx, g.my, g.mz};
    int t1 = (axis == 0) ? 1 : 0, t2 = (axis == 2) ? 1 : 2;
    int a1 = blockIdx.x * blockDim.x + threadIdx.x + 1, a2 = blockIdx.y + 1, side = blockIdx.z;
    if (a1 >= m[t1] - 1 || a2 >= m[t2] - 1) return;
    int patch = 2 * axis + side, n = m[axis];
    int idx[3], src[3];
    idx[t1] = src[t1] = a1;
    idx[t2] = src[t2] = a2;
    idx[axis] = side ? n - 1 : 0;
    int bc = B.bc[patch];
    T v;
    if (bc == TOSCA_BCNUT_ZEROGRADIENT) { src[axis] = side ? n - 2 : 1; v = nut[g.L(src[2], src[1], src[0])]; }
    else if (bc == TOSCA_BCNUT_FIXEDVALUE) v = T(B.val[patch]);
    else if (bc == TOSCA_BCNUT_PERIODIC && g.per[axis] != TOSCA_PERIODIC_NONE) { src[axis] = side ? 1 : n - 2; v = nut[g.L(src[2], src[1], src[0])]; }
    else return;
    nut[g.L(idx[2], idx[1], idx[0])] = v;
}

// ---------------------------------------------------------------------------------------------
// Buoyancy
// ---------------------------------------------------------------------------------------------
__global__ void k_buoyancy(const double *tmprt, const double *icsi, const double *jeta, const double *kzet,
                           double tRef, double scale, double *btheta, double *gcont, Geo g) {
    int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y * blockDim.y + threadIdx.y, k = blockIdx.z;
    if (i >= g.mx || j >= g.my || k >= g.mz) return;
    const int c = g.L(k, j, i), r = g.G(k, j, i);
    const double gx = 0, gy = 0, gz = -9.81;
    double gc[3];
    const bool face_rng = (i < g.mx - 1 && j < g.my - 1 && k < g.mz - 1);
    V3 a = ld3(icsi, c), b = ld3(jeta, c), d = ld3(kzet, c);
    gc[0] = (gx * a.x + gy * a.y + gz * a.z);
    gc[1] = (gx * b.x + gy * b.y + gz * b.z);
    gc[2] = (gx * d.x + gy * d.y + gz * d.z);
    if (gcont && face_rng) {
        if (j > 0 && k > 0) gcont[3 * r] = gc[0];
        if (i > 0 && k > 0) gcont[3 * r + 1] = gc[1];
        if (i > 0 && j > 0) gcont[3 * r + 2] = gc[2];
    }
    double out[3] = {0.0, 0.0, 0.0};
    if (i > 0 && j > 0 && k > 0 && face_rng) {
        double t = tmprt[c];
        double tI = 0.5 * (t + tmprt[g.L(k, j, i + 1)]);
        double tJ = 0.5 * (t + tmprt[g.L(k, j + 1, i)]);
        double tK = 0.5 * (t + tmprt[g.L(k + 1, j, i)]);
        out[0] += scale * (gc[0] * (tRef - tI) / tRef);
        out[1] += scale * (gc[1] * (tRef - tJ) / tRef);
        out[2] += scale * (gc[2] * (tRef - tK) / tRef);
    }
    btheta[3 * r] = out[0]; btheta[3 * r + 1] = out[1]; btheta[3 * r + 2] = out[2];
}

// ---------------------------------------------------------------------------------------------
// host-side launch helpers of this path
// ---------------------------------------------------------------------------------------------
static inline dim3 blk2() { return dim3(64, 4, 1); }
static inline dim3 grd(int ni, int nj, int nk) { return dim3((ni + 63) / 64, (nj + 3) / 4, nk); }

template <class T>
void launch_wrap(T *a, int ncomp, const Geo &g, hipStream_t s) {
    const int m[3] = {g.mx, g.my, g.mz};
    for (int axis = 0; axis < 3; axis++) {
        if (g.per[axis] != TOSCA_PERIODIC_WRAP) continue;
        int t1 = (axis == 0) ? 1 : 0, t2 = (axis == 2) ? 1 : 2;
        int n1 = m[t1] + ((g.per[t1] == TOSCA_PERIODIC_WRAP) ? 2 * GW : 0);
        int n2 = m[t2] + ((g.per[t2] == TOSCA_PERIODIC_WRAP) ? 2 * GW : 0);
        TOSCA_LAUNCH("k_wrap_axis", k_wrap_axis<T>, dim3((n1 + 127) / 128, n2, 2 * GW), dim3(128), 0, s, a, ncomp, g, axis);
    }
}
template void launch_wrap<double>(double *, int, const Geo &, hipStream_t);
template void launch_wrap<float>(float *, int, const Geo &, hipStream_t);

void launch_reset_cell_periodic(double *a, int ncomp, const Geo &g, hipStream_t s) {
    const int m[3] = {g.mx, g.my, g.mz};
    for (int axis = 0; axis < 3; axis++) {
        if (g.per[axis] == TOSCA_PERIODIC_NONE) continue;
        int t1 = (axis == 0) ? 1 : 0, t2 = (axis == 2) ? 1 : 2;
        TOSCA_LAUNCH("k_reset_axis", k_reset_axis, dim3((m[t1] + 127) / 128, m[t2], 2), dim3(128), 0, s, a, ncomp, g, axis);
    }
    launch_wrap(a, ncomp, g, s);
}

// resetFacePeriodicFluxesVector (src/include/inline.h:5357-5451): on each periodic axis the
// points of index 0 take the values at m-2 (type 1; type 2 reads ghost -2, which holds m-2),
// then the ghost scatter; the side-0 launch of k_reset_axis
void launch_reset_face_periodic(double *a, int ncomp, const Geo &g, hipStream_t s) {
    const int m[3] = {g.mx, g.my, g.mz};
    for (int axis = 0; axis < 3; axis++) {
        if (g.per[axis] == TOSCA_PERIODIC_NONE) continue;
        int t1 = (axis == 0) ? 1 : 0, t2 = (axis == 2) ? 1 : 2;
        TOSCA_LAUNCH("k_reset_axis", k_reset_axis, dim3((m[t1] + 127) / 128, m[t2], 1), dim3(128), 0, s, a, ncomp, g, axis);
    }
    launch_wrap(a, ncomp, g, s);
}

void launch_fill(double *a, long n, double v, hipStream_t s) {
    TOSCA_LAUNCH("k_fill", k_fill, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a, n, v);
}

void ref_formu(const Geo &g, const tosca_mesh *m, const tosca_params *p, const double *ucont, const double *ucat,
               const double *nut, double *rhs, double scale, int mode, double *ws, hipStream_t s) {
    const long n = g.nloc();
    FaceArgs A;
    A.ucont = ucont; A.ucat = ucat; A.nut = nut;
    A.icsi = m->icsi; A.ieta = m->ieta; A.izet = m->izet;
    A.jcsi = m->jcsi; A.jeta = m->jeta; A.jzet = m->jzet;
    A.kcsi = m->kcsi; A.keta = m->keta; A.kzet = m->kzet;
    A.iaj = m->iaj; A.jaj = m->jaj; A.kaj = m->kaj;
    A.tauI = p->tau_wall_i; A.tauJ = p->tau_wall_j;
    for (int q = 0; q < 3; q++) { A.div[q] = ws + (long)q * 3 * n; A.visc[q] = ws + (long)(3 + q) * 3 * n; }
    for (int q = 0; q < 6; q++) A.bc_u[q] = p->bc_u[q];
    A.quick = p->div_scheme == TOSCA_DIV_QUICK;
    A.les = p->les_active != 0;
    A.mode = mode;
    A.nu = p->nu;
    double *fp = ws + 18 * n;
    CellArgs C;
    for (int q = 0; q < 3; q++) { C.div[q] = A.div[q]; C.visc[q] = A.visc[q]; }
    TOSCA_LAUNCH("k_formu_faces", k_formu_faces, grd(g.mx - 1, g.my - 1, g.mz - 1), blk2(), 0, s, A, g);
    TOSCA_LAUNCH("k_formu_cells", k_formu_cells, grd(g.mx - 2, g.my - 2, g.mz - 2), blk2(), 0, s, C, fp, mode, g);
    TOSCA_LAUNCH("k_formu_dot", k_formu_dot, grd(g.mx - 2, g.my - 2, g.mz - 2), blk2(), 0, s, fp, m->csi, m->eta, m->zet,
                       m->iaj, m->jaj, m->kaj, rhs, scale, g);
}

void ref_update_cs(const Geo &g, const tosca_mesh *m, const tosca_params *p, const double *ucat, double *cs,
                   double *ws, hipStream_t s) {
    const long n = g.nloc();
    if (p->les_model == TOSCA_LES_SMAGORINSKY) { launch_fill(cs, n, 0.0289, s); return; }
    double *S = ws, *LM = ws + 10 * n, *MM = ws + 11 * n, *QN = ws + 12 * n, *NN = ws + 13 * n;
    double *part = ws + 14 * n;
    dim3 gi = grd(g.mx - 2, g.my - 2, g.mz - 2);
    TOSCA_LAUNCH("k_strain", k_strain, gi, blk2(), 0, s, ucat, m->csi, m->eta, m->zet, m->aj, S, g);
    TOSCA_LAUNCH("k_test_filter3", k_test_filter<3>, gi, blk2(), 0, s, ucat, m->csi, m->eta, m->zet, m->aj, S, LM, MM, g);
    if (p->les_model == TOSCA_LES_DSM) {
        TOSCA_LAUNCH("k_box_clip", k_box_clip, gi, blk2(), 0, s, LM, MM, m->aj, cs, p->max_cs, g);
    } else {
        TOSCA_LAUNCH("k_test_filter5", k_test_filter<5>, gi, blk2(), 0, s, ucat, m->csi, m->eta, m->zet, m->aj, S, QN, NN, g);
        TOSCA_LAUNCH("k_plane_partial", k_plane_partial, dim3(g.my - 2, PLANE_CHUNKS), dim3(PLANE_THREADS), 0, s, LM, MM, QN, NN,
                     part, g);
        TOSCA_LAUNCH("k_dpasd_cs", k_dpasd_cs, gi, blk2(), 0, s, part, cs, p->max_cs, g);
    }
    launch_wrap(cs, 1, g, s);
    launch_reset_cell_periodic(cs, 1, g, s);
}

void ref_update_nut(const Geo &g, const tosca_mesh *m, const tosca_params *p, const double *ucat, const double *cs,
                    double *nut, double *ksgs, hipStream_t s) {
    const long n = g.nloc();
    launch_fill(nut, n, 0.0, s);
    if (ksgs) launch_fill(ksgs, n, 0.0, s);
    TOSCA_LAUNCH("k_nut", k_nut, grd(g.mx - 2, g.my - 2, g.mz - 2), blk2(), 0, s, ucat, m->csi, m->eta, m->zet, m->aj,
                       cs, nut, ksgs, g);
    launch_nut_bcs(nut, p, g, s);
    if (ksgs) launch_wrap(ksgs, 1, g, s);
}

template <class T>
void launch_nut_bcs(T *nut, const tosca_params *p, const Geo &g, hipStream_t s) {
    NutBC B;
    for (int q = 0; q < 6; q++) { B.bc[q] = p->bc_nut[q]; B.val[q] = p->nut_value[q]; }
    const int mm[3] = {g.mx, g.my, g.mz};
    for (int axis = 0; axis < 3; axis++) {
        int t1 = (axis == 0) ? 1 : 0, t2 = (axis == 2) ? 1 : 2;
        TOSCA_LAUNCH("k_nut_bc", k_nut_bc<T>, dim3((mm[t1] + 127) / 128, mm[t2] - 2, 2), dim3(128), 0, s, nut, B, g, axis);
    }
    launch_wrap(nut, 1, g, s);
}
template void launch_nut_bcs<double>(double *, const tosca_params *, const Geo &, hipStream_t);
template void launch_nut_bcs<float>(float *, const tosca_params *, const Geo &, hipStream_t);

void ref_buoyancy(const Geo &g, const tosca_mesh *m, const tosca_params *p, const double *theta, double scale,
                  double *btheta, double *gcont, hipStream_t s) {
    TOSCA_LAUNCH("k_buoyancy", k_buoyancy, grd(g.mx, g.my, g.mz), blk2(), 0, s, theta, m->icsi, m->jeta, m->kzet,
                       p->t_ref, scale, btheta, gcont, g);
}

}  // namespace tosca
