// fused_les.hip -- dynamic Smagorinsky on gfx950 in two kernels.
//
// The reference's UpdateCs (src/les.c:197-1977) makes four HBM sweeps with nine workspace
// vectors: strain rows lSx/lSy/lSz/lS, the 27-point test filter into lLM/lMM, ghost scatters,
// and the 27-point box average into lCs; UpdateNut (src/les.c:2152-2514) then recomputes |S|.
//   k_dsm_lm: one workgroup per 16x8x4 tile of cells.  The strain, |S| and the contravariant
//     velocity of the tile plus a one-cell halo (18x10x6 points) are evaluated once into LDS;
//     each thread then gathers its 27 neighbours from LDS for the masked Simpson test filter of
//     the 27 Germano quantities and contracts them (les.c:466-801) into LM, MM.  |S| of the
//     cell is kept for the eddy viscosity.
//   k_dsm_cs_nut: 3x3x3 volume-weighted Simpson box average of LM/MM with the reference's
//     periodic remap and boundary-ghost weights (les.c:1885-1950), clip to [0, maxCs], and
//     nu_t = Cs Delta^2 |S| (les.c:2510-2514) -- Cs and nu_t written in one pass.
#include "fused.h"

namespace tosca {

namespace {
constexpr int TX = 16, TY = 8, TZ = 4;
constexpr int RX = TX + 2, RY = TY + 2, RZ = TZ + 2, RN = RX * RY * RZ;   // tile + 1-cell halo
constexpr int NQ = 14;   // u v w U V W Sxx Sxy Sxz Syy Syz Szz |S| 1/aj

__device__ __forceinline__ bool interior_cell(const Geo &g, int i, int j, int k) {
    return i >= 1 && i <= g.mx - 2 && j >= 1 && j <= g.my - 2 && k >= 1 && k <= g.mz - 2;
}
}  // namespace

__global__ void __launch_bounds__(TX * TY * TZ) k_dsm_lm(Geo g, const double *ucat, const double *csi,
                                                         const double *eta, const double *zet, const double *aj,
                                                         double *LM, double *MM, double *Sabs) {
    __shared__ double P[NQ][RN];
    const int tid = threadIdx.x + TX * (threadIdx.y + TY * threadIdx.z);
    const int i0 = 1 + blockIdx.x * TX, j0 = 1 + blockIdx.y * TY, k0 = 1 + blockIdx.z * TZ;

    // phase 1: point quantities of the tile + halo (interior cells only; others are masked)
    for (int idx = tid; idx < RN; idx += TX * TY * TZ) {
        const int ri = idx % RX, rj = (idx / RX) % RY, rk = idx / (RX * RY);
        const int I = i0 - 1 + ri, J = j0 - 1 + rj, K = k0 - 1 + rk;
        double q[NQ];
        for (int n = 0; n < NQ; n++) q[n] = 0.0;
        if (interior_cell(g, I, J, K)) {
            const int c = g.L(K, J, I);
            const V3 uu = ld3(ucat, c), cc = ld3(csi, c), ee = ld3(eta, c), zz = ld3(zet, c);
            const double a = aj[c];
            const Cart d = du_dxyz(cc, ee, zz, a, du_center(g, ucat, I, J, K));
            q[0] = uu.x; q[1] = uu.y; q[2] = uu.z;
            q[3] = uu.x * cc.x + uu.y * cc.y + uu.z * cc.z;
            q[4] = uu.x * ee.x + uu.y * ee.y + uu.z * ee.z;
            q[5] = uu.x * zz.x + uu.y * zz.y + uu.z * zz.z;
            q[6] = 0.5 * (d.du_dx + d.du_dx); q[7] = 0.5 * (d.du_dy + d.dv_dx); q[8] = 0.5 * (d.du_dz + d.dw_dx);
            q[9] = 0.5 * (d.dv_dy + d.dv_dy); q[10] = 0.5 * (d.dv_dz + d.dw_dy); q[11] = 0.5 * (d.dw_dz + d.dw_dz);
            q[12] = strain_mag(d);
            q[13] = 1. / a;
        }
        for (int n = 0; n < NQ; n++) P[n][idx] = q[n];
    }
    __syncthreads();

    // phase 2: masked Simpson test filter (integrateTestfilterSimpson, inline.h:3592-3611)
    const int i = i0 + threadIdx.x, j = j0 + threadIdx.y, k = k0 + threadIdx.z;
    if (!interior_cell(g, i, j, k)) return;
    const int ctr = ((threadIdx.z + 1) * RY + (threadIdx.y + 1)) * RX + (threadIdx.x + 1);
    // sum_vol in the reference's loop order (p over i, q over j, r over k; les.c:618-655)
    double sum_vol = 0.0;
    for (int p = 0; p < 3; p++)
        for (int qq = 0; qq < 3; qq++)
            for (int r = 0; r < 3; r++) {
                if (!interior_cell(g, i - 1 + p, j - 1 + qq, k - 1 + r)) continue;
                const double cf = (p == 1 ? 0.5 : 0.25) * (qq == 1 ? 0.5 : 0.25) * (r == 1 ? 0.5 : 0.25);
                const int n = ((threadIdx.z + r) * RY + (threadIdx.y + qq)) * RX + (threadIdx.x + p);
                sum_vol += P[13][n] * 8.0 * cf;
            }
    double acc[27];
    for (int n = 0; n < 27; n++) acc[n] = 0.0;
    double wsum = 0.0;
    for (int R = 0; R < 3; R++)
        for (int Q = 0; Q < 3; Q++)
            for (int Pp = 0; Pp < 3; Pp++) {
                if (!interior_cell(g, i - 1 + Pp, j - 1 + Q, k - 1 + R)) continue;
                const double sw = (R == 1 ? 4. : 1.) * (Q == 1 ? 4. : 1.) * (Pp == 1 ? 4. : 1.);
                const int n = ((threadIdx.z + R) * RY + (threadIdx.y + Q)) * RX + (threadIdx.x + Pp);
                const double u = P[0][n], v = P[1][n], w = P[2][n], U = P[3][n], V = P[4][n], W = P[5][n];
                const double Sm = P[12][n];
                wsum += sw;
                acc[0] += sw * u; acc[1] += sw * v; acc[2] += sw * w;
                acc[3] += sw * U; acc[4] += sw * V; acc[5] += sw * W;
                acc[6] += sw * (U * u); acc[7] += sw * (U * v); acc[8] += sw * (U * w);
                acc[9] += sw * (V * u); acc[10] += sw * (V * v); acc[11] += sw * (V * w);
                acc[12] += sw * (W * u); acc[13] += sw * (W * v); acc[14] += sw * (W * w);
#pragma unroll
                for (int s = 0; s < 6; s++) {
                    const double S = P[6 + s][n];
                    acc[15 + s] += sw * S;
                    acc[21 + s] += sw * (S * Sm);
                }
            }
    double F[27];
    for (int n = 0; n < 27; n++) F[n] = acc[n] / wsum;
    double L[3][3], Sh[3][3], SSh[3][3];
    L[0][0] = F[6] - F[3] * F[0]; L[0][1] = F[7] - F[3] * F[1]; L[0][2] = F[8] - F[3] * F[2];
    L[1][0] = F[9] - F[4] * F[0]; L[1][1] = F[10] - F[4] * F[1]; L[1][2] = F[11] - F[4] * F[2];
    L[2][0] = F[12] - F[5] * F[0]; L[2][1] = F[13] - F[5] * F[1]; L[2][2] = F[14] - F[5] * F[2];
    // symmetric strain: 15..20 = xx xy xz yy yz zz
    Sh[0][0] = F[15]; Sh[0][1] = Sh[1][0] = F[16]; Sh[0][2] = Sh[2][0] = F[17];
    Sh[1][1] = F[18]; Sh[1][2] = Sh[2][1] = F[19]; Sh[2][2] = F[20];
    SSh[0][0] = F[21]; SSh[0][1] = SSh[1][0] = F[22]; SSh[0][2] = SSh[2][0] = F[23];
    SSh[1][1] = F[24]; SSh[1][2] = SSh[2][1] = F[25]; SSh[2][2] = F[26];
    const int c = g.L(k, j, i);
    const double filter = pow(1. / aj[c], 1. / 3.);
    const double test_filter = pow(sum_vol, 1. / 3.);
    double nm, dn;
    germano(L, Sh, SSh, filter, test_filter, ld3(csi, c), ld3(eta, c), ld3(zet, c), nm, dn);
    LM[c] = nm;
    MM[c] = dn;
    Sabs[c] = P[12][ctr];
}

namespace {
// box-average index remap (les.c:1885-1931), see box_remap in ref_kernels.hip
__device__ __forceinline__ bool remap_axis(int &I, int centre, int m, int ptype, bool jquirk, bool &ghost) {
    if (ptype != TOSCA_PERIODIC_NONE) {
        if (I == 0) I = m - 2;
        else if (I == m - 1) I = 1;
        return true;
    }
    if (jquirk) {
        if (I == 0 || centre == m - 1) return false;
        if (I == m - 1) ghost = true;
        return true;
    }
    return !(I == 0 || I == m - 1);
}
}  // namespace

__global__ void __launch_bounds__(256) k_dsm_cs_nut(Geo g, const double *LM, const double *MM, const double *Sabs, const double *aj,
                             double maxCs, double *Cs, double *nut) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1, j = blockIdx.y * blockDim.y + threadIdx.y + 1,
              k = blockIdx.z + 1;
    if (i > g.mx - 2 || j > g.my - 2 || k > g.mz - 2) return;
    double wsum = 0, lsum = 0, msum = 0;
    for (int R = 0; R < 3; R++)
        for (int Q = 0; Q < 3; Q++)
            for (int P = 0; P < 3; P++) {
                int K = k + R - 1, J = j + Q - 1, I = i + P - 1;
                double w = 1. / aj[g.L(K, J, I)];
                bool ghost = false;
                bool ok = remap_axis(I, i, g.mx, g.per[0], false, ghost);
                ok = remap_axis(J, j, g.my, g.per[1], true, ghost) && ok;
                ok = remap_axis(K, k, g.mz, g.per[2], false, ghost) && ok;
                if (!ok) w = 0;
                const double sw = (R == 1 ? 4. : 1.) * (Q == 1 ? 4. : 1.) * (P == 1 ? 4. : 1.);
                double l = 0, mm = 0;
                if (!ghost && w != 0) { const int c = g.L(K, J, I); l = LM[c]; mm = MM[c]; }
                wsum += sw * w;
                lsum += sw * w * l;
                msum += sw * w * mm;
            }
    const double C = 0.5 * (lsum / wsum) / ((msum / wsum) + 1e-10);
    const int c = g.L(k, j, i);
    const double cs = pmin(pmax(C, 0.0), maxCs);
    Cs[c] = cs;
    if (nut) {
        const double filter = pow(1. / aj[c], 1. / 3.);
        nut[c] = cs * pow(filter, 2.0) * Sabs[c];
    }
}

__global__ void __launch_bounds__(256) k_nut_fused(Geo g, const double *ucat, const double *csi, const double *eta, const double *zet,
                            const double *aj, const double *Cs, double *nut, double *ksgs) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1, j = blockIdx.y * blockDim.y + threadIdx.y + 1,
              k = blockIdx.z + 1;
    if (i > g.mx - 2 || j > g.my - 2 || k > g.mz - 2) return;
    const int c = g.L(k, j, i);
    const V3 cc = ld3(csi, c), ee = ld3(eta, c), zz = ld3(zet, c);
    const double a = aj[c];
    const double S = strain_mag(du_dxyz(cc, ee, zz, a, du_center(g, ucat, i, j, k)));
    const double filter = pow(1. / a, 1. / 3.);
    const double nt = Cs[c] * pow(filter, 2.0) * S;
    nut[c] = nt;
    if (ksgs) {
        const double V = 1.0 / a;
        const double dx = V / sqrt(zz.x * zz.x + zz.y * zz.y + zz.z * zz.z);
        const double dy = V / sqrt(cc.x * cc.x + cc.y * cc.y + cc.z * cc.z);
        const double dz = V / sqrt(ee.x * ee.x + ee.y * ee.y + ee.z * ee.z);
        const double d2 = 3.0 / (1.0 / (dx * dx) + 1.0 / (dy * dy) + 1.0 / (dz * dz));
        ksgs[c] = pow(nt, 2.0) / (pow(0.094, 2.0) * d2);
    }
}

// zero an ni x nj x nk box of an array with row length nx and plane size nx*ny (ncomp values per point)
__global__ void k_zero_box(double *a, int ncomp, int nx, int ny, int i0, int j0, int k0, int ni) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ni * ncomp) return;
    const long p = ((long)(k0 + blockIdx.z) * ny + (j0 + blockIdx.y)) * nx + i0;
    a[p * ncomp + t] = 0.0;
}

void zero_shell(double *a, int ncomp, int nx, int ny, int nz, int lo, int hi, hipStream_t s) {
    // boxes (in array coordinates) covering everything but [lo+1 .. n-hi-2] on every axis:
    // `lo` / `hi` extra layers outside the boundary cells on the low / high side
    const int w0 = lo + 1, w1 = hi + 1;            // shell thickness per side
    auto box = [&](int i0, int j0, int k0, int ni, int nj, int nk) {
        if (ni <= 0 || nj <= 0 || nk <= 0) return;
        TOSCA_LAUNCH("k_zero_box", k_zero_box, dim3((ni * ncomp + 255) / 256, nj, nk), dim3(256), 0, s, a, ncomp, nx, ny,
                     i0, j0, k0, ni);
    };
    box(0, 0, 0, nx, ny, w0);                                   // low k slab
    box(0, 0, nz - w1, nx, ny, w1);                             // high k slab
    box(0, 0, w0, nx, w0, nz - w0 - w1);                        // low j slab
    box(0, ny - w1, w0, nx, w1, nz - w0 - w1);                  // high j slab
    box(0, w0, w0, w0, ny - w0 - w1, nz - w0 - w1);             // low i slab
    box(nx - w1, w0, w0, w1, ny - w0 - w1, nz - w0 - w1);       // high i slab
}

void zero_outside_interior(double *a, const Geo &g, hipStream_t s) {
    zero_shell(a, 1, g.gxm, g.gym, g.gzm, GW, GW, s);
}

void fused_dsm(const Geo &g, const tosca_mesh *m, const double *ucat, double maxCs, double *cs, double *nut,
               double *ws, hipStream_t s) {
    const long n = g.nloc();
    double *LM = ws, *MM = ws + n, *Sabs = ws + 2 * n;
    dim3 grid((g.mx - 2 + TX - 1) / TX, (g.my - 2 + TY - 1) / TY, (g.mz - 2 + TZ - 1) / TZ);
    TOSCA_LAUNCH("k_dsm_lm", k_dsm_lm, grid, dim3(TX, TY, TZ), 0, s, g, ucat, m->csi, m->eta, m->zet, m->aj, LM, MM,
                 Sabs);
    TOSCA_LAUNCH("k_dsm_cs_nut", k_dsm_cs_nut, dim3((g.mx - 2 + 63) / 64, (g.my - 2 + 3) / 4, g.mz - 2), dim3(64, 4),
                 0, s, g, LM, MM, Sabs, m->aj, maxCs, cs, nut);
}

void fused_nut(const Geo &g, const tosca_mesh *m, const double *ucat, const double *cs, double *nut, double *ksgs,
               hipStream_t s) {
    TOSCA_LAUNCH("k_nut_fused", k_nut_fused, dim3((g.mx - 2 + 63) / 64, (g.my - 2 + 3) / 4, g.mz - 2), dim3(64, 4), 0,
                 s, g, ucat, m->csi, m->eta, m->zet, m->aj, cs, nut, ksgs);
}

void nut_bcs(double *nut, const tosca_params *p, const Geo &g, hipStream_t s) { launch_nut_bcs(nut, p, g, s); }

}  // namespace tosca
