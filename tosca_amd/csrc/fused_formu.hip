// fused_formu.hip -- FormU (formMode 0/1/2) + Buoyancy in one k-marching kernel for gfx950.
//
// The reference evaluates FormU in three sweeps over HBM (face fluxes into six workspace vectors,
// their divergence into lFp, the face-dot into Rhs; src/solvers/ueqn_solvers.c:173-1035) plus a
// separate Buoyancy sweep (src/sources/ueqn_sources.c:2726-2880).  Here one workgroup owns a
// 62 x (NR-2) tile of (i, j) cells and marches along k; per plane it
//   1. computes the i-, j- and k-face fluxes (convective + viscous) of its columns,
//   2. takes the i-face flux of the left neighbour from the lane below (cross-lane shuffle: a
//      wave is one 64-wide row of i), the j-face flux of the row below from LDS and the k-face
//      flux of the plane below from registers, and forms fp (the cell divergence, div and visc
//      differences summed separately as in the reference),
//   3. dots fp of the previous plane with the cell-centred basis for Rhs (+=), and adds the
//      buoyancy source of the same cells into bTheta.
// Face metrics are read exactly once per face, fp and the face fluxes never reach HBM.
// Periodic seams: face 0 of a periodic axis is evaluated as face m-2 and the ghost cell m-1 as
// cell 1, which is what resetFacePeriodicFluxesVector / resetCellPeriodicFluxes make the
// reference read.
#include "fused.h"

namespace tosca {

namespace {

struct Flux { V3 d, v; };

__device__ __forceinline__ V3 add3(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }

// convective (div) and viscous (visc) flux of one face.  dir: 0 = i face (k,j,i)|(k,j,i+1),
// 1 = j face, 2 = k face.  (i,j,k) is the face index after the periodic remap.
template <int DIR>
__device__ __forceinline__ Flux face_flux(const FormUArgs &A, const Geo &g, int i, int j, int k) {
    const int c = g.L(k, j, i);
    const double *u = A.ucat;
    Flux r;
    r.d = V3{0, 0, 0};
    r.v = V3{0, 0, 0};
    const int up = (DIR == 0) ? g.L(k, j, i + 1) : (DIR == 1) ? g.L(k, j + 1, i) : g.L(k + 1, j, i);
    if (A.mode != 2) {
        const double V = A.ucont[3 * c + DIR];
        V3 f;
        if (A.quick) {
            int L, R;
            if (DIR == 0) { face2cell4(g.per[0], i, g.mx, L, R); f = quick(ld3(u, g.L(k, j, L)), ld3(u, c), ld3(u, up), ld3(u, g.L(k, j, R)), V); }
            else if (DIR == 1) { face2cell4(g.per[1], j, g.my, L, R); f = quick(ld3(u, g.L(k, L, i)), ld3(u, c), ld3(u, up), ld3(u, g.L(k, R, i)), V); }
            else { face2cell4(g.per[2], k, g.mz, L, R); f = quick(ld3(u, g.L(L, j, i)), ld3(u, c), ld3(u, up), ld3(u, g.L(R, j, i)), V); }
        } else {
            V3 a = ld3(u, c), b = ld3(u, up);
            f = V3{0.5 * (a.x + b.x), 0.5 * (a.y + b.y), 0.5 * (a.z + b.z)};
        }
        r.d = V3{-V * f.x, -V * f.y, -V * f.z};
    }
    if (A.mode != 1) {
        Grad gr = (DIR == 0) ? du_i(g, u, i, j, k) : (DIR == 1) ? du_j(g, u, i, j, k) : du_k(g, u, i, j, k);
        double nuEff = A.nu;
        V3 base{0, 0, 0};
        if (A.les) {
            const double nt = 0.5 * (A.nut[c] + A.nut[up]);
            const int f = (DIR == 0) ? i : (DIR == 1) ? j : k;
            const int m = (DIR == 0) ? g.mx : (DIR == 1) ? g.my : g.mz;
            const int bl = A.bc_u[2 * DIR], br = A.bc_u[2 * DIR + 1];
            if (DIR < 2 && ((bl == TOSCA_BCU_WALLFUNCTION && f == 0) || (br == TOSCA_BCU_WALLFUNCTION && f == m - 2))) {
                if (DIR == 0) { int t = k * g.my + j, P = g.mz * g.my; base = V3{-A.tauI[t], -A.tauI[P + t], -A.tauI[2 * P + t]}; }
                else { int t = k * g.mx + i, P = g.mz * g.mx; base = V3{-A.tauJ[t], -A.tauJ[P + t], -A.tauJ[2 * P + t]}; }
                nuEff = A.nu;
            } else if ((bl == TOSCA_BCU_SLIP && f == 0) || (br == TOSCA_BCU_SLIP && f == m - 2)) {
                nuEff = 0.0;
            } else {
                nuEff = A.nu + nt;
            }
        }
        const double *fc = (DIR == 0) ? A.icsi : (DIR == 1) ? A.jcsi : A.kcsi;
        const double *fe = (DIR == 0) ? A.ieta : (DIR == 1) ? A.jeta : A.keta;
        const double *fz = (DIR == 0) ? A.izet : (DIR == 1) ? A.jzet : A.kzet;
        const double *fa = (DIR == 0) ? A.iaj : (DIR == 1) ? A.jaj : A.kaj;
        r.v = add3(base, visc_flux(gr, ld3(fc, c), ld3(fe, c), ld3(fz, c), DIR, fa[c], nuEff));
    }
    return r;
}

__device__ __forceinline__ double shfl_up1(double x) { return __shfl_up(x, 1, 64); }
__device__ __forceinline__ double shfl_dn1(double x) { return __shfl_down(x, 1, 64); }
__device__ __forceinline__ V3 shfl_up3(V3 a) { return V3{shfl_up1(a.x), shfl_up1(a.y), shfl_up1(a.z)}; }
__device__ __forceinline__ V3 shfl_dn3(V3 a) { return V3{shfl_dn1(a.x), shfl_dn1(a.y), shfl_dn1(a.z)}; }

// fp differences in the reference's order: ((((a - b) + c) - d) + e) - f, div and visc apart
__device__ __forceinline__ double fpsum(double d1, double d1m, double d2, double d2m, double d3, double d3m,
                                        double v1, double v1m, double v2, double v2m, double v3, double v3m, int mode) {
    double f = 0.0;
    if (mode != 2) f += (d1 - d1m + d2 - d2m + d3 - d3m);
    if (mode != 1) f += (v1 - v1m + v2 - v2m + v3 - v3m);
    return f;
}

}  // namespace

template <int NR>
__global__ void __launch_bounds__(64 * NR) k_fused_formu(FormUArgs A, Geo g, int kchunk) {
    constexpr int TI = 62, TJ = NR - 2;
    __shared__ double sF2[6][NR][64];     // j-face fluxes of the plane (div xyz, visc xyz)
    __shared__ double sFp[3][NR][64];     // fp of the previous plane

    const int lane = threadIdx.x, r = threadIdx.y;
    const int i0 = 1 + blockIdx.x * TI, j0 = 1 + blockIdx.y * TJ;
    const int kb = 1 + blockIdx.z * kchunk;
    const int ke = min(kb + kchunk, g.mz - 1);      // output planes [kb, ke)
    const int iraw = i0 - 1 + lane, jraw = j0 - 1 + r;
    const bool iper = g.per[0] != TOSCA_PERIODIC_NONE, jper = g.per[1] != TOSCA_PERIODIC_NONE,
               kper = g.per[2] != TOSCA_PERIODIC_NONE;
    // column this thread evaluates (periodic ghost m-1 acts as cell 1)
    const int ic = (iraw == g.mx - 1 && iper) ? 1 : iraw;
    const int jc = (jraw == g.my - 1 && jper) ? 1 : jraw;
    const bool colI = ic >= 1 && ic <= g.mx - 2, colJ = jc >= 1 && jc <= g.my - 2;
    // i-face of this lane (face 0 of a periodic axis is face m-2)
    const int fi = (ic == 0 && iper) ? g.mx - 2 : ic;
    const bool doF1 = r >= 1 && colJ && fi >= 0 && fi <= g.mx - 2;
    const int fj = (jc == 0 && jper) ? g.my - 2 : jc;
    const bool doF2 = lane >= 1 && colI && fj >= 0 && fj <= g.my - 2;
    const bool cell = lane >= 1 && r >= 1 && colI && colJ;           // fp of (ic, jc) is formed
    const bool out = cell && lane <= TI && r <= TJ && iraw <= g.mx - 2 && jraw <= g.my - 2;
    // + side neighbours of an output cell on a non-periodic boundary read their own fp
    const bool ownI = (iraw == g.mx - 2) && !iper, ownJ = (jraw == g.my - 2) && !jper;

    Flux F3p{V3{0, 0, 0}, V3{0, 0, 0}};
    if (cell) {
        const int kf = (kb - 1 == 0 && kper) ? g.mz - 2 : kb - 1;
        F3p = face_flux<2>(A, g, ic, jc, kf);
    }
    V3 fpPrev{0, 0, 0};
    for (int t = kb; t <= ke; t++) {
        // plane evaluated at this step: the periodic ghost plane m-1 is plane 1
        const bool ghostK = (t == g.mz - 1);
        const int kc = (ghostK && kper) ? 1 : t;
        const bool evalPlane = !(ghostK && !kper);
        Flux F1{V3{0, 0, 0}, V3{0, 0, 0}}, F2{V3{0, 0, 0}, V3{0, 0, 0}}, F3{V3{0, 0, 0}, V3{0, 0, 0}};
        if (evalPlane) {
            if (doF1) F1 = face_flux<0>(A, g, fi, jc, kc);
            if (doF2) F2 = face_flux<1>(A, g, ic, fj, kc);
            if (cell) F3 = face_flux<2>(A, g, ic, jc, kc);
        }
        // exchange: F1 of the lane below by shuffle (every lane takes part), F2 of the row below by LDS
        V3 F1dm = shfl_up3(F1.d), F1vm = shfl_up3(F1.v);
        sF2[0][r][lane] = F2.d.x; sF2[1][r][lane] = F2.d.y; sF2[2][r][lane] = F2.d.z;
        sF2[3][r][lane] = F2.v.x; sF2[4][r][lane] = F2.v.y; sF2[5][r][lane] = F2.v.z;
        __syncthreads();
        V3 fp{0, 0, 0};
        if (cell) {
            if (evalPlane) {
                const int rm = r - 1;
                V3 F2dm{sF2[0][rm][lane], sF2[1][rm][lane], sF2[2][rm][lane]};
                V3 F2vm{sF2[3][rm][lane], sF2[4][rm][lane], sF2[5][rm][lane]};
                fp.x = fpsum(F1.d.x, F1dm.x, F2.d.x, F2dm.x, F3.d.x, F3p.d.x, F1.v.x, F1vm.x, F2.v.x, F2vm.x, F3.v.x, F3p.v.x, A.mode);
                fp.y = fpsum(F1.d.y, F1dm.y, F2.d.y, F2dm.y, F3.d.y, F3p.d.y, F1.v.y, F1vm.y, F2.v.y, F2vm.y, F3.v.y, F3p.v.y, A.mode);
                fp.z = fpsum(F1.d.z, F1dm.z, F2.d.z, F2dm.z, F3.d.z, F3p.d.z, F1.v.z, F1vm.z, F2.v.z, F2vm.z, F3.v.z, F3p.v.z, A.mode);
            } else {
                fp = fpPrev;                      // non-periodic top: ghost fp copies cell m-2
            }
        }
        // Rhs of the previous plane (t-1): fp(t-1) own / i+1 (shuffle) / j+1 (LDS), fp(t) own
        V3 fpI = shfl_dn3(fpPrev);
        if (t > kb && out) {
            const int k = t - 1;
            const int c = g.L(k, jraw, iraw);
            V3 fJ{sFp[0][r + 1][lane], sFp[1][r + 1][lane], sFp[2][r + 1][lane]};
            if (ownI) fpI = fpPrev;
            if (ownJ) fJ = fpPrev;
            const V3 fK = fp;
            const int ro = g.G(k, jraw, iraw);
            V3 a = ld3(A.csi, c), b = ld3(A.csi, g.L(k, jraw, iraw + 1));
            double v1 = a.x * fpPrev.x + a.y * fpPrev.y + a.z * fpPrev.z;
            double v2 = b.x * fpI.x + b.y * fpI.y + b.z * fpI.z;
            double rx = A.scale * A.iaj[c] * (0.5 * (v1 + v2));
            a = ld3(A.eta, c); b = ld3(A.eta, g.L(k, jraw + 1, iraw));
            v1 = a.x * fpPrev.x + a.y * fpPrev.y + a.z * fpPrev.z;
            v2 = b.x * fJ.x + b.y * fJ.y + b.z * fJ.z;
            double ry = A.scale * A.jaj[c] * (0.5 * (v1 + v2));
            a = ld3(A.zet, c); b = ld3(A.zet, g.L(k + 1, jraw, iraw));
            v1 = a.x * fpPrev.x + a.y * fpPrev.y + a.z * fpPrev.z;
            v2 = b.x * fK.x + b.y * fK.y + b.z * fK.z;
            double rz = A.scale * A.kaj[c] * (0.5 * (v1 + v2));
            A.rhs[3 * ro] += rx;
            A.rhs[3 * ro + 1] += ry;
            A.rhs[3 * ro + 2] += rz;
            if (A.btheta) {
                // Boussinesq source on the three + faces of the cell (ueqn_sources.c:2812-2870):
                // g . S with g = (0, 0, -9.81) is -9.81 * S.z exactly
                const double th = A.theta[c], tr = A.tRef;
                const double tI = 0.5 * (th + A.theta[g.L(k, jraw, iraw + 1)]);
                const double tJ = 0.5 * (th + A.theta[g.L(k, jraw + 1, iraw)]);
                const double tK = 0.5 * (th + A.theta[g.L(k + 1, jraw, iraw)]);
                const double gx = -9.81 * A.icsi[3 * c + 2], gy = -9.81 * A.jeta[3 * c + 2], gz = -9.81 * A.kzet[3 * c + 2];
                A.btheta[3 * ro] = A.bscale * (gx * (tr - tI) / tr);
                A.btheta[3 * ro + 1] = A.bscale * (gy * (tr - tJ) / tr);
                A.btheta[3 * ro + 2] = A.bscale * (gz * (tr - tK) / tr);
            }
        }
        __syncthreads();
        sFp[0][r][lane] = fp.x; sFp[1][r][lane] = fp.y; sFp[2][r][lane] = fp.z;
        fpPrev = fp;
        F3p = F3;
    }
}

// zero bTheta on the boundary shell of the owned range (the reference zeroes the whole vector)
__global__ void k_btheta_shell(double *bt, Geo g) {
    int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, k = blockIdx.z;
    if (i >= g.mx) return;
    bool shell = i == 0 || i == g.mx - 1 || j == 0 || j == g.my - 1 || k == 0 || k == g.mz - 1;
    if (!shell) return;
    int r = g.G(k, j, i);
    bt[3 * r] = 0.0; bt[3 * r + 1] = 0.0; bt[3 * r + 2] = 0.0;
}

void fused_formu(const Geo &g, const FormUArgs &A, hipStream_t s) {
    constexpr int NR = 8;
    const int TI = 62, TJ = NR - 2;
    const int nk = g.mz - 2;
    // k chunks: enough workgroups to fill 256 CUs several times over
    const int tiles = ((g.mx - 2 + TI - 1) / TI) * ((g.my - 2 + TJ - 1) / TJ);
    int chunks = (2048 + tiles - 1) / tiles;
    chunks = max(1, min(chunks, nk / 8 > 0 ? nk / 8 : 1));
    const int kchunk = (nk + chunks - 1) / chunks;
    chunks = (nk + kchunk - 1) / kchunk;
    dim3 grid((g.mx - 2 + TI - 1) / TI, (g.my - 2 + TJ - 1) / TJ, chunks);
    TOSCA_LAUNCH("k_fused_formu", k_fused_formu<NR>, grid, dim3(64, NR), 0, s, A, g, kchunk);
    if (A.btheta)
        TOSCA_LAUNCH("k_btheta_shell", k_btheta_shell, dim3((g.mx + 63) / 64, g.my, g.mz), dim3(64), 0, s, A.btheta, g);
}

}  // namespace tosca
