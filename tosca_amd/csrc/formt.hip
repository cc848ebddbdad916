// formt.hip -- FormT (src/solvers/teqn_solvers.c:6-592), the divergence and viscous fluxes of the
// potential-temperature equation and their cell budgets, on gfx950.  Compiled with FP contraction
// off: every face flux is evaluated in the reference's expression order.
//
//   k_formt_faces   the fused i/j/k face loop (teqn_solvers.c:117-481): WENO3 upwinded advective
//                   flux and the diffusive flux with the effective diffusivity of the LES closure
//                   (nu/Pr, the subgrid diffusivity k_t, or nu_t over the stability-corrected
//                   turbulent Prandtl number) and the wall heat flux of thetaWallFunction patches
//   (ghosts)        DMLocalToLocal and resetFacePeriodicFluxesVector on both flux arrays (:487-503)
//   k_formt_cells   the flux budgets of the interior cells times the cell Jacobian (:511-537)
//
// Flux arrays are two local vector fields in the context workspace; the faces kernel covers the
// whole ghosted range and writes zero where the reference leaves its VecSet(0) value, so no
// separate clear is needed.
#include "kernels.h"

namespace tosca {

namespace {

struct FormTArgs {
    const double *ucont, *t, *nut, *kt;
    const double *icsi, *ieta, *izet, *jcsi, *jeta, *jzet, *kcsi, *keta, *kzet;
    const double *iaj, *jaj, *kaj;
    const double *qI, *qJ;       // wall heat flux [3][mz][my] (i patches), [3][mz][mx] (j patches)
    double *div, *visc;          // local vector fields
    double nu, pr, tref;
    int les, ktm, wall[4];
};

// Compute_dscalar_i / _j / _k (src/include/inline.h:3210-3342) with no IBM or overset cells: the
// difference across the face and the two four-point transverse averages, which become two-point
// one-sided averages on the face rows next to a non-periodic boundary edge.
__device__ __forceinline__ void dscalar_i(const Geo &g, const double *K, int i, int j, int k, double &dc, double &de,
                                          double &dz) {
    const bool iN = g.nonper(0), jN = g.nonper(1), kN = g.nonper(2);
    const bool ib = (i == 0 || i == g.mx - 2);
    dc = K[g.L(k, j, i + 1)] - K[g.L(k, j, i)];
    if (j == g.my - 2 && ib && (jN || iN))
        de = (K[g.L(k, j, i + 1)] + K[g.L(k, j, i)] - K[g.L(k, j - 1, i + 1)] - K[g.L(k, j - 1, i)]) * 0.5;
    else if (j == 1 && ib && (jN || iN))
        de = (K[g.L(k, j + 1, i + 1)] + K[g.L(k, j + 1, i)] - K[g.L(k, j, i + 1)] - K[g.L(k, j, i)]) * 0.5;
    else
        de = (K[g.L(k, j + 1, i + 1)] + K[g.L(k, j + 1, i)] - K[g.L(k, j - 1, i + 1)] - K[g.L(k, j - 1, i)]) * 0.25;
    if (k == g.mz - 2 && ib && (kN || iN))
        dz = (K[g.L(k, j, i + 1)] + K[g.L(k, j, i)] - K[g.L(k - 1, j, i + 1)] - K[g.L(k - 1, j, i)]) * 0.5;
    else if (k == 1 && ib && (kN || iN))
        dz = (K[g.L(k + 1, j, i + 1)] + K[g.L(k + 1, j, i)] - K[g.L(k, j, i + 1)] - K[g.L(k, j, i)]) * 0.5;
    else
        dz = (K[g.L(k + 1, j, i + 1)] + K[g.L(k + 1, j, i)] - K[g.L(k - 1, j, i + 1)] - K[g.L(k - 1, j, i)]) * 0.25;
}

__device__ __forceinline__ void dscalar_j(const Geo &g, const double *K, int i, int j, int k, double &dc, double &de,
                                          double &dz) {
    const bool iN = g.nonper(0), jN = g.nonper(1), kN = g.nonper(2);
    const bool jb = (j == 0 || j == g.my - 2);
    if (i == g.mx - 2 && jb && (iN || jN))
        dc = (K[g.L(k, j + 1, i)] + K[g.L(k, j, i)] - K[g.L(k, j + 1, i - 1)] - K[g.L(k, j, i - 1)]) * 0.5;
    else if (i == 1 && jb && (iN || jN))
        dc = (K[g.L(k, j + 1, i + 1)] + K[g.L(k, j, i + 1)] - K[g.L(k, j + 1, i)] - K[g.L(k, j, i)]) * 0.5;
    else
        dc = (K[g.L(k, j + 1, i + 1)] + K[g.L(k, j, i + 1)] - K[g.L(k, j + 1, i - 1)] - K[g.L(k, j, i - 1)]) * 0.25;
    de = K[g.L(k, j + 1, i)] - K[g.L(k, j, i)];
    if (k == g.mz - 2 && jb && (kN || jN))
        dz = (K[g.L(k, j + 1, i)] + K[g.L(k, j, i)] - K[g.L(k - 1, j + 1, i)] - K[g.L(k - 1, j, i)]) * 0.5;
    else if (k == 1 && jb && (kN || jN))
        dz = (K[g.L(k + 1, j + 1, i)] + K[g.L(k + 1, j, i)] - K[g.L(k, j + 1, i)] - K[g.L(k, j, i)]) * 0.5;
    else
        dz = (K[g.L(k + 1, j + 1, i)] + K[g.L(k + 1, j, i)] - K[g.L(k - 1, j + 1, i)] - K[g.L(k - 1, j, i)]) * 0.25;
}

__device__ __forceinline__ void dscalar_k(const Geo &g, const double *K, int i, int j, int k, double &dc, double &de,
                                          double &dz) {
    const bool iN = g.nonper(0), jN = g.nonper(1), kN = g.nonper(2);
    const bool kb = (k == 0 || k == g.mz - 2);
    if (i == g.mx - 2 && kb && (iN || kN))
        dc = (K[g.L(k + 1, j, i)] + K[g.L(k, j, i)] - K[g.L(k + 1, j, i - 1)] - K[g.L(k, j, i - 1)]) * 0.5;
    else if (i == 1 && kb && (iN || kN))
        dc = (K[g.L(k + 1, j, i + 1)] + K[g.L(k, j, i + 1)] - K[g.L(k + 1, j, i)] - K[g.L(k, j, i)]) * 0.5;
    else
        dc = (K[g.L(k + 1, j, i + 1)] + K[g.L(k, j, i + 1)] - K[g.L(k + 1, j, i - 1)] - K[g.L(k, j, i - 1)]) * 0.25;
    if (j == g.my - 2 && kb && (jN || kN))
        de = (K[g.L(k + 1, j, i)] + K[g.L(k, j, i)] - K[g.L(k + 1, j - 1, i)] - K[g.L(k, j - 1, i)]) * 0.5;
    else if (j == 1 && kb && (jN || kN))
        de = (K[g.L(k + 1, j + 1, i)] + K[g.L(k, j + 1, i)] - K[g.L(k + 1, j, i)] - K[g.L(k, j, i)]) * 0.5;
    else
        de = (K[g.L(k + 1, j + 1, i)] + K[g.L(k, j + 1, i)] - K[g.L(k + 1, j - 1, i)] - K[g.L(k, j - 1, i)]) * 0.25;
    dz = K[g.L(k + 1, j, i)] - K[g.L(k, j, i)];
}

// getFace2Cell4Stencil{Csi,Eta,Zet} (src/include/inline.h:2190-2330): the outer cells of the
// four-cell stencil of face f on an axis with m points
__device__ __forceinline__ void face_stencil(int ptype, int f, int m, int &L, int &R) {
    if (f == 0 || f == m - 2) {
        if (ptype == TOSCA_PERIODIC_COPY) { L = m - 3; R = 2; }
        else if (ptype == TOSCA_PERIODIC_WRAP && f == m - 2) { R = m + 2; L = f - 1; }
        else if (ptype == TOSCA_PERIODIC_WRAP && f == 0) { L = -3; R = f + 2; }
        else if (f == 0) { L = f; R = f + 2; }
        else { L = f - 1; R = f + 1; }
    } else {
        L = f - 1; R = f + 2;
    }
}

// weno3 (src/include/inline.h:5147-5179): third-order WENO face value upwinded on the sign of the
// face flux; the reference squares with pow(x, 2.), here the correctly rounded product
__device__ __forceinline__ double weno3(double f0, double f1, double f2, double f3, double wavespeed) {
    double fL, fC, fR;
    if (wavespeed > 0) { fL = f0; fC = f1; fR = f2; }
    else { fL = f3; fC = f2; fR = f1; }
    const double d0 = 2. / 3., d1 = 1. / 3.;
    const double eps = 1.e-6;
    const double beta0 = (fC - fR) * (fC - fR);
    const double beta1 = (fL - fC) * (fL - fC);
    const double alpha0 = d0 / ((eps + beta0) * (eps + beta0));
    const double alpha1 = d1 / ((eps + beta1) * (eps + beta1));
    const double sumalpha = alpha0 + alpha1;
    const double w0 = alpha0 / sumalpha, w1 = alpha1 / sumalpha;
    const double u0 = (fC * 0.5 + fR * 0.5);
    const double u1 = (-fL * 0.5 + fC * 1.5);
    return w0 * u0 + w1 * u1;
}

// effective diffusivity of a face between cells c and cn (teqn_solvers.c:168-200 and the j / k
// copies): nu/Pr without LES; with LES (nu set to 0) the face-averaged k_t of the models that
// carry one, else nu_t over the turbulent Prandtl number 1 / (1 + 2 l / delta), with the
// stability length l = min(delta, 7.6 nu_t / delta sqrt(T_ref / |g . grad T|)) in stable
// stratification (g . grad T < 0) and l = delta otherwise
__device__ __forceinline__ double kappa_eff(const FormTArgs &A, long c, long cn, double ajc, double dtde) {
    if (!A.les) return A.nu / A.pr;
    const double nu = 0.0;
    const double nut = 0.5 * (A.nut[c] + A.nut[cn]);
    if (A.ktm) {
        const double diff = 0.5 * (A.kt[c] + A.kt[cn]);
        return (nu / A.pr) + diff;
    }
    const double gradTdotG = dtde * (-9.81);
    const double delta = pow(1. / ajc, 1. / 3.);
    double l;
    if (gradTdotG < 0.) {
        const double ls = 7.6 * nut / delta * sqrt(A.tref / fabs(gradTdotG));
        l = delta < ls ? delta : ls;
    }
    else l = delta;
    const double Prt = 1.0 / (1.0 + (2.0 * l / delta));
    return (nu / A.pr) + (nut / Prt);
}

// The advective (d) and diffusive (v) flux of one face, as the reference's face loop evaluates it
// (teqn_solvers.c:127-246 / 250-370 / 374-478).  Face (k, j, i) of direction DIR lies between
// cell (k, j, i) and its + neighbour along DIR.
__device__ __forceinline__ void flux_i(const FormTArgs &A, const Geo &g, int i, int j, int k, double &d, double &v) {
    const double *T = A.t;
    const long c = g.L(k, j, i);
    const double ajc = A.iaj[c];
    const V3 cs = ld3(A.icsi, c), et = ld3(A.ieta, c), ze = ld3(A.izet, c);
    const double g11 = cs.x * cs.x + cs.y * cs.y + cs.z * cs.z;
    const double g21 = et.x * cs.x + et.y * cs.y + et.z * cs.z;
    const double g31 = ze.x * cs.x + ze.y * cs.y + ze.z * cs.z;
    double dtdc, dtde, dtdz;
    dscalar_i(g, T, i, j, k, dtdc, dtde, dtdz);
    int iL, iR;
    face_stencil(g.per[0], i, g.mx, iL, iR);
    const double u = A.ucont[3 * c];
    d = -u * weno3(T[g.L(k, j, iL)], T[c], T[g.L(k, j, i + 1)], T[g.L(k, j, iR)], u);
    double kap = kappa_eff(A, c, g.L(k, j, i + 1), ajc, dtde);
    v = 0.0;
    if (A.les && ((A.wall[0] && i == 0) || (A.wall[1] && i == g.mx - 2))) {
        const double signQ = (i == 0) ? -1.0 : 1.0;
        const long pl = (long)g.mz * g.my, o = (long)k * g.my + j;
        v = signQ * (A.qI[o] * cs.x + A.qI[pl + o] * cs.y + A.qI[2 * pl + o] * cs.z);
        kap = 0.0;
    }
    v += (g11 * dtdc + g21 * dtde + g31 * dtdz) * ajc * kap;
}

__device__ __forceinline__ void flux_j(const FormTArgs &A, const Geo &g, int i, int j, int k, double &d, double &v) {
    const double *T = A.t;
    const long c = g.L(k, j, i);
    const double ajc = A.jaj[c];
    const V3 cs = ld3(A.jcsi, c), et = ld3(A.jeta, c), ze = ld3(A.jzet, c);
    const double g11 = cs.x * et.x + cs.y * et.y + cs.z * et.z;
    const double g21 = et.x * et.x + et.y * et.y + et.z * et.z;
    const double g31 = ze.x * et.x + ze.y * et.y + ze.z * et.z;
    double dtdc, dtde, dtdz;
    dscalar_j(g, T, i, j, k, dtdc, dtde, dtdz);
    int jL, jR;
    face_stencil(g.per[1], j, g.my, jL, jR);
    const double u = A.ucont[3 * c + 1];
    d = -u * weno3(T[g.L(k, jL, i)], T[c], T[g.L(k, j + 1, i)], T[g.L(k, jR, i)], u);
    double kap = kappa_eff(A, c, g.L(k, j + 1, i), ajc, dtde);
    v = 0.0;
    if (A.les && ((A.wall[2] && j == 0) || (A.wall[3] && j == g.my - 2))) {
        const double signQ = (j == 0) ? -1.0 : 1.0;
        const long pl = (long)g.mz * g.mx, o = (long)k * g.mx + i;
        v = signQ * (A.qJ[o] * et.x + A.qJ[pl + o] * et.y + A.qJ[2 * pl + o] * et.z);
        kap = 0.0;
    }
    v += (g11 * dtdc + g21 * dtde + g31 * dtdz) * ajc * kap;
}

__device__ __forceinline__ void flux_k(const FormTArgs &A, const Geo &g, int i, int j, int k, double &d, double &v) {
    const double *T = A.t;
    const long c = g.L(k, j, i);
    const double ajc = A.kaj[c];
    const V3 cs = ld3(A.kcsi, c), et = ld3(A.keta, c), ze = ld3(A.kzet, c);
    const double g11 = cs.x * ze.x + cs.y * ze.y + cs.z * ze.z;
    const double g21 = et.x * ze.x + et.y * ze.y + et.z * ze.z;
    const double g31 = ze.x * ze.x + ze.y * ze.y + ze.z * ze.z;
    double dtdc, dtde, dtdz;
    dscalar_k(g, T, i, j, k, dtdc, dtde, dtdz);
    int kL, kR;
    face_stencil(g.per[2], k, g.mz, kL, kR);
    const double u = A.ucont[3 * c + 2];
    d = -u * weno3(T[g.L(kL, j, i)], T[c], T[g.L(k + 1, j, i)], T[g.L(kR, j, i)], u);
    const double kap = kappa_eff(A, c, g.L(k + 1, j, i), ajc, dtde);
    v = 0.0;
    v += (g11 * dtdc + g21 * dtde + g31 * dtdz) * ajc * kap;
}

// One thread per local (ghosted) point; points off the reference's face loop get zero fluxes.
__global__ void k_formt_faces(FormTArgs A, Geo g) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x - GW, j = blockIdx.y * blockDim.y + threadIdx.y - GW,
              k = blockIdx.z - GW;
    if (i >= g.mx + GW || j >= g.my + GW) return;
    const long c = g.L(k, j, i);
    V3 dv{0.0, 0.0, 0.0}, vs{0.0, 0.0, 0.0};
    const bool in = i >= 0 && j >= 0 && k >= 0 && i < g.mx - 1 && j < g.my - 1 && k < g.mz - 1;
    if (in) {
        if (j != 0 && k != 0) flux_i(A, g, i, j, k, dv.x, vs.x);
        if (i != 0 && k != 0) flux_j(A, g, i, j, k, dv.y, vs.y);
        if (i != 0 && j != 0) flux_k(A, g, i, j, k, dv.z, vs.z);
    }
    st3(A.div, c, dv);
    st3(A.visc, c, vs);
}

// rhs = scale ((formMode != 2 ? div budget : 0) + (formMode != 1 ? visc budget : 0)) aj, interior
// cells only (the reference writes nothing else of Rhs)
__global__ void k_formt_cells(const double *div, const double *visc, const double *aj, double *rhs, double scale,
                              int mode, Geo g) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1, j = blockIdx.y * blockDim.y + threadIdx.y + 1,
              k = blockIdx.z + 1;
    if (i > g.mx - 2 || j > g.my - 2 || k > g.mz - 2) return;
    const long c = g.L(k, j, i), ci = g.L(k, j, i - 1), cj = g.L(k, j - 1, i), ck = g.L(k - 1, j, i);
    const double divBudget = div[3 * c] - div[3 * ci] + div[3 * c + 1] - div[3 * cj + 1] + div[3 * c + 2] - div[3 * ck + 2];
    const double viscBudget =
        visc[3 * c] - visc[3 * ci] + visc[3 * c + 1] - visc[3 * cj + 1] + visc[3 * c + 2] - visc[3 * ck + 2];
    rhs[g.G(k, j, i)] = scale * ((mode != 2 ? divBudget : 0.0) + (mode != 1 ? viscBudget : 0.0)) * aj[c];
}

bool kt_model(int m) {
    return m == TOSCA_LES_AMD || m == TOSCA_LES_DSM || m == TOSCA_LES_DLASI || m == TOSCA_LES_DLASD ||
           m == TOSCA_LES_DPASD || m == TOSCA_LES_BAMD;
}

}  // namespace

namespace {
FormTArgs formt_args(const Geo &g, const tosca_mesh *m, const tosca_params *p, const tosca_teqn_params *tp,
                     const double *ucont, const double *theta, const double *nut, const double *kt) {
    FormTArgs A;
    A.ucont = ucont; A.t = theta; A.nut = nut; A.kt = kt;
    A.icsi = m->icsi; A.ieta = m->ieta; A.izet = m->izet;
    A.jcsi = m->jcsi; A.jeta = m->jeta; A.jzet = m->jzet;
    A.kcsi = m->kcsi; A.keta = m->keta; A.kzet = m->kzet;
    A.iaj = m->iaj; A.jaj = m->jaj; A.kaj = m->kaj;
    A.qI = tp->q_wall_i; A.qJ = tp->q_wall_j;
    A.div = nullptr; A.visc = nullptr;
    A.nu = p->nu; A.pr = tp->pr; A.tref = p->t_ref;
    A.les = p->les_active != 0;
    A.ktm = A.les && kt_model(p->les_model);
    for (int q = 0; q < 4; q++) A.wall[q] = tp->wall_t[q];
    return A;
}
}  // namespace

void ref_formt(const Geo &g, const tosca_mesh *m, const tosca_params *p, const tosca_teqn_params *tp,
               const double *ucont, const double *theta, const double *nut, const double *kt, double *rhs,
               double scale, int mode, double *ws, hipStream_t s) {
    const long n = g.nloc();
    FormTArgs A = formt_args(g, m, p, tp, ucont, theta, nut, kt);
    A.div = ws; A.visc = ws + 3 * n;
    const dim3 blk(64, 4, 1);
    TOSCA_LAUNCH("k_formt_faces", k_formt_faces, dim3((g.gxm + 63) / 64, (g.gym + 3) / 4, g.gzm), blk, 0, s, A, g);
    launch_wrap(A.div, 3, g, s);
    launch_wrap(A.visc, 3, g, s);
    launch_reset_face_periodic(A.div, 3, g, s);
    launch_reset_face_periodic(A.visc, 3, g, s);
    TOSCA_LAUNCH("k_formt_cells", k_formt_cells, dim3((g.mx - 2 + 63) / 64, (g.my - 2 + 3) / 4, g.mz - 2), blk, 0, s,
                 A.div, A.visc, m->aj, rhs, scale, mode, g);
}

}  // namespace tosca
