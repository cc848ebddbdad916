// tosca_oracle.cpp -- CPU oracle for the momentum right-hand side + LES path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker the HIP library is compared against; it is
// never linked into, called by, or used as a fallback for the product library.  Only tests/,
// the smoke check and the CPU-baseline leg of the benchmark load it.
//
// It restates, pass for pass and with the same workspace arrays, loop order and FP64 expression
// order, the reference routines of sebastipa/TOSCA (paths relative to the reference root):
//   FormU            src/solvers/ueqn_solvers.c:6-1083
//   UpdateCs         src/les.c:165-2115 (Smagorinsky :172-179, DSM :197-811 + :1779-1977,
//                                        DPASD second filter :812-1174, plane averages :1979-2048)
//   UpdateNut        src/les.c:2119-2578 (Smagorinsky / dynamic branch :2510-2535)
//   UpdateNutBCs     src/boundary.c:2630-2929 (zeroGradient / fixedValue / periodic subset)
//   Buoyancy         src/sources/ueqn_sources.c:2726-2880
// and, for the routines next to that path:
//   ExplicitRhsU, UeqnFE, UeqnRK4, IMEXMatVec, UeqnRK3CN_W/_SO, UeqnABCN, UeqnSNES's constant
//                    right-hand side and SNESFuncEval   src/solvers/ueqn_solvers.c:1087-1838
//   contravariantToCartesian   src/ueqn.c:528-635
//   KSPGMRES as kspIMEX runs it (PETSc 3.15: classical Gram-Schmidt, Givens rotations)
//   UpdateContravariantBCs / UpdateCartesianBCs / UpdateTemperatureBCs   src/boundary.c:614-2628
//   the Cabot wall law   src/wallfunctions.c:280-585
//   the Nieuwstadt inflow profile   src/inflow.c:2732-2811, with the complex gamma, digamma and
//                    hypergeometric functions of src/include/inline.h:372-525
//   the power-law and mapped inflow-database ghosts of UpdateCartesianBCs   src/boundary.c:1228-1622,
//                    with findInterpolationWeights(WithExtrap) and scaleHyperTang{Bot,Top}
//                    (inline.h:5873-5999)
//   FormT            src/solvers/teqn_solvers.c:6-595
//   GradP            src/peqn.c:3437-3773
//   projectBladeForce (uniformADM branch) and bodyForceCartesian2Contravariant
//                    src/turbines/turbines_force.c:717-932, 1741-1835
// and the helpers of src/include/inline.h they call (cited at each helper below).
//
// Every function is pinned by tests under tests/test_oracle_*.py against facts outside its code
// (closed forms, invariants, library routines, quadrature, brute force).  FormU's boundary
// behaviour -- the one-sided corner rows of Compute_du_i/j/k (inline.h:2762-2930), QUICK clipping
// at non-periodic boundary faces (inline.h:2215-2227), slip and wall-function faces, the fp ghost
// copies and the cell-centred metric of the face dot -- is pinned on bounded lattices with a
// linear velocity and NaN in every ghost point those rules never read
// (tests/test_oracle_boundary_pins.py).  PETSc's
// single-rank DMLocalToLocal scatter is restated from PETSc's documented DMDA semantics:
// a ghost point of a DM_BOUNDARY_PERIODIC axis holds the owned value at (index mod m); ghost
// points of a DM_BOUNDARY_GHOSTED axis are left untouched.
//
// Build: g++ -O3 -ffp-contract=off -fopenmp -shared -fPIC (see oracle/Makefile).  FP contraction
// is off so every product and sum is rounded exactly as the reference's gcc build without FMA.

#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>
#include <algorithm>
#include <functional>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/tosca_rhs.h"

// The FormU / UpdateNut / Buoyancy part is written over a scalar type Real: double here, and a
// dual number in tosca_oracle_tangent.cpp, which includes this file to obtain the exact FP64
// tangent (JVP) of the same code.
#ifndef ORACLE_TANGENT
using Real = double;
#define ORACLE_MESH tosca_mesh
#define ORACLE_API(name) name
#endif

namespace {

using std::pow;
using std::sqrt;

struct V3 { Real x, y, z; };
#ifndef ORACLE_TANGENT
static_assert(sizeof(V3) == 24, "Cmpnts layout");
#endif

thread_local std::string g_err;

struct Lay {
    int64_t mx, my, mz;
    int64_t gxs, gys, gzs, gxm, gym, gzm;
    int per[3];      // periodic type per axis i, j, k
    inline int64_t L(int64_t k, int64_t j, int64_t i) const {
        return ((k - gzs) * gym + (j - gys)) * gxm + (i - gxs);
    }
    inline int64_t G(int64_t k, int64_t j, int64_t i) const { return (k * my + j) * mx + i; }
    int64_t nloc() const { return gxm * gym * gzm; }
    bool i_per() const { return per[0] == TOSCA_PERIODIC_COPY; }
    bool ii_per() const { return per[0] == TOSCA_PERIODIC_WRAP; }
    bool j_per() const { return per[1] == TOSCA_PERIODIC_COPY; }
    bool jj_per() const { return per[1] == TOSCA_PERIODIC_WRAP; }
    bool k_per() const { return per[2] == TOSCA_PERIODIC_COPY; }
    bool kk_per() const { return per[2] == TOSCA_PERIODIC_WRAP; }
    bool anyper() const { return per[0] || per[1] || per[2]; }
};

int check_layout(const tosca_dmda_info *in, const tosca_params *p, Lay &l) {
    if (!in || !p) { g_err = "null info/params"; return TOSCA_ERR_ARG; }
    if (in->xs != 0 || in->ys != 0 || in->zs != 0 || in->xm != in->mx || in->ym != in->my ||
        in->zm != in->mz || in->gxs != -3 || in->gys != -3 || in->gzs != -3 ||
        in->gxm != in->mx + 6 || in->gym != in->my + 6 || in->gzm != in->mz + 6 ||
        in->mx < 3 || in->my < 3 || in->mz < 3) {
        g_err = "layout must be a whole-domain DMDA with ghost width 3";
        return TOSCA_ERR_LAYOUT;
    }
    for (int a = 0; a < 3; a++)
        if (p->periodic[a] < 0 || p->periodic[a] > 2) { g_err = "bad periodic type"; return TOSCA_ERR_ARG; }
    if (p->ibm_active || p->overset_active) { g_err = "IBM / overset not supported"; return TOSCA_ERR_UNSUPPORTED; }
    l.mx = in->mx; l.my = in->my; l.mz = in->mz;
    l.gxs = in->gxs; l.gys = in->gys; l.gzs = in->gzs;
    l.gxm = in->gxm; l.gym = in->gym; l.gzm = in->gzm;
    for (int a = 0; a < 3; a++) l.per[a] = p->periodic[a];
    return TOSCA_OK;
}

// ---- PETSc single-rank DMLocalToLocal (INSERT_VALUES) ------------------------------------------
// Ghost points whose every out-of-range axis is DM_BOUNDARY_PERIODIC (type-2 periodicity) take the
// owned value at the wrapped index; other ghost points are not touched.
template <typename T>
void local_to_local(const Lay &l, T *a) {
    const int64_t m[3] = {l.mx, l.my, l.mz};
    for (int64_t k = l.gzs; k < l.gzs + l.gzm; k++)
        for (int64_t j = l.gys; j < l.gys + l.gym; j++)
            for (int64_t i = l.gxs; i < l.gxs + l.gxm; i++) {
                int64_t idx[3] = {i, j, k};
                bool ghost = false, fill = true;
                for (int ax = 0; ax < 3; ax++) {
                    if (idx[ax] < 0 || idx[ax] >= m[ax]) {
                        ghost = true;
                        if (l.per[ax] != TOSCA_PERIODIC_WRAP) fill = false;
                        idx[ax] = ((idx[ax] % m[ax]) + m[ax]) % m[ax];
                    }
                }
                if (ghost && fill) a[l.L(k, j, i)] = a[l.L(idx[2], idx[1], idx[0])];
            }
}

// ---- resetFacePeriodicFluxesVector, localToLocal (src/include/inline.h:5357-5451) ---------------
void reset_face_periodic(const Lay &l, V3 *v) {
    if (l.anyper()) {
        for (int64_t k = 0; k < l.mz; k++)
            for (int64_t j = 0; j < l.my; j++)
                for (int64_t i = 0; i < l.mx; i++) {
                    int flag = 0; int64_t a = i, b = j, c = k;
                    if (l.i_per() && i == 0) a = l.mx - 2, flag = 1;
                    if (l.j_per() && j == 0) b = l.my - 2, flag = 1;
                    if (l.k_per() && k == 0) c = l.mz - 2, flag = 1;
                    if (l.ii_per() && i == 0) a = -2, flag = 1;
                    if (l.jj_per() && j == 0) b = -2, flag = 1;
                    if (l.kk_per() && k == 0) c = -2, flag = 1;
                    if (flag) v[l.L(k, j, i)] = v[l.L(c, b, a)];
                }
    }
    local_to_local(l, v);
}

// ---- resetCellPeriodicFluxes, localToLocal (src/include/inline.h:5453-5606) ---------------------
template <typename T>
void reset_cell_periodic(const Lay &l, T *v) {
    if (l.anyper()) {
        for (int64_t k = 0; k < l.mz; k++)
            for (int64_t j = 0; j < l.my; j++)
                for (int64_t i = 0; i < l.mx; i++) {
                    int flag = 0; int64_t a = i, b = j, c = k;
                    if (l.i_per() && i == 0) a = l.mx - 2, flag = 1;
                    else if (l.i_per() && i == l.mx - 1) a = 1, flag = 1;
                    if (l.j_per() && j == 0) b = l.my - 2, flag = 1;
                    else if (l.j_per() && j == l.my - 1) b = 1, flag = 1;
                    if (l.k_per() && k == 0) c = l.mz - 2, flag = 1;
                    else if (l.k_per() && k == l.mz - 1) c = 1, flag = 1;
                    if (l.ii_per() && i == 0) a = -2, flag = 1;
                    else if (l.ii_per() && i == l.mx - 1) a = l.mx + 1, flag = 1;
                    if (l.jj_per() && j == 0) b = -2, flag = 1;
                    else if (l.jj_per() && j == l.my - 1) b = l.my + 1, flag = 1;
                    if (l.kk_per() && k == 0) c = -2, flag = 1;
                    else if (l.kk_per() && k == l.mz - 1) c = l.mz + 1, flag = 1;
                    if (flag) v[l.L(k, j, i)] = v[l.L(c, b, a)];
                }
    }
    local_to_local(l, v);
}

// ---- convective interpolants (src/include/inline.h:4314-4324, 5089-5109, 683) --------------------
inline V3 centralVec(V3 f0, V3 f1) { return {0.5 * (f0.x + f1.x), 0.5 * (f0.y + f1.y), 0.5 * (f0.z + f1.z)}; }

inline V3 quadraticUpwindVec(V3 f0, V3 f1, V3 f2, V3 f3, Real wavespeed) {
    V3 fUU, fU, fD;
    if (wavespeed > 0) { fUU = f0; fU = f1; fD = f2; }
    else               { fUU = f3; fU = f2; fD = f1; }
    // nSum(nScale(0.5, nSum(fU, fD)), nScale(1./8., nSub(nScale(2.0, fU), nSum(fD, fUU))))
    V3 r;
    r.x = 0.5 * (fU.x + fD.x) + (1. / 8.) * (2.0 * fU.x - (fD.x + fUU.x));
    r.y = 0.5 * (fU.y + fD.y) + (1. / 8.) * (2.0 * fU.y - (fD.y + fUU.y));
    r.z = 0.5 * (fU.z + fD.z) + (1. / 8.) * (2.0 * fU.z - (fD.z + fUU.z));
    return r;
}

inline V3 nScale(Real c, V3 v) { return {c * v.x, c * v.y, c * v.z}; }

// ---- getFace2Cell4Stencil{Csi,Eta,Zet} with nvert = meshTag = 0 (src/include/inline.h:2190-2380)
inline void face2cell4(int ptype, int64_t f, int64_t m, int64_t *L, int64_t *R) {
    if (f == 0 || f == m - 2) {
        if (ptype == TOSCA_PERIODIC_COPY) { *L = m - 3; *R = 2; }
        else if (ptype == TOSCA_PERIODIC_WRAP && f == m - 2) { *R = m + 2; *L = f - 1; }
        else if (ptype == TOSCA_PERIODIC_WRAP && f == 0) { *L = -3; *R = f + 2; }
        else {
            if (f == 0) { *L = f; *R = f + 2; }
            if (f == m - 2) { *L = f - 1; *R = f + 1; }
        }
    } else {
        *L = f - 1; *R = f + 2;
    }
}

struct Grad { Real dudc, dvdc, dwdc, dude, dvde, dwde, dudz, dvdz, dwdz; };

// ---- Compute_du_i / _j / _k with nvert = meshTag = 0 (src/include/inline.h:2748-2931) ------------
inline Grad du_i(const Lay &l, const V3 *u, int64_t i, int64_t j, int64_t k) {
    const bool iN = !l.i_per() && !l.ii_per(), jN = !l.j_per() && !l.jj_per(), kN = !l.k_per() && !l.kk_per();
    auto U = [&](int64_t kk, int64_t jj, int64_t ii) { return u[l.L(kk, jj, ii)]; };
    Grad g;
    g.dudc = U(k, j, i + 1).x - U(k, j, i).x;
    g.dvdc = U(k, j, i + 1).y - U(k, j, i).y;
    g.dwdc = U(k, j, i + 1).z - U(k, j, i).z;
    if (j == l.my - 2 && (i == 0 || i == l.mx - 2) && (jN || iN)) {
        g.dude = (U(k, j, i + 1).x + U(k, j, i).x - U(k, j - 1, i + 1).x - U(k, j - 1, i).x) * 0.5;
        g.dvde = (U(k, j, i + 1).y + U(k, j, i).y - U(k, j - 1, i + 1).y - U(k, j - 1, i).y) * 0.5;
        g.dwde = (U(k, j, i + 1).z + U(k, j, i).z - U(k, j - 1, i + 1).z - U(k, j - 1, i).z) * 0.5;
    } else if (j == 1 && (i == 0 || i == l.mx - 2) && (jN || iN)) {
        g.dude = (U(k, j + 1, i + 1).x + U(k, j + 1, i).x - U(k, j, i + 1).x - U(k, j, i).x) * 0.5;
        g.dvde = (U(k, j + 1, i + 1).y + U(k, j + 1, i).y - U(k, j, i + 1).y - U(k, j, i).y) * 0.5;
        g.dwde = (U(k, j + 1, i + 1).z + U(k, j + 1, i).z - U(k, j, i + 1).z - U(k, j, i).z) * 0.5;
    } else {
        g.dude = (U(k, j + 1, i + 1).x + U(k, j + 1, i).x - U(k, j - 1, i + 1).x - U(k, j - 1, i).x) * 0.25;
        g.dvde = (U(k, j + 1, i + 1).y + U(k, j + 1, i).y - U(k, j - 1, i + 1).y - U(k, j - 1, i).y) * 0.25;
        g.dwde = (U(k, j + 1, i + 1).z + U(k, j + 1, i).z - U(k, j - 1, i + 1).z - U(k, j - 1, i).z) * 0.25;
    }
    if (k == l.mz - 2 && (i == 0 || i == l.mx - 2) && (kN || iN)) {
        g.dudz = (U(k, j, i + 1).x + U(k, j, i).x - U(k - 1, j, i + 1).x - U(k - 1, j, i).x) * 0.5;
        g.dvdz = (U(k, j, i + 1).y + U(k, j, i).y - U(k - 1, j, i + 1).y - U(k - 1, j, i).y) * 0.5;
        g.dwdz = (U(k, j, i + 1).z + U(k, j, i).z - U(k - 1, j, i + 1).z - U(k - 1, j, i).z) * 0.5;
    } else if (k == 1 && (i == 0 || i == l.mx - 2) && (kN || iN)) {
        g.dudz = (U(k + 1, j, i + 1).x + U(k + 1, j, i).x - U(k, j, i + 1).x - U(k, j, i).x) * 0.5;
        g.dvdz = (U(k + 1, j, i + 1).y + U(k + 1, j, i).y - U(k, j, i + 1).y - U(k, j, i).y) * 0.5;
        g.dwdz = (U(k + 1, j, i + 1).z + U(k + 1, j, i).z - U(k, j, i + 1).z - U(k, j, i).z) * 0.5;
    } else {
        g.dudz = (U(k + 1, j, i + 1).x + U(k + 1, j, i).x - U(k - 1, j, i + 1).x - U(k - 1, j, i).x) * 0.25;
        g.dvdz = (U(k + 1, j, i + 1).y + U(k + 1, j, i).y - U(k - 1, j, i + 1).y - U(k - 1, j, i).y) * 0.25;
        g.dwdz = (U(k + 1, j, i + 1).z + U(k + 1, j, i).z - U(k - 1, j, i + 1).z - U(k - 1, j, i).z) * 0.25;
    }
    return g;
}

inline Grad du_j(const Lay &l, const V3 *u, int64_t i, int64_t j, int64_t k) {
    const bool iN = !l.i_per() && !l.ii_per(), jN = !l.j_per() && !l.jj_per(), kN = !l.k_per() && !l.kk_per();
    auto U = [&](int64_t kk, int64_t jj, int64_t ii) { return u[l.L(kk, jj, ii)]; };
    Grad g;
    if (i == l.mx - 2 && (j == 0 || j == l.my - 2) && (iN || jN)) {
        g.dudc = (U(k, j + 1, i).x + U(k, j, i).x - U(k, j + 1, i - 1).x - U(k, j, i - 1).x) * 0.5;
        g.dvdc = (U(k, j + 1, i).y + U(k, j, i).y - U(k, j + 1, i - 1).y - U(k, j, i - 1).y) * 0.5;
        g.dwdc = (U(k, j + 1, i).z + U(k, j, i).z - U(k, j + 1, i - 1).z - U(k, j, i - 1).z) * 0.5;
    } else if (i == 1 && (j == 0 || j == l.my - 2) && (iN || jN)) {
        g.dudc = (U(k, j + 1, i + 1).x + U(k, j, i + 1).x - U(k, j + 1, i).x - U(k, j, i).x) * 0.5;
        g.dvdc = (U(k, j + 1, i + 1).y + U(k, j, i + 1).y - U(k, j + 1, i).y - U(k, j, i).y) * 0.5;
        g.dwdc = (U(k, j + 1, i + 1).z + U(k, j, i + 1).z - U(k, j + 1, i).z - U(k, j, i).z) * 0.5;
    } else {
        g.dudc = (U(k, j + 1, i + 1).x + U(k, j, i + 1).x - U(k, j + 1, i - 1).x - U(k, j, i - 1).x) * 0.25;
        g.dvdc = (U(k, j + 1, i + 1).y + U(k, j, i + 1).y - U(k, j + 1, i - 1).y - U(k, j, i - 1).y) * 0.25;
        g.dwdc = (U(k, j + 1, i + 1).z + U(k, j, i + 1).z - U(k, j + 1, i - 1).z - U(k, j, i - 1).z) * 0.25;
    }
    g.dude = U(k, j + 1, i).x - U(k, j, i).x;
    g.dvde = U(k, j + 1, i).y - U(k, j, i).y;
    g.dwde = U(k, j + 1, i).z - U(k, j, i).z;
    if (k == l.mz - 2 && (j == 0 || j == l.my - 2) && (kN || jN)) {
        g.dudz = (U(k, j + 1, i).x + U(k, j, i).x - U(k - 1, j + 1, i).x - U(k - 1, j, i).x) * 0.5;
        g.dvdz = (U(k, j + 1, i).y + U(k, j, i).y - U(k - 1, j + 1, i).y - U(k - 1, j, i).y) * 0.5;
        g.dwdz = (U(k, j + 1, i).z + U(k, j, i).z - U(k - 1, j + 1, i).z - U(k - 1, j, i).z) * 0.5;
    } else if (k == 1 && (j == 0 || j == l.my - 2) && (kN || jN)) {
        g.dudz = (U(k + 1, j + 1, i).x + U(k + 1, j, i).x - U(k, j + 1, i).x - U(k, j, i).x) * 0.5;
        g.dvdz = (U(k + 1, j + 1, i).y + U(k + 1, j, i).y - U(k, j + 1, i).y - U(k, j, i).y) * 0.5;
        g.dwdz = (U(k + 1, j + 1, i).z + U(k + 1, j, i).z - U(k, j + 1, i).z - U(k, j, i).z) * 0.5;
    } else {
        g.dudz = (U(k + 1, j + 1, i).x + U(k + 1, j, i).x - U(k - 1, j + 1, i).x - U(k - 1, j, i).x) * 0.25;
        g.dvdz = (U(k + 1, j + 1, i).y + U(k + 1, j, i).y - U(k - 1, j + 1, i).y - U(k - 1, j, i).y) * 0.25;
        g.dwdz = (U(k + 1, j + 1, i).z + U(k + 1, j, i).z - U(k - 1, j + 1, i).z - U(k - 1, j, i).z) * 0.25;
    }
    return g;
}

inline Grad du_k(const Lay &l, const V3 *u, int64_t i, int64_t j, int64_t k) {
    const bool iN = !l.i_per() && !l.ii_per(), jN = !l.j_per() && !l.jj_per(), kN = !l.k_per() && !l.kk_per();
    auto U = [&](int64_t kk, int64_t jj, int64_t ii) { return u[l.L(kk, jj, ii)]; };
    Grad g;
    if (i == l.mx - 2 && (k == 0 || k == l.mz - 2) && (iN || kN)) {
        g.dudc = (U(k + 1, j, i).x + U(k, j, i).x - U(k + 1, j, i - 1).x - U(k, j, i - 1).x) * 0.5;
        g.dvdc = (U(k + 1, j, i).y + U(k, j, i).y - U(k + 1, j, i - 1).y - U(k, j, i - 1).y) * 0.5;
        g.dwdc = (U(k + 1, j, i).z + U(k, j, i).z - U(k + 1, j, i - 1).z - U(k, j, i - 1).z) * 0.5;
    } else if (i == 1 && (k == 0 || k == l.mz - 2) && (iN || kN)) {
        g.dudc = (U(k + 1, j, i + 1).x + U(k, j, i + 1).x - U(k + 1, j, i).x - U(k, j, i).x) * 0.5;
        g.dvdc = (U(k + 1, j, i + 1).y + U(k, j, i + 1).y - U(k + 1, j, i).y - U(k, j, i).y) * 0.5;
        g.dwdc = (U(k + 1, j, i + 1).z + U(k, j, i + 1).z - U(k + 1, j, i).z - U(k, j, i).z) * 0.5;
    } else {
        g.dudc = (U(k + 1, j, i + 1).x + U(k, j, i + 1).x - U(k + 1, j, i - 1).x - U(k, j, i - 1).x) * 0.25;
        g.dvdc = (U(k + 1, j, i + 1).y + U(k, j, i + 1).y - U(k + 1, j, i - 1).y - U(k, j, i - 1).y) * 0.25;
        g.dwdc = (U(k + 1, j, i + 1).z + U(k, j, i + 1).z - U(k + 1, j, i - 1).z - U(k, j, i - 1).z) * 0.25;
    }
    if (j == l.my - 2 && (k == 0 || k == l.mz - 2) && (jN || kN)) {
        g.dude = (U(k + 1, j, i).x + U(k, j, i).x - U(k + 1, j - 1, i).x - U(k, j - 1, i).x) * 0.5;
        g.dvde = (U(k + 1, j, i).y + U(k, j, i).y - U(k + 1, j - 1, i).y - U(k, j - 1, i).y) * 0.5;
        g.dwde = (U(k + 1, j, i).z + U(k, j, i).z - U(k + 1, j - 1, i).z - U(k, j - 1, i).z) * 0.5;
    } else if (j == 1 && (k == 0 || k == l.mz - 2) && (jN || kN)) {
        g.dude = (U(k + 1, j + 1, i).x + U(k, j + 1, i).x - U(k + 1, j, i).x - U(k, j, i).x) * 0.5;
        g.dvde = (U(k + 1, j + 1, i).y + U(k, j + 1, i).y - U(k + 1, j, i).y - U(k, j, i).y) * 0.5;
        g.dwde = (U(k + 1, j + 1, i).z + U(k, j + 1, i).z - U(k + 1, j, i).z - U(k, j, i).z) * 0.5;
    } else {
        g.dude = (U(k + 1, j + 1, i).x + U(k, j + 1, i).x - U(k + 1, j - 1, i).x - U(k, j - 1, i).x) * 0.25;
        g.dvde = (U(k + 1, j + 1, i).y + U(k, j + 1, i).y - U(k + 1, j - 1, i).y - U(k, j - 1, i).y) * 0.25;
        g.dwde = (U(k + 1, j + 1, i).z + U(k, j + 1, i).z - U(k + 1, j - 1, i).z - U(k, j - 1, i).z) * 0.25;
    }
    g.dudz = U(k + 1, j, i).x - U(k, j, i).x;
    g.dvdz = U(k + 1, j, i).y - U(k, j, i).y;
    g.dwdz = U(k + 1, j, i).z - U(k, j, i).z;
    return g;
}

// ---- Compute_du_center with nvert = meshTag = 0 (src/include/inline.h:3344-3418) ----------------
inline Grad du_center(const Lay &l, const V3 *u, int64_t i, int64_t j, int64_t k) {
    auto U = [&](int64_t kk, int64_t jj, int64_t ii) { return u[l.L(kk, jj, ii)]; };
    const bool iN = !l.i_per() && !l.ii_per(), jN = !l.j_per() && !l.jj_per(), kN = !l.k_per() && !l.kk_per();
    Grad g;
    if (iN && i == l.mx - 2) {
        g.dudc = (U(k, j, i).x - U(k, j, i - 1).x); g.dvdc = (U(k, j, i).y - U(k, j, i - 1).y); g.dwdc = (U(k, j, i).z - U(k, j, i - 1).z);
    } else if (iN && i == 1) {
        g.dudc = (U(k, j, i + 1).x - U(k, j, i).x); g.dvdc = (U(k, j, i + 1).y - U(k, j, i).y); g.dwdc = (U(k, j, i + 1).z - U(k, j, i).z);
    } else {
        g.dudc = (U(k, j, i + 1).x - U(k, j, i - 1).x) * 0.5; g.dvdc = (U(k, j, i + 1).y - U(k, j, i - 1).y) * 0.5; g.dwdc = (U(k, j, i + 1).z - U(k, j, i - 1).z) * 0.5;
    }
    if (jN && j == l.my - 2) {
        g.dude = (U(k, j, i).x - U(k, j - 1, i).x); g.dvde = (U(k, j, i).y - U(k, j - 1, i).y); g.dwde = (U(k, j, i).z - U(k, j - 1, i).z);
    } else if (jN && j == 1) {
        g.dude = (U(k, j + 1, i).x - U(k, j, i).x); g.dvde = (U(k, j + 1, i).y - U(k, j, i).y); g.dwde = (U(k, j + 1, i).z - U(k, j, i).z);
    } else {
        g.dude = (U(k, j + 1, i).x - U(k, j - 1, i).x) * 0.5; g.dvde = (U(k, j + 1, i).y - U(k, j - 1, i).y) * 0.5; g.dwde = (U(k, j + 1, i).z - U(k, j - 1, i).z) * 0.5;
    }
    if (kN && k == l.mz - 2) {
        g.dudz = (U(k, j, i).x - U(k - 1, j, i).x); g.dvdz = (U(k, j, i).y - U(k - 1, j, i).y); g.dwdz = (U(k, j, i).z - U(k - 1, j, i).z);
    } else if (kN && k == 1) {
        g.dudz = (U(k + 1, j, i).x - U(k, j, i).x); g.dvdz = (U(k + 1, j, i).y - U(k, j, i).y); g.dwdz = (U(k + 1, j, i).z - U(k, j, i).z);
    } else {
        g.dudz = (U(k + 1, j, i).x - U(k - 1, j, i).x) * 0.5; g.dvdz = (U(k + 1, j, i).y - U(k - 1, j, i).y) * 0.5; g.dwdz = (U(k + 1, j, i).z - U(k - 1, j, i).z) * 0.5;
    }
    return g;
}

struct Cart { Real du_dx, du_dy, du_dz, dv_dx, dv_dy, dv_dz, dw_dx, dw_dy, dw_dz; };

// ---- Compute_du_dxyz (src/include/inline.h:3635-3663) --------------------------------------------
inline Cart du_dxyz(V3 c, V3 e, V3 z, Real ajc, const Grad &g) {
    Cart r;
    r.du_dx = (g.dudc * c.x + g.dude * e.x + g.dudz * z.x) * ajc;
    r.du_dy = (g.dudc * c.y + g.dude * e.y + g.dudz * z.y) * ajc;
    r.du_dz = (g.dudc * c.z + g.dude * e.z + g.dudz * z.z) * ajc;
    r.dv_dx = (g.dvdc * c.x + g.dvde * e.x + g.dvdz * z.x) * ajc;
    r.dv_dy = (g.dvdc * c.y + g.dvde * e.y + g.dvdz * z.y) * ajc;
    r.dv_dz = (g.dvdc * c.z + g.dvde * e.z + g.dvdz * z.z) * ajc;
    r.dw_dx = (g.dwdc * c.x + g.dwde * e.x + g.dwdz * z.x) * ajc;
    r.dw_dy = (g.dwdc * c.y + g.dwde * e.y + g.dwdz * z.y) * ajc;
    r.dw_dz = (g.dwdc * c.z + g.dwde * e.z + g.dwdz * z.z) * ajc;
    return r;
}

// |S| = sqrt(2 S:S) (src/les.c:342-354, 2265-2271)
inline Real strain_mag(const Cart &d) {
    Real Sxx = 0.5 * (d.du_dx + d.du_dx), Sxy = 0.5 * (d.du_dy + d.dv_dx), Sxz = 0.5 * (d.du_dz + d.dw_dx);
    Real Syx = Sxy, Syy = 0.5 * (d.dv_dy + d.dv_dy), Syz = 0.5 * (d.dv_dz + d.dw_dy);
    Real Szx = Sxz, Szy = Syz, Szz = 0.5 * (d.dw_dz + d.dw_dz);
    return sqrt(2.0 * (Sxx * Sxx + Sxy * Sxy + Sxz * Sxz + Syx * Syx + Syy * Syy + Syz * Syz +
                            Szx * Szx + Szy * Szy + Szz * Szz));
}

inline Real nMag(V3 v) { return sqrt(v.x * v.x + v.y * v.y + v.z * v.z); }

// PetscMin / PetscMax
inline double pmax(double a, double b) { return (a < b) ? b : a; }
inline double pmin(double a, double b) { return (a < b) ? a : b; }

#ifndef ORACLE_TANGENT
// ---- integrateTestfilterSimpson / 5x5 (src/include/inline.h:3592-3633) ---------------------------
double simpson3(const double val[3][3][3], const double w[3][3][3]) {
    double wsum = 0, valsum = 0;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            for (int k = 0; k < 3; k++) {
                double sw = 1.0;
                if (i == 1) sw *= 4.;
                if (j == 1) sw *= 4.;
                if (k == 1) sw *= 4.;
                wsum += sw * w[i][j][k];
                valsum += sw * w[i][j][k] * val[i][j][k];
            }
    return valsum / wsum;
}

double simpson5(const double val[5][5][5], const double w[5][5][5]) {
    double wsum = 0, valsum = 0;
    for (int i = 0; i < 5; i++)
        for (int j = 0; j < 5; j++)
            for (int k = 0; k < 5; k++) {
                double sw = 1.0;
                sw *= (i == 0 || i == 4) ? 1 : (i % 2 == 0 ? 2 : 4);
                sw *= (j == 0 || j == 4) ? 1 : (j % 2 == 0 ? 2 : 4);
                sw *= (k == 0 || k == 4) ? 1 : (k % 2 == 0 ? 2 : 4);
                wsum += sw * w[i][j][k];
                valsum += sw * w[i][j][k] * val[i][j][k];
            }
    return valsum / wsum;
}

// ---- Calculate_Covariant_metrics (src/include/inline.h:1210-1229) --------------------------------
void covariant(const double g[3][3], double G[3][3]) {
    const double a11 = g[0][0], a12 = g[0][1], a13 = g[0][2];
    const double a21 = g[1][0], a22 = g[1][1], a23 = g[1][2];
    const double a31 = g[2][0], a32 = g[2][1], a33 = g[2][2];
    double det = a11 * (a33 * a22 - a32 * a23) - a21 * (a33 * a12 - a32 * a13) + a31 * (a23 * a12 - a22 * a13);
    G[0][0] = (a33 * a22 - a32 * a23) / det, G[0][1] = -(a33 * a12 - a32 * a13) / det, G[0][2] = (a23 * a12 - a22 * a13) / det;
    G[1][0] = -(a33 * a21 - a31 * a23) / det, G[1][1] = (a33 * a11 - a31 * a13) / det, G[1][2] = -(a23 * a11 - a21 * a13) / det;
    G[2][0] = (a32 * a21 - a31 * a22) / det, G[2][1] = -(a32 * a11 - a31 * a12) / det, G[2][2] = (a22 * a11 - a21 * a12) / det;
}

// Binomial test-filter kernels of src/les.c:597-612 (n=2) and :961-997 (n=4): products of the 1-D
// rows (1 2 1)/4 and (1 4 6 4 1)/16; every entry is a dyadic rational, equal to the printed tables.
const double bin3[3] = {0.25, 0.5, 0.25};
const double bin5[5] = {1.0 / 16, 4.0 / 16, 6.0 / 16, 4.0 / 16, 1.0 / 16};

// Germano contraction shared by the 3- and 5-point test filters (src/les.c:715-799, :1089-1164):
// returns num = L_ba M_aq G_bq and denom = M_nm M_nl G_ml.
void germano(const double L[3][3], const double Shat_ij[3][3], const double SShat_ij[3][3],
             double filter, double test_filter, V3 c0, V3 e0, V3 z0, double *num_out, double *den_out) {
    double S_hat = 0;
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) S_hat += std::pow(Shat_ij[a][b], 2.);
    S_hat = std::sqrt(2 * S_hat);
    double gg[3][3] = {{c0.x, c0.y, c0.z}, {e0.x, e0.y, e0.z}, {z0.x, z0.y, z0.z}}, ggc[3][3], G[3][3];
    covariant(gg, ggc);
    double xcsi = ggc[0][0], xeta = ggc[0][1], xzet = ggc[0][2];
    double ycsi = ggc[1][0], yeta = ggc[1][1], yzet = ggc[1][2];
    double zcsi = ggc[2][0], zeta = ggc[2][1], zzet = ggc[2][2];
    G[0][0] = xcsi * xcsi + ycsi * ycsi + zcsi * zcsi;
    G[1][1] = xeta * xeta + yeta * yeta + zeta * zeta;
    G[2][2] = xzet * xzet + yzet * yzet + zzet * zzet;
    G[0][1] = G[1][0] = xeta * xcsi + yeta * ycsi + zeta * zcsi;
    G[0][2] = G[2][0] = xzet * xcsi + yzet * ycsi + zzet * zcsi;
    G[1][2] = G[2][1] = xeta * xzet + yeta * yzet + zeta * zzet;
    double Mc[3][3], M[3][3];
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++)
            Mc[a][b] = (-std::pow(test_filter, 2.) * S_hat * Shat_ij[a][b] + std::pow(filter, 2.) * SShat_ij[a][b]);
    for (int a = 0; a < 3; a++) {
        M[a][0] = Mc[a][0] * c0.x + Mc[a][1] * c0.y + Mc[a][2] * c0.z;
        M[a][1] = Mc[a][0] * e0.x + Mc[a][1] * e0.y + Mc[a][2] * e0.z;
        M[a][2] = Mc[a][0] * z0.x + Mc[a][1] * z0.y + Mc[a][2] * z0.z;
    }
    double num = 0, denom = 0;
    for (int q = 0; q < 3; q++)
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) num += L[b][a] * M[a][q] * G[b][q];
    for (int m = 0; m < 3; m++)
        for (int n = 0; n < 3; n++)
            for (int l = 0; l < 3; l++) denom += M[n][m] * M[n][l] * G[m][l];
    *num_out = num;
    *den_out = denom;
}

// DSM / DPASD test filter at one interior cell, NP = 3 (src/les.c:466-801) or 5 (:820-1164).
template <int NP>
void test_filter_cell(const Lay &l, const tosca_mesh *m, const V3 *ucat, const V3 *Ax, const V3 *Ay,
                      const V3 *Az, const double *Sabs, int64_t i, int64_t j, int64_t k,
                      double *num, double *den) {
    const V3 *csi = (const V3 *)m->csi, *eta = (const V3 *)m->eta, *zet = (const V3 *)m->zet;
    const double *aj = m->aj;
    constexpr int H = NP / 2;
    // 40 stencil quantities (the reference keeps each in its own [NP][NP][NP] array)
    static thread_local double q[40][NP][NP][NP];
    double w[NP][NP][NP];
    for (int p = -H; p <= H; p++)
        for (int qq = -H; qq <= H; qq++)
            for (int r = -H; r <= H; r++) {
                int R = r + H, Q = qq + H, P = p + H;
                int64_t K = k + r, J = j + qq, I = i + p;
                int64_t c = l.L(K, J, I);
                double u = ucat[c].x, v = ucat[c].y, ww = ucat[c].z;
                double U = u * csi[c].x + v * csi[c].y + ww * csi[c].z;
                double V = u * eta[c].x + v * eta[c].y + ww * eta[c].z;
                double W = u * zet[c].x + v * zet[c].y + ww * zet[c].z;
                double du_dx = Ax[c].x, du_dy = Ax[c].y, du_dz = Ax[c].z;
                double dv_dx = Ay[c].x, dv_dy = Ay[c].y, dv_dz = Ay[c].z;
                double dw_dx = Az[c].x, dw_dy = Az[c].y, dw_dz = Az[c].z;
                double Sxx = 0.5 * (du_dx + du_dx), Sxy = 0.5 * (du_dy + dv_dx), Sxz = 0.5 * (du_dz + dw_dx);
                double Syx = Sxy, Syy = 0.5 * (dv_dy + dv_dy), Syz = 0.5 * (dv_dz + dw_dy);
                double Szx = Sxz, Szy = Syz, Szz = 0.5 * (dw_dz + dw_dz);
                double S = Sabs[c];
                double vals[40] = {u, v, ww, U, V, W,
                                   U * u, U * v, U * ww, V * u, V * v, V * ww, W * u, W * v, W * ww,
                                   Sxx, Sxy, Sxz, Syx, Syy, Syz, Szx, Szy, Szz,
                                   Sxx * S, Sxy * S, Sxz * S, Syx * S, Syy * S, Syz * S, Szx * S, Szy * S, Szz * S,
                                   S, 0, 0, 0, 0, 0, 0};
                for (int n = 0; n < 34; n++) q[n][R][Q][P] = vals[n];
            }
    double sum_vol = 0;
    for (int p = -H; p <= H; p++)
        for (int qq = -H; qq <= H; qq++)
            for (int r = -H; r <= H; r++) {
                int R = r + H, Q = qq + H, P = p + H;
                int64_t K = k + r, J = j + qq, I = i + p;
                // nvert = meshTag = 0: fluid and calculated everywhere; exclude DMDA boundary ghosts
                bool inside = (NP == 3) ? (I != 0 && I != l.mx - 1 && J != 0 && J != l.my - 1 && K != 0 && K != l.mz - 1)
                                        : (I > 0 && I < l.mx - 1 && J > 0 && J < l.my - 1 && K > 0 && K < l.mz - 1);
                if (inside) {
                    double coef = (NP == 3) ? bin3[R] * bin3[Q] * bin3[P] : bin5[R] * bin5[Q] * bin5[P];
                    double scale = (NP == 3) ? 8.0 : 64.0;
                    sum_vol += 1. / aj[l.L(K, J, I)] * scale * coef;
                    w[R][Q][P] = 1;
                } else {
                    w[R][Q][P] = 0;
                }
            }
    double filter = std::pow(1. / aj[l.L(k, j, i)], 1. / 3.);
    double test_filter = std::pow(sum_vol, 1. / 3.);
    auto F = [&](int n) {
        if (NP == 3) return simpson3(reinterpret_cast<const double(*)[3][3]>(q[n]), reinterpret_cast<const double(*)[3][3]>(w));
        return simpson5(reinterpret_cast<const double(*)[5][5]>(q[n]), reinterpret_cast<const double(*)[5][5]>(w));
    };
    double _u = 0, _v = 0, _w = 0, _U = 0, _V = 0, _W = 0;
    _U = F(3); _V = F(4); _W = F(5);
    _u = F(0); _v = F(1); _w = F(2);
    double L[3][3];
    L[0][0] = F(6) - _U * _u; L[0][1] = F(7) - _U * _v; L[0][2] = F(8) - _U * _w;
    L[1][0] = F(9) - _V * _u; L[1][1] = F(10) - _V * _v; L[1][2] = F(11) - _V * _w;
    L[2][0] = F(12) - _W * _u; L[2][1] = F(13) - _W * _v; L[2][2] = F(14) - _W * _w;
    double Sh[3][3], SSh[3][3];
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) Sh[a][b] = F(15 + 3 * a + b);
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) SSh[a][b] = F(24 + 3 * a + b);
    int64_t c0 = l.L(k, j, i);
    germano(L, Sh, SSh, filter, test_filter, csi[c0], eta[c0], zet[c0], num, den);
}

#endif  // ORACLE_TANGENT

#ifdef _OPENMP
#define OMP_FOR _Pragma("omp parallel for schedule(static)")
#else
#define OMP_FOR
#endif

}  // namespace

#ifndef ORACLE_TANGENT
extern "C" {
#else
namespace {
#endif

#ifndef ORACLE_TANGENT
const char *oracle_last_error(void) { return g_err.c_str(); }

void oracle_set_threads(int n) {
#ifdef _OPENMP
    omp_set_num_threads(n > 0 ? n : 1);
#else
    (void)n;
#endif
}

int oracle_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
#endif  // ORACLE_TANGENT

// ================================================================================================
// FormU (src/solvers/ueqn_solvers.c:6-1083)
// ================================================================================================
int ORACLE_API(oracle_formu)(const tosca_dmda_info *info, const ORACLE_MESH *m, const tosca_params *p,
                             const Real *ucont_, const Real *ucat_, const Real *nut, Real *rhs_,
                             double scale, int formMode) {
    Lay l;
    int rc = check_layout(info, p, l);
    if (rc) return rc;
    if (formMode < 0 || formMode > 2) { g_err = "bad formMode"; return TOSCA_ERR_ARG; }
    if (p->div_scheme != TOSCA_DIV_CENTRAL && p->div_scheme != TOSCA_DIV_QUICK) { g_err = "scheme"; return TOSCA_ERR_UNSUPPORTED; }
    if (formMode != 1 && (((p->bc_u[0] == TOSCA_BCU_WALLFUNCTION || p->bc_u[1] == TOSCA_BCU_WALLFUNCTION) && !p->tau_wall_i) ||
                          ((p->bc_u[2] == TOSCA_BCU_WALLFUNCTION || p->bc_u[3] == TOSCA_BCU_WALLFUNCTION) && !p->tau_wall_j))) {
        g_err = "wall-function patch without tau_wall";
        return TOSCA_ERR_ARG;
    }
    const int64_t mx = l.mx, my = l.my, mz = l.mz, n = l.nloc();
    const V3 *ucont = (const V3 *)ucont_, *ucat = (const V3 *)ucat_;
    const V3 *csi = (const V3 *)m->csi, *eta = (const V3 *)m->eta, *zet = (const V3 *)m->zet;
    const V3 *icsi = (const V3 *)m->icsi, *ieta = (const V3 *)m->ieta, *izet = (const V3 *)m->izet;
    const V3 *jcsi = (const V3 *)m->jcsi, *jeta = (const V3 *)m->jeta, *jzet = (const V3 *)m->jzet;
    const V3 *kcsi = (const V3 *)m->kcsi, *keta = (const V3 *)m->keta, *kzet = (const V3 *)m->kzet;
    const Real *iaj = m->iaj, *jaj = m->jaj, *kaj = m->kaj;
    V3 *rhs = (V3 *)rhs_;
    const bool les = p->les_active != 0;
    const Real nu = p->nu;
    const bool quick = p->div_scheme == TOSCA_DIV_QUICK;

    std::vector<V3> fp(n, V3{0, 0, 0});
    std::vector<V3> div1(n, V3{0, 0, 0}), div2(n, V3{0, 0, 0}), div3(n, V3{0, 0, 0});
    std::vector<V3> visc1(n, V3{0, 0, 0}), visc2(n, V3{0, 0, 0}), visc3(n, V3{0, 0, 0});

    const double *tauI = p->tau_wall_i, *tauJ = p->tau_wall_j;
    const int64_t tIx = 0, tIy = mz * my, tIz = 2 * mz * my;   // plane offsets [x|y|z][zm][ym]
    const int64_t tJx = 0, tJy = mz * mx, tJz = 2 * mz * mx;   // plane offsets [x|y|z][zm][xm]

    // viscous flux of one face (shared expression of :196-215, :410-429, :624-643, :386-392 ...)
    auto visc_flux = [](const Grad &g, V3 c, V3 e, V3 z, int dir, Real ajc, Real nuEff, V3 &out) {
        Real csi0 = c.x, csi1 = c.y, csi2 = c.z, eta0 = e.x, eta1 = e.y, eta2 = e.z, zet0 = z.x, zet1 = z.y, zet2 = z.z;
        Real g11, g21, g31;
        if (dir == 0) { g11 = csi0 * csi0 + csi1 * csi1 + csi2 * csi2; g21 = eta0 * csi0 + eta1 * csi1 + eta2 * csi2; g31 = zet0 * csi0 + zet1 * csi1 + zet2 * csi2; }
        else if (dir == 1) { g11 = csi0 * eta0 + csi1 * eta1 + csi2 * eta2; g21 = eta0 * eta0 + eta1 * eta1 + eta2 * eta2; g31 = zet0 * eta0 + zet1 * eta1 + zet2 * eta2; }
        else { g11 = csi0 * zet0 + csi1 * zet1 + csi2 * zet2; g21 = eta0 * zet0 + eta1 * zet1 + eta2 * zet2; g31 = zet0 * zet0 + zet1 * zet1 + zet2 * zet2; }
        Real r11 = g.dudc * csi0 + g.dude * eta0 + g.dudz * zet0;
        Real r21 = g.dvdc * csi0 + g.dvde * eta0 + g.dvdz * zet0;
        Real r31 = g.dwdc * csi0 + g.dwde * eta0 + g.dwdz * zet0;
        Real r12 = g.dudc * csi1 + g.dude * eta1 + g.dudz * zet1;
        Real r22 = g.dvdc * csi1 + g.dvde * eta1 + g.dvdz * zet1;
        Real r32 = g.dwdc * csi1 + g.dwde * eta1 + g.dwdz * zet1;
        Real r13 = g.dudc * csi2 + g.dude * eta2 + g.dudz * zet2;
        Real r23 = g.dvdc * csi2 + g.dvde * eta2 + g.dvdz * zet2;
        Real r33 = g.dwdc * csi2 + g.dwde * eta2 + g.dwdz * zet2;
        Real n0, n1, n2;   // the face's own normal: csi (i), eta (j) or zet (k)
        if (dir == 0) { n0 = csi0; n1 = csi1; n2 = csi2; }
        else if (dir == 1) { n0 = eta0; n1 = eta1; n2 = eta2; }
        else { n0 = zet0; n1 = zet1; n2 = zet2; }
        out.x += (g11 * g.dudc + g21 * g.dude + g31 * g.dudz + r11 * n0 + r21 * n1 + r31 * n2) * ajc * (nuEff);
        out.y += (g11 * g.dvdc + g21 * g.dvde + g31 * g.dvdz + r12 * n0 + r22 * n1 + r32 * n2) * ajc * (nuEff);
        out.z += (g11 * g.dwdc + g21 * g.dwde + g31 * g.dwdz + r13 * n0 + r23 * n1 + r33 * n2) * ajc * (nuEff);
    };

    // ---- face pass (:173-808) ---------------------------------------------------------------
    OMP_FOR
    for (int64_t k = 0; k < mz; k++)
        for (int64_t j = 0; j < my; j++)
            for (int64_t i = 0; i < mx; i++) {
                if (i == mx - 1 || j == my - 1 || k == mz - 1) continue;
                const int64_t c = l.L(k, j, i);
                // i faces
                if (j != 0 && k != 0) {
                    Grad g{};
                    Real ajc = 0.0;
                    if (formMode != 1) { ajc = iaj[c]; g = du_i(l, ucat, i, j, k); }
                    if (formMode != 2) {
                        if (quick) {
                            int64_t iL, iR; face2cell4(l.per[0], i, mx, &iL, &iR);
                            div1[c] = nScale(-ucont[c].x, quadraticUpwindVec(ucat[l.L(k, j, iL)], ucat[c], ucat[l.L(k, j, i + 1)], ucat[l.L(k, j, iR)], ucont[c].x));
                        } else {
                            div1[c] = nScale(-ucont[c].x, centralVec(ucat[c], ucat[l.L(k, j, i + 1)]));
                        }
                    }
                    if (formMode != 1) {
                        Real nuEff;
                        if (les) {
                            Real nt = 0.5 * (nut[c] + nut[l.L(k, j, i + 1)]);
                            if ((p->bc_u[0] == TOSCA_BCU_WALLFUNCTION && i == 0) || (p->bc_u[1] == TOSCA_BCU_WALLFUNCTION && i == mx - 2)) {
                                int64_t t = k * my + j;
                                visc1[c].x = -tauI[tIx + t]; visc1[c].y = -tauI[tIy + t]; visc1[c].z = -tauI[tIz + t];
                                nuEff = nu;
                            } else if ((p->bc_u[0] == TOSCA_BCU_SLIP && i == 0) || (p->bc_u[1] == TOSCA_BCU_SLIP && i == mx - 2)) {
                                nuEff = 0.0;
                            } else {
                                nuEff = nu + nt;
                            }
                        } else {
                            nuEff = nu;
                        }
                        visc_flux(g, icsi[c], ieta[c], izet[c], 0, ajc, nuEff, visc1[c]);
                    }
                }
                // j faces
                if (i != 0 && k != 0) {
                    Grad g{};
                    Real ajc = 0.0;
                    if (formMode != 1) { ajc = jaj[c]; g = du_j(l, ucat, i, j, k); }
                    if (formMode != 2) {
                        if (quick) {
                            int64_t jL, jR; face2cell4(l.per[1], j, my, &jL, &jR);
                            div2[c] = nScale(-ucont[c].y, quadraticUpwindVec(ucat[l.L(k, jL, i)], ucat[c], ucat[l.L(k, j + 1, i)], ucat[l.L(k, jR, i)], ucont[c].y));
                        } else {
                            div2[c] = nScale(-ucont[c].y, centralVec(ucat[c], ucat[l.L(k, j + 1, i)]));
                        }
                    }
                    if (formMode != 1) {
                        Real nuEff;
                        if (les) {
                            Real nt = 0.5 * (nut[c] + nut[l.L(k, j + 1, i)]);
                            if ((p->bc_u[2] == TOSCA_BCU_WALLFUNCTION && j == 0) || (p->bc_u[3] == TOSCA_BCU_WALLFUNCTION && j == my - 2)) {
                                int64_t t = k * mx + i;
                                visc2[c].x = -tauJ[tJx + t]; visc2[c].y = -tauJ[tJy + t]; visc2[c].z = -tauJ[tJz + t];
                                nuEff = nu;
                            } else if ((p->bc_u[2] == TOSCA_BCU_SLIP && j == 0) || (p->bc_u[3] == TOSCA_BCU_SLIP && j == my - 2)) {
                                nuEff = 0.0;
                            } else {
                                nuEff = nu + nt;
                            }
                        } else {
                            nuEff = nu;
                        }
                        visc_flux(g, jcsi[c], jeta[c], jzet[c], 1, ajc, nuEff, visc2[c]);
                    }
                }
                // k faces
                if (i != 0 && j != 0) {
                    Grad g{};
                    Real ajc = 0.0;
                    if (formMode != 1) { ajc = kaj[c]; g = du_k(l, ucat, i, j, k); }
                    if (formMode != 2) {
                        if (quick) {
                            int64_t kL, kR; face2cell4(l.per[2], k, mz, &kL, &kR);
                            div3[c] = nScale(-ucont[c].z, quadraticUpwindVec(ucat[l.L(kL, j, i)], ucat[c], ucat[l.L(k + 1, j, i)], ucat[l.L(kR, j, i)], ucont[c].z));
                        } else {
                            div3[c] = nScale(-ucont[c].z, centralVec(ucat[c], ucat[l.L(k + 1, j, i)]));
                        }
                    }
                    if (formMode != 1) {
                        Real nuEff;
                        if (les) {
                            Real nt = 0.5 * (nut[c] + nut[l.L(k + 1, j, i)]);
                            if ((p->bc_u[4] == TOSCA_BCU_SLIP && k == 0) || (p->bc_u[5] == TOSCA_BCU_SLIP && k == mz - 2)) nuEff = 0.0;
                            else nuEff = nu + nt;
                        } else {
                            nuEff = nu;
                        }
                        visc_flux(g, kcsi[c], keta[c], kzet[c], 2, ajc, nuEff, visc3[c]);
                    }
                }
            }

    // ---- ghost scatter + periodic face resets (:829-857) -------------------------------------
    V3 *arrs[6] = {div1.data(), div2.data(), div3.data(), visc1.data(), visc2.data(), visc3.data()};
    for (auto *a : arrs) local_to_local(l, a);
    for (auto *a : arrs) reset_face_periodic(l, a);

    // ---- cell pass (:871-993) ----------------------------------------------------------------
    const bool iN = !l.i_per() && !l.ii_per(), jN = !l.j_per() && !l.jj_per(), kN = !l.k_per() && !l.kk_per();
    OMP_FOR
    for (int64_t k = 1; k < mz - 1; k++)
        for (int64_t j = 1; j < my - 1; j++)
            for (int64_t i = 1; i < mx - 1; i++) {
                const int64_t c = l.L(k, j, i), cim = l.L(k, j, i - 1), cjm = l.L(k, j - 1, i), ckm = l.L(k - 1, j, i);
                V3 &f = fp[c];
                if (formMode != 2) {
                    f.x += (div1[c].x - div1[cim].x + div2[c].x - div2[cjm].x + div3[c].x - div3[ckm].x);
                    f.y += (div1[c].y - div1[cim].y + div2[c].y - div2[cjm].y + div3[c].y - div3[ckm].y);
                    f.z += (div1[c].z - div1[cim].z + div2[c].z - div2[cjm].z + div3[c].z - div3[ckm].z);
                }
                if (formMode != 1) {
                    f.x += (visc1[c].x - visc1[cim].x + visc2[c].x - visc2[cjm].x + visc3[c].x - visc3[ckm].x);
                    f.y += (visc1[c].y - visc1[cim].y + visc2[c].y - visc2[cjm].y + visc3[c].y - visc3[ckm].y);
                    f.z += (visc1[c].z - visc1[cim].z + visc2[c].z - visc2[cjm].z + visc3[c].z - visc3[ckm].z);
                }
            }
    // boundary-ghost copies of the same pass (:949-990), serial: they read finished interior cells
    for (int64_t k = 1; k < mz - 1; k++)
        for (int64_t j = 1; j < my - 1; j++)
            for (int64_t i = 1; i < mx - 1; i++) {
                const int64_t c = l.L(k, j, i);
                if (i == 1 && iN) fp[l.L(k, j, i - 1)] = fp[c];
                if (j == 1 && jN) fp[l.L(k, j - 1, i)] = fp[c];
                if (k == 1 && kN) fp[l.L(k - 1, j, i)] = fp[c];
                if (i == mx - 2 && iN) fp[l.L(k, j, i + 1)] = fp[c];
                if (j == my - 2 && jN) fp[l.L(k, j + 1, i)] = fp[c];
                if (k == mz - 2 && kN) fp[l.L(k + 1, j, i)] = fp[c];
            }
    local_to_local(l, fp.data());
    reset_cell_periodic(l, fp.data());

    // ---- face-dot pass (:1006-1035) ----------------------------------------------------------
    OMP_FOR
    for (int64_t k = 1; k < mz - 1; k++)
        for (int64_t j = 1; j < my - 1; j++)
            for (int64_t i = 1; i < mx - 1; i++) {
                const int64_t c = l.L(k, j, i);
                const int64_t ci = l.L(k, j, i + 1), cj = l.L(k, j + 1, i), ck = l.L(k + 1, j, i);
                V3 &r = rhs[l.G(k, j, i)];
                Real v1, v2;
                v1 = csi[c].x * fp[c].x + csi[c].y * fp[c].y + csi[c].z * fp[c].z;
                v2 = csi[ci].x * fp[ci].x + csi[ci].y * fp[ci].y + csi[ci].z * fp[ci].z;
                r.x += scale * iaj[c] * (0.5 * (v1 + v2));
                v1 = eta[c].x * fp[c].x + eta[c].y * fp[c].y + eta[c].z * fp[c].z;
                v2 = eta[cj].x * fp[cj].x + eta[cj].y * fp[cj].y + eta[cj].z * fp[cj].z;
                r.y += scale * jaj[c] * (0.5 * (v1 + v2));
                v1 = zet[c].x * fp[c].x + zet[c].y * fp[c].y + zet[c].z * fp[c].z;
                v2 = zet[ck].x * fp[ck].x + zet[ck].y * fp[ck].y + zet[ck].z * fp[ck].z;
                r.z += scale * kaj[c] * (0.5 * (v1 + v2));
            }
    return TOSCA_OK;
}

#ifndef ORACLE_TANGENT
// ================================================================================================
// UpdateCs (src/les.c:165-2115): SMAGORINSKY, DSM, DPASD
// ================================================================================================
int oracle_update_cs(const tosca_dmda_info *info, const tosca_mesh *m, const tosca_params *p,
                     const double *ucat_, double *Cs) {
    Lay l;
    int rc = check_layout(info, p, l);
    if (rc) return rc;
    const int64_t mx = l.mx, my = l.my, mz = l.mz, n = l.nloc();
    if (p->les_model == TOSCA_LES_SMAGORINSKY) {
        // VecSet(lCs, std_cs) over the whole local vector, then a ghost scatter (:172-179)
        for (int64_t c = 0; c < n; c++) Cs[c] = 0.0289;
        return TOSCA_OK;
    }
    if (p->les_model != TOSCA_LES_DSM && p->les_model != TOSCA_LES_DPASD) {
        g_err = "LES model not supported";
        return TOSCA_ERR_UNSUPPORTED;
    }
    const V3 *ucat = (const V3 *)ucat_;
    const V3 *csi = (const V3 *)m->csi, *eta = (const V3 *)m->eta, *zet = (const V3 *)m->zet;
    const double *aj = m->aj;

    std::vector<V3> Ax(n, V3{0, 0, 0}), Ay(n, V3{0, 0, 0}), Az(n, V3{0, 0, 0});
    std::vector<double> Sabs(n, 0.0);

    // 1 - strain rate (:277-374)
    OMP_FOR
    for (int64_t k = 1; k < mz - 1; k++)
        for (int64_t j = 1; j < my - 1; j++)
            for (int64_t i = 1; i < mx - 1; i++) {
                const int64_t c = l.L(k, j, i);
                Grad g = du_center(l, ucat, i, j, k);
                Cart d = du_dxyz(csi[c], eta[c], zet[c], aj[c], g);
                Sabs[c] = strain_mag(d);
                Ax[c] = {d.du_dx, d.du_dy, d.du_dz};
                Ay[c] = {d.dv_dx, d.dv_dy, d.dv_dz};
                Az[c] = {d.dw_dx, d.dw_dy, d.dw_dz};
            }
    local_to_local(l, Ax.data()); local_to_local(l, Ay.data());
    local_to_local(l, Az.data()); local_to_local(l, Sabs.data());
    reset_cell_periodic(l, Ax.data()); reset_cell_periodic(l, Ay.data());
    reset_cell_periodic(l, Az.data()); reset_cell_periodic(l, Sabs.data());

    // 2 - first test filter -> LM, MM (:452-811)
    std::vector<double> LM(n, 0.0), MM(n, 0.0);
    OMP_FOR
    for (int64_t k = 1; k < mz - 1; k++)
        for (int64_t j = 1; j < my - 1; j++)
            for (int64_t i = 1; i < mx - 1; i++) {
                const int64_t c = l.L(k, j, i);
                test_filter_cell<3>(l, m, ucat, Ax.data(), Ay.data(), Az.data(), Sabs.data(), i, j, k, &LM[c], &MM[c]);
            }
    local_to_local(l, LM.data()); local_to_local(l, MM.data());

    std::vector<double> QN, NN;
    if (p->les_model == TOSCA_LES_DPASD) {
        // second test filter -> QN, NN (:812-1174)
        QN.assign(n, 0.0); NN.assign(n, 0.0);
        OMP_FOR
        for (int64_t k = 1; k < mz - 1; k++)
            for (int64_t j = 1; j < my - 1; j++)
                for (int64_t i = 1; i < mx - 1; i++) {
                    const int64_t c = l.L(k, j, i);
                    test_filter_cell<5>(l, m, ucat, Ax.data(), Ay.data(), Az.data(), Sabs.data(), i, j, k, &QN[c], &NN[c]);
                }
        local_to_local(l, QN.data()); local_to_local(l, NN.data());
    }
    local_to_local(l, LM.data()); local_to_local(l, MM.data());
    reset_cell_periodic(l, LM.data()); reset_cell_periodic(l, MM.data());

    // 3 - Cs (:1779-1977)
    OMP_FOR
    for (int64_t k = 1; k < mz - 1; k++)
        for (int64_t j = 1; j < my - 1; j++)
            for (int64_t i = 1; i < mx - 1; i++) {
                double C = 0.0;
                if (p->les_model == TOSCA_LES_DSM) {
                    double weight[3][3][3], LM0[3][3][3], MM0[3][3][3];
                    for (int a = -1; a <= 1; a++)
                        for (int b = -1; b <= 1; b++)
                            for (int cc = -1; cc <= 1; cc++) {
                                int R = cc + 1, Q = b + 1, P = a + 1;
                                int64_t K = k + cc, J = j + b, I = i + a;
                                weight[R][Q][P] = 1. / aj[l.L(K, J, I)];
                                if (l.i_per()) { if (I == 0) I = mx - 2; else if (I == mx - 1) I = 1; }
                                else if (l.ii_per()) { if (I == 0) I = -2; else if (I == mx - 1) I = mx + 1; }
                                else if (I == 0 || I == mx - 1) weight[R][Q][P] = 0;
                                if (l.j_per()) { if (J == 0) J = my - 2; else if (J == my - 1) J = 1; }
                                else if (l.jj_per()) { if (J == 0) J = -2; else if (J == my - 1) J = my + 1; }
                                else if (J == 0 || j == my - 1) weight[R][Q][P] = 0;   // lower-case j as in :1911
                                if (l.k_per()) { if (K == 0) K = mz - 2; else if (K == mz - 1) K = 1; }
                                else if (l.kk_per()) { if (K == 0) K = -2; else if (K == mz - 1) K = mz + 1; }
                                else if (K == 0 || K == mz - 1) weight[R][Q][P] = 0;
                                LM0[R][Q][P] = LM[l.L(K, J, I)];
                                MM0[R][Q][P] = MM[l.L(K, J, I)];
                            }
                    double LM_avg = simpson3(LM0, weight), MM_avg = simpson3(MM0, weight);
                    C = 0.5 * LM_avg / (MM_avg + 1e-10);
                }
                Cs[l.L(k, j, i)] = pmin(pmax(C, 0.0), p->max_cs);
            }

    if (p->les_model == TOSCA_LES_DPASD) {
        // j-plane averages (:1979-2048); single rank, so MPI_Allreduce is the identity
        std::vector<double> J_LM(my, 0.0), J_MM(my, 0.0), J_QN(my, 0.0), J_NN(my, 0.0);
        std::vector<int64_t> count(my, 0);
        for (int64_t k = 1; k < mz - 1; k++)
            for (int64_t j = 1; j < my - 1; j++)
                for (int64_t i = 1; i < mx - 1; i++) {
                    const int64_t c = l.L(k, j, i);
                    J_LM[j] += LM[c]; J_MM[j] += MM[c]; J_QN[j] += QN[c]; J_NN[j] += NN[c];
                    count[j]++;
                }
        for (int64_t j = 0; j < my; j++)
            if (count[j] > 0) {
                J_LM[j] = J_LM[j] / count[j]; J_MM[j] = J_MM[j] / count[j];
                J_QN[j] = J_QN[j] / count[j]; J_NN[j] = J_NN[j] / count[j];
            }
        for (int64_t k = 1; k < mz - 1; k++)
            for (int64_t j = 1; j < my - 1; j++)
                for (int64_t i = 1; i < mx - 1; i++) {
                    double beta = (J_QN[j] * J_MM[j]) / (J_NN[j] * J_LM[j] + 1e-10);
                    double v = 0.5 * (J_LM[j] / (J_MM[j] + 1e-10)) / beta;
                    Cs[l.L(k, j, i)] = pmin(pmax(v, 0.0), p->max_cs);
                }
    }
    local_to_local(l, Cs);
    reset_cell_periodic(l, Cs);
    return TOSCA_OK;
}

#endif  // ORACLE_TANGENT

// ================================================================================================
// UpdateNut (src/les.c:2119-2578) + UpdateNutBCs (src/boundary.c:2630-2929)
// ================================================================================================
int ORACLE_API(oracle_update_nut)(const tosca_dmda_info *info, const ORACLE_MESH *m, const tosca_params *p,
                                  const Real *ucat_, const Real *Cs, Real *nut, Real *ksgs) {
    Lay l;
    int rc = check_layout(info, p, l);
    if (rc) return rc;
    if (p->les_model != TOSCA_LES_SMAGORINSKY && p->les_model != TOSCA_LES_DSM && p->les_model != TOSCA_LES_DPASD) {
        g_err = "LES model not supported";
        return TOSCA_ERR_UNSUPPORTED;
    }
    const int64_t mx = l.mx, my = l.my, mz = l.mz, n = l.nloc();
    const V3 *ucat = (const V3 *)ucat_;
    const V3 *csi = (const V3 *)m->csi, *eta = (const V3 *)m->eta, *zet = (const V3 *)m->zet;
    const Real *aj = m->aj;
    std::vector<Real> ksg_local;
    Real *ksg = ksgs;
    if (!ksg) { ksg_local.assign(n, 0.0); ksg = ksg_local.data(); }
    for (int64_t c = 0; c < n; c++) { nut[c] = 0.0; ksg[c] = 0.0; }
    OMP_FOR
    for (int64_t k = 1; k < mz - 1; k++)
        for (int64_t j = 1; j < my - 1; j++)
            for (int64_t i = 1; i < mx - 1; i++) {
                const int64_t c = l.L(k, j, i);
                Grad g = du_center(l, ucat, i, j, k);
                Cart d = du_dxyz(csi[c], eta[c], zet[c], aj[c], g);
                Real S = strain_mag(d);
                Real filter = pow(1. / aj[c], 1. / 3.);
                nut[c] = Cs[c] * pow(filter, 2.0) * S;
                Real Ck = 0.094;
                Real V = 1.0 / aj[c];
                Real A_x = nMag(zet[c]), A_y = nMag(csi[c]), A_z = nMag(eta[c]);
                Real delta_x = V / A_x, delta_y = V / A_y, delta_z = V / A_z;
                Real invDelta2Sum = 1.0 / (delta_x * delta_x) + 1.0 / (delta_y * delta_y) + 1.0 / (delta_z * delta_z);
                Real delta2_harm = 3.0 / invDelta2Sum;
                ksg[c] = pow(nut[c], 2.0) / (pow(Ck, 2.0) * delta2_harm);
            }
    local_to_local(l, nut);
    local_to_local(l, ksg);
    // UpdateNutBCs (:2812-2917), face-adjacent ghosts of the interior cells
    const int *bc = p->bc_nut;
    const double *val = p->nut_value;
    for (int64_t k = 1; k < mz - 1; k++)
        for (int64_t j = 1; j < my - 1; j++)
            for (int64_t i = 1; i < mx - 1; i++) {
                const int64_t c = l.L(k, j, i);
                if (bc[0] == TOSCA_BCNUT_ZEROGRADIENT && i == 1) nut[l.L(k, j, i - 1)] = nut[c];
                if (bc[1] == TOSCA_BCNUT_ZEROGRADIENT && i == mx - 2) nut[l.L(k, j, i + 1)] = nut[c];
                if (bc[2] == TOSCA_BCNUT_ZEROGRADIENT && j == 1) nut[l.L(k, j - 1, i)] = nut[c];
                if (bc[3] == TOSCA_BCNUT_ZEROGRADIENT && j == my - 2) nut[l.L(k, j + 1, i)] = nut[c];
                if (bc[4] == TOSCA_BCNUT_ZEROGRADIENT && k == 1) nut[l.L(k - 1, j, i)] = nut[c];
                if (bc[5] == TOSCA_BCNUT_ZEROGRADIENT && k == mz - 2) nut[l.L(k + 1, j, i)] = nut[c];
                if (bc[0] == TOSCA_BCNUT_FIXEDVALUE && i == 1) nut[l.L(k, j, i - 1)] = val[0];
                if (bc[1] == TOSCA_BCNUT_FIXEDVALUE && i == mx - 2) nut[l.L(k, j, i + 1)] = val[1];
                if (bc[2] == TOSCA_BCNUT_FIXEDVALUE && j == 1) nut[l.L(k, j - 1, i)] = val[2];
                if (bc[3] == TOSCA_BCNUT_FIXEDVALUE && j == my - 2) nut[l.L(k, j + 1, i)] = val[3];
                if (bc[4] == TOSCA_BCNUT_FIXEDVALUE && k == 1) nut[l.L(k - 1, j, i)] = val[4];
                if (bc[5] == TOSCA_BCNUT_FIXEDVALUE && k == mz - 2) nut[l.L(k + 1, j, i)] = val[5];
                if (bc[0] == TOSCA_BCNUT_PERIODIC && i == 1) {
                    if (l.i_per()) nut[l.L(k, j, i - 1)] = nut[l.L(k, j, mx - 2)];
                    else if (l.ii_per()) nut[l.L(k, j, i - 1)] = nut[l.L(k, j, -2)];
                }
                if (bc[1] == TOSCA_BCNUT_PERIODIC && i == mx - 2) {
                    if (l.i_per()) nut[l.L(k, j, i + 1)] = nut[l.L(k, j, 1)];
                    else if (l.ii_per()) nut[l.L(k, j, i + 1)] = nut[l.L(k, j, mx + 1)];
                }
                if (bc[2] == TOSCA_BCNUT_PERIODIC && j == 1) {
                    if (l.j_per()) nut[l.L(k, j - 1, i)] = nut[l.L(k, my - 2, i)];
                    else if (l.jj_per()) nut[l.L(k, j - 1, i)] = nut[l.L(k, -2, i)];
                }
                if (bc[3] == TOSCA_BCNUT_PERIODIC && j == my - 2) {
                    if (l.j_per()) nut[l.L(k, j + 1, i)] = nut[l.L(k, 1, i)];
                    else if (l.jj_per()) nut[l.L(k, j + 1, i)] = nut[l.L(k, my + 1, i)];
                }
                if (bc[4] == TOSCA_BCNUT_PERIODIC && k == 1) {
                    if (l.k_per()) nut[l.L(k - 1, j, i)] = nut[l.L(mz - 2, j, i)];
                    else if (l.kk_per()) nut[l.L(k - 1, j, i)] = nut[l.L(-2, j, i)];
                }
                if (bc[5] == TOSCA_BCNUT_PERIODIC && k == mz - 2) {
                    if (l.k_per()) nut[l.L(k + 1, j, i)] = nut[l.L(1, j, i)];
                    else if (l.kk_per()) nut[l.L(k + 1, j, i)] = nut[l.L(mz + 1, j, i)];
                }
            }
    local_to_local(l, nut);
    return TOSCA_OK;
}

// ================================================================================================
// Buoyancy (src/sources/ueqn_sources.c:2726-2880)
// ================================================================================================
int ORACLE_API(oracle_buoyancy)(const tosca_dmda_info *info, const ORACLE_MESH *m, const tosca_params *p,
                                const Real *tmprt, double scale, Real *btheta_, Real *gcont_) {
    Lay l;
    int rc = check_layout(info, p, l);
    if (rc) return rc;
    const int64_t mx = l.mx, my = l.my, mz = l.mz;
    const V3 *icsi = (const V3 *)m->icsi, *jeta = (const V3 *)m->jeta, *kzet = (const V3 *)m->kzet;
    V3 *btheta = (V3 *)btheta_;
    std::vector<V3> gloc;
    V3 *gcont = (V3 *)gcont_;
    if (!gcont) { gloc.assign(mx * my * mz, V3{0, 0, 0}); gcont = gloc.data(); }
    const V3 gravity{0, 0, -9.81};
    const Real tRef = p->t_ref;
    for (int64_t k = 0; k < mz - 1; k++)
        for (int64_t j = 0; j < my - 1; j++)
            for (int64_t i = 0; i < mx - 1; i++) {
                const int64_t c = l.L(k, j, i), g = l.G(k, j, i);
                if (j > 0 && k > 0) gcont[g].x = (gravity.x * icsi[c].x + gravity.y * icsi[c].y + gravity.z * icsi[c].z);
                if (i > 0 && k > 0) gcont[g].y = (gravity.x * jeta[c].x + gravity.y * jeta[c].y + gravity.z * jeta[c].z);
                if (i > 0 && j > 0) gcont[g].z = (gravity.x * kzet[c].x + gravity.y * kzet[c].y + gravity.z * kzet[c].z);
            }
    for (int64_t c = 0; c < mx * my * mz; c++) btheta[c] = V3{0, 0, 0};
    for (int64_t k = 1; k < mz - 1; k++)
        for (int64_t j = 1; j < my - 1; j++)
            for (int64_t i = 1; i < mx - 1; i++) {
                const int64_t c = l.L(k, j, i), g = l.G(k, j, i);
                Real tempI = 0.5 * (tmprt[c] + tmprt[l.L(k, j, i + 1)]);
                Real tempJ = 0.5 * (tmprt[c] + tmprt[l.L(k, j + 1, i)]);
                Real tempK = 0.5 * (tmprt[c] + tmprt[l.L(k + 1, j, i)]);
                btheta[g].x += scale * (gcont[g].x * (tRef - tempI) / tRef);
                btheta[g].y += scale * (gcont[g].y * (tRef - tempJ) / tRef);
                btheta[g].z += scale * (gcont[g].z * (tRef - tempK) / tRef);
            }
    return TOSCA_OK;
}

#ifndef ORACLE_TANGENT
// One evaluation of the benchmark's unit of work, in the reference's call order:
// UpdateCs -> UpdateNut -> Buoyancy -> FormU(formMode 0) (src/main.c:111-112,181,185).
int oracle_rhs_les_step(const tosca_dmda_info *info, const tosca_mesh *m, const tosca_params *p,
                        const double *ucont, const double *ucat, const double *theta,
                        double *rhs, double scale, double *cs, double *nut, double *btheta,
                        double buoy_scale) {
    int rc = oracle_update_cs(info, m, p, ucat, cs);
    if (rc) return rc;
    rc = oracle_update_nut(info, m, p, ucat, cs, nut, nullptr);
    if (rc) return rc;
    rc = oracle_buoyancy(info, m, p, theta, buoy_scale, btheta, nullptr);
    if (rc) return rc;
    return oracle_formu(info, m, p, ucont, ucat, nut, rhs, scale, 0);
}
#endif  // ORACLE_TANGENT

// ================================================================================================
// ExplicitRhsU (src/solvers/ueqn_solvers.c:1276-1407), the terms of the momentum right-hand side
// this path covers.  Written over Real like FormU, so the tangent build differentiates it (and
// the explicit Wray Runge-Kutta predictor built on it, below).  On one rank a PETSc global vector is the owned part of its local vector, so
// the "globalToLocal" resets below act on the local arrays and the GlobalToLocal scatter that
// follows each of them is the periodic ghost fill of local_to_local.
// ================================================================================================

// resetNoPenetrationFluxes (src/include/inline.h:5262-5355): loops k < mz-1, j < my-1, i < mx-1
static void reset_no_penetration(const Lay &l, const tosca_params *p, V3 *ucont) {
    const int64_t mx = l.mx, my = l.my, mz = l.mz;
    const int *bc = p->bc_u;   // iLeft iRight jLeft jRight kLeft kRight
    auto zeroes = [](int b) { return b == TOSCA_BCU_NOSLIP || b == TOSCA_BCU_WALLFUNCTION || b == TOSCA_BCU_SLIP; };
    for (int64_t k = 0; k < mz - 1; k++)
        for (int64_t j = 0; j < my - 1; j++)
            for (int64_t i = 0; i < mx - 1; i++) {
                V3 &u = ucont[l.L(k, j, i)];
                // noSlip, velocityWallFunction, slip: no flux through the patch (:5290-5311)
                if (i == 0 && zeroes(bc[0])) u.x = 0.0;
                if (i == mx - 2 && zeroes(bc[1])) u.x = 0.0;
                if (j == 0 && zeroes(bc[2])) u.y = 0.0;
                if (j == my - 2 && zeroes(bc[3])) u.y = 0.0;
                if (k == 0 && zeroes(bc[4])) u.z = 0.0;
                if (k == mz - 2 && zeroes(bc[5])) u.z = 0.0;
                // zeroGradient: reverse (inflowing) flux removed (:5313-5336)
                if (bc[0] == TOSCA_BCU_ZEROGRADIENT && i == 0 && u.x > 0.0) u.x = 0.0;
                if (bc[1] == TOSCA_BCU_ZEROGRADIENT && i == mx - 2 && u.x < 0.0) u.x = 0.0;
                if (bc[2] == TOSCA_BCU_ZEROGRADIENT && j == 0 && u.y > 0.0) u.y = 0.0;
                if (bc[3] == TOSCA_BCU_ZEROGRADIENT && j == my - 2 && u.y < 0.0) u.y = 0.0;
                if (bc[4] == TOSCA_BCU_ZEROGRADIENT && k == 0 && u.z > 0.0) u.z = 0.0;
                if (bc[5] == TOSCA_BCU_ZEROGRADIENT && k == mz - 2 && u.z < 0.0) u.z = 0.0;
            }
    local_to_local(l, ucont);     // DMGlobalToLocal (:5343-5344)
}

// contravariantToCartesian (src/ueqn.c:528-635): interior cells, Cramer's rule on the cell basis
static void contravariant_to_cartesian(const Lay &l, const ORACLE_MESH *m, const V3 *lucont, V3 *ucat) {
    const V3 *csi = (const V3 *)m->csi, *eta = (const V3 *)m->eta, *zet = (const V3 *)m->zet;
    for (int64_t k = 1; k < l.mz - 1; k++)
        for (int64_t j = 1; j < l.my - 1; j++)
            for (int64_t i = 1; i < l.mx - 1; i++) {
                const int64_t c = l.L(k, j, i);
                Real mat[3][3] = {{csi[c].x, csi[c].y, csi[c].z}, {eta[c].x, eta[c].y, eta[c].z},
                                  {zet[c].x, zet[c].y, zet[c].z}};
                Real q[3];
                q[0] = 0.5 * (lucont[l.L(k, j, i - 1)].x + lucont[c].x);     // central (inline.h:4307-4312)
                q[1] = 0.5 * (lucont[l.L(k, j - 1, i)].y + lucont[c].y);
                q[2] = 0.5 * (lucont[l.L(k - 1, j, i)].z + lucont[c].z);
                const Real det = mat[0][0] * (mat[1][1] * mat[2][2] - mat[1][2] * mat[2][1]) -
                                 mat[0][1] * (mat[1][0] * mat[2][2] - mat[1][2] * mat[2][0]) +
                                 mat[0][2] * (mat[1][0] * mat[2][1] - mat[1][1] * mat[2][0]);
                const Real det0 = q[0] * (mat[1][1] * mat[2][2] - mat[1][2] * mat[2][1]) -
                                  q[1] * (mat[0][1] * mat[2][2] - mat[0][2] * mat[2][1]) +
                                  q[2] * (mat[0][1] * mat[1][2] - mat[0][2] * mat[1][1]);
                const Real det1 = -q[0] * (mat[1][0] * mat[2][2] - mat[1][2] * mat[2][0]) +
                                  q[1] * (mat[0][0] * mat[2][2] - mat[0][2] * mat[2][0]) -
                                  q[2] * (mat[0][0] * mat[1][2] - mat[0][2] * mat[1][0]);
                const Real det2 = q[0] * (mat[1][0] * mat[2][1] - mat[1][1] * mat[2][0]) -
                                  q[1] * (mat[0][0] * mat[2][1] - mat[0][1] * mat[2][0]) +
                                  q[2] * (mat[0][0] * mat[1][1] - mat[0][1] * mat[1][0]);
                ucat[c] = V3{det0 / det, det1 / det, det2 / det};
            }
    local_to_local(l, ucat);      // DMGlobalToLocal (:621-622)
}

// resetNonResolvedCellFaces (src/include/inline.h:1674-1838) on a global vector (no IBM / overset)
static void reset_nonresolved(const Lay &l, const tosca_params *p, V3 *v, const V3 *ucont_o) {
    const int64_t mx = l.mx, my = l.my, mz = l.mz;
    const int *bc = p->bc_u;
    const bool iper = l.per[0] != TOSCA_PERIODIC_NONE, jper = l.per[1] != TOSCA_PERIODIC_NONE,
               kper = l.per[2] != TOSCA_PERIODIC_NONE;
    for (int64_t k = 0; k < mz; k++)
        for (int64_t j = 0; j < my; j++)
            for (int64_t i = 0; i < mx; i++) {
                const int64_t g = l.G(k, j, i);
                const V3 uo = ucont_o[g];
                if (i == 0) { if (!(bc[0] == TOSCA_BCU_ZEROGRADIENT || uo.x < 0)) v[g].x = 0.0; }
                else if (i == mx - 1) v[g].x = 0.0;
                else if (i == mx - 2 && !iper) { if (!(bc[1] == TOSCA_BCU_ZEROGRADIENT || uo.x > 0)) v[g].x = 0.0; }
                if (j == 0) { if (!(bc[2] == TOSCA_BCU_ZEROGRADIENT || uo.y < 0)) v[g].y = 0.0; }
                else if (j == my - 1) v[g].y = 0.0;
                else if (j == my - 2 && !jper) { if (!(bc[3] == TOSCA_BCU_ZEROGRADIENT || uo.y > 0)) v[g].y = 0.0; }
                if (k == 0) { if (!(bc[4] == TOSCA_BCU_ZEROGRADIENT || uo.z < 0)) v[g].z = 0.0; }
                else if (k == mz - 1) v[g].z = 0.0;
                else if (k == mz - 2 && !kper) { if (!(bc[5] == TOSCA_BCU_ZEROGRADIENT || uo.z > 0)) v[g].z = 0.0; }
            }
}

// VecAXPY(y, a, x): y = y + a x on every owned point
static void axpy(int64_t n, double a, const V3 *x, V3 *y) {
    for (int64_t g = 0; g < n; g++) {
        y[g].x = y[g].x + a * x[g].x;
        y[g].y = y[g].y + a * x[g].y;
        y[g].z = y[g].z + a * x[g].z;
    }
}

int ORACLE_API(oracle_explicit_rhs_u)(const tosca_dmda_info *info, const ORACLE_MESH *m, const tosca_params *p,
                                      Real *ucont_, Real *ucat_, const Real *nut, const Real *dp,
                                      const Real *btheta, const Real *btheta_o, int ab2, const Real *farm,
                                      const Real *ucont_o, Real *rhs_, int formMode) {
    Lay l;
    int rc = check_layout(info, p, l);
    if (rc) return rc;
    if (!ucont_ || !ucat_ || !dp || !ucont_o || !rhs_ || (btheta && ab2 && !btheta_o)) {
        g_err = "explicit_rhs_u: missing array";
        return TOSCA_ERR_ARG;
    }
    V3 *ucont = (V3 *)ucont_, *ucat = (V3 *)ucat_, *rhs = (V3 *)rhs_;
    const int64_t ng = l.mx * l.my * l.mz;
    reset_no_penetration(l, p, ucont);                        // :1281
    reset_face_periodic(l, ucont);                            // :1284
    contravariant_to_cartesian(l, m, ucont, ucat);            // :1287
    reset_cell_periodic(l, ucat);                             // :1301
    for (int64_t g = 0; g < ng; g++) rhs[g] = V3{0, 0, 0};   // :1304
    axpy(ng, -1.0, (const V3 *)dp, rhs);                      // :1307
    if (btheta) {                                             // :1328-1336
        if (ab2) {
            axpy(ng, 3.0 / 2.0, (const V3 *)btheta, rhs);
            axpy(ng, -1.0 / 2.0, (const V3 *)btheta_o, rhs);
        } else {
            axpy(ng, 1.0, (const V3 *)btheta, rhs);
        }
    }
    if (farm) axpy(ng, 1.0, (const V3 *)farm, rhs);          // :1341-1344
    rc = ORACLE_API(oracle_formu)(info, m, p, ucont_, ucat_, nut, rhs_, 1.0, formMode);   // :1347
    if (rc) return rc;
    reset_nonresolved(l, p, rhs, (const V3 *)ucont_o);       // :1403
    return TOSCA_OK;
}

// ================================================================================================
// Explicit momentum predictors and the IMEX operator around ExplicitRhsU
// (src/solvers/ueqn_solvers.c:1411-1504, 1739-1801)
// ================================================================================================

// DMGlobalToLocal of a global vector on one rank: the owned range copied, then the type-2 wrap
static void global_vec_to_local(const Lay &l, const V3 *g, V3 *loc) {
    for (int64_t k = 0; k < l.mz; k++)
        for (int64_t j = 0; j < l.my; j++)
            for (int64_t i = 0; i < l.mx; i++) loc[l.L(k, j, i)] = g[l.G(k, j, i)];
    local_to_local(l, loc);
}

// the owned range of a local vector as a global vector
static void local_vec_to_global(const Lay &l, const V3 *loc, V3 *g) {
    for (int64_t k = 0; k < l.mz; k++)
        for (int64_t j = 0; j < l.my; j++)
            for (int64_t i = 0; i < l.mx; i++) g[l.G(k, j, i)] = loc[l.L(k, j, i)];
}

// Explicit Wray Runge-Kutta predictor: the three-stage low-storage scheme of UeqnRK3CN_W
// (src/solvers/ueqn_solvers.c:1508-1583) with its coefficients beta = (8/15, 5/12, 3/4),
// delta = (0, 17/60, 5/12) (:1518-1520) and the viscous term explicit as well, i.e. with the
// whole right-hand side in the beta / delta terms (beta_i - delta_i = 2 gamma_i).  Stage i:
//   Rhs = ExplicitRhsU(formMode) (its resets update Ucont in place);
//   bU = Ucont; bU += dt beta_i Rhs; bU += -1.0 dt delta_i Rhs_o  (VecCopy / VecAXPY order, :1533-1541);
//   Ucont = bU; GlobalToLocal; Rhs_o = Rhs (:1563-1569).
// n_stages stages run back to back; every third stage begins a time step, where the time loop sets
// Ucont_o = Ucont (the input of resetNonResolvedCellFaces) and the predictor sets Rhs_o = 0
// (:1522).  The sources and nu_t are held fixed.  ucont_o, rhs, rhs_o: global (rhs: last Rhs).
int ORACLE_API(oracle_ueqn_rk3w)(const tosca_dmda_info *info, const ORACLE_MESH *m, const tosca_params *p,
                                 Real *ucont_, Real *ucat_, const Real *nut, const Real *dp, const Real *btheta,
                                 const Real *btheta_o, int ab2, const Real *farm, double dt, int n_stages,
                                 Real *ucont_o_, Real *rhs_, Real *rhs_o_, int formMode) {
    Lay l;
    int rc = check_layout(info, p, l);
    if (rc) return rc;
    if (!ucont_ || !ucat_ || !ucont_o_ || !rhs_ || !rhs_o_ || n_stages < 0) {
        g_err = "ueqn_rk3w: missing array or negative stage count";
        return TOSCA_ERR_ARG;
    }
    const double beta[3] = {8.0 / 15.0, 5.0 / 12.0, 3.0 / 4.0};
    const double delta[3] = {0.0, 17.0 / 60.0, 5.0 / 12.0};
    const int64_t ng = l.mx * l.my * l.mz;
    V3 *ucont = (V3 *)ucont_, *rhs = (V3 *)rhs_, *rhs_o = (V3 *)rhs_o_, *uo = (V3 *)ucont_o_;
    std::vector<V3> bU(ng);
    for (int s = 0; s < n_stages; s++) {
        const int i = s % 3;
        if (i == 0) {
            local_vec_to_global(l, ucont, uo);
            for (int64_t g = 0; g < ng; g++) rhs_o[g] = V3{0.0, 0.0, 0.0};
        }
        rc = ORACLE_API(oracle_explicit_rhs_u)(info, m, p, ucont_, ucat_, nut, dp, btheta, btheta_o, ab2, farm,
                                               ucont_o_, rhs_, formMode);
        if (rc) return rc;
        local_vec_to_global(l, ucont, bU.data());
        axpy(ng, dt * beta[i], rhs, bU.data());
        axpy(ng, -1.0 * dt * delta[i], rhs_o, bU.data());
        global_vec_to_local(l, bU.data(), ucont);
        for (int64_t g = 0; g < ng; g++) rhs_o[g] = rhs[g];
    }
    return TOSCA_OK;
}

#ifndef ORACLE_TANGENT

// UeqnFE (:1411-1441): Rhs = ExplicitRhsU; VecScale(Rhs, dt); VecAXPY(Rhs, 1, Ucont_o);
// Ucont = Rhs; GlobalToLocal.  rhs holds the new Ucont on return, as the reference's Rhs does.
int oracle_ueqn_fe(const tosca_dmda_info *info, const tosca_mesh *m, const tosca_params *p, double *ucont_,
                   double *ucat_, const double *nut, const double *dp, const double *btheta, const double *btheta_o,
                   int ab2, const double *farm, const double *ucont_o, double dt, double *rhs_, int formMode) {
    int rc = oracle_explicit_rhs_u(info, m, p, ucont_, ucat_, nut, dp, btheta, btheta_o, ab2, farm, ucont_o, rhs_,
                                   formMode);
    if (rc) return rc;
    Lay l;
    check_layout(info, p, l);
    V3 *rhs = (V3 *)rhs_;
    const int64_t ng = l.mx * l.my * l.mz;
    for (int64_t g = 0; g < ng; g++) {
        rhs[g].x = dt * rhs[g].x;
        rhs[g].y = dt * rhs[g].y;
        rhs[g].z = dt * rhs[g].z;
    }
    axpy(ng, 1.0, (const V3 *)ucont_o, rhs);
    global_vec_to_local(l, rhs, (V3 *)ucont_);
    return TOSCA_OK;
}

// UeqnRK4 (:1446-1504): classical four-stage Runge-Kutta on ExplicitRhsU with the sources held
// fixed; stage i > 0 starts from Ucont = a_i dt Rhs + Ucont_o (VecWAXPY), and
// Utmp = Ucont_o + sum_i b_i dt K_i is accumulated with VecAXPY.  utmp is a global work vector.
int oracle_ueqn_rk4(const tosca_dmda_info *info, const tosca_mesh *m, const tosca_params *p, double *ucont_,
                    double *ucat_, const double *nut, const double *dp, const double *btheta, const double *btheta_o,
                    int ab2, const double *farm, const double *ucont_o_, double dt, double *rhs_, double *utmp_,
                    int formMode) {
    Lay l;
    int rc = check_layout(info, p, l);
    if (rc) return rc;
    if (!utmp_) { g_err = "ueqn_rk4: missing utmp"; return TOSCA_ERR_ARG; }
    const double b[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};
    const double a[4] = {0.0, 0.5, 0.5, 1.0};
    const int64_t ng = l.mx * l.my * l.mz;
    V3 *rhs = (V3 *)rhs_, *utmp = (V3 *)utmp_;
    const V3 *uo = (const V3 *)ucont_o_;
    for (int64_t g = 0; g < ng; g++) utmp[g] = uo[g];
    std::vector<V3> w(ng);
    for (int i = 0; i < 4; i++) {
        if (i != 0) {
            const double al = a[i] * dt;
            for (int64_t g = 0; g < ng; g++) {
                w[g].x = al * rhs[g].x + uo[g].x;
                w[g].y = al * rhs[g].y + uo[g].y;
                w[g].z = al * rhs[g].z + uo[g].z;
            }
            global_vec_to_local(l, w.data(), (V3 *)ucont_);
        }
        rc = oracle_explicit_rhs_u(info, m, p, ucont_, ucat_, nut, dp, btheta, btheta_o, ab2, farm, ucont_o_, rhs_,
                                   formMode);
        if (rc) return rc;
        axpy(ng, dt * b[i], rhs, utmp);
    }
    global_vec_to_local(l, utmp, (V3 *)ucont_);
    return TOSCA_OK;
}

// IMEXMatVec (:1739-1801): Ucont = v and its ghosts; the ExplicitRhsU resets and
// contravariantToCartesian; Av = 0; FormU(Av, -coef, 2) with coef = dt scale;
// resetNonResolvedCellFaces(Av); VecAXPY(Av, 1, v)
int oracle_imex_matvec(const tosca_dmda_info *info, const tosca_mesh *m, const tosca_params *p, const double *v_,
                       double *ucont_, double *ucat_, const double *nut, const double *ucont_o, double coef,
                       double *av_) {
    Lay l;
    int rc = check_layout(info, p, l);
    if (rc) return rc;
    if (!v_ || !ucont_ || !ucat_ || !ucont_o || !av_) { g_err = "imex_matvec: missing array"; return TOSCA_ERR_ARG; }
    const int64_t ng = l.mx * l.my * l.mz;
    V3 *ucont = (V3 *)ucont_, *ucat = (V3 *)ucat_, *av = (V3 *)av_;
    global_vec_to_local(l, (const V3 *)v_, ucont);            // :1763-1767
    reset_no_penetration(l, p, ucont);                        // :1770
    reset_face_periodic(l, ucont);                            // :1773
    contravariant_to_cartesian(l, m, ucont, ucat);            // :1777
    reset_cell_periodic(l, ucat);                             // :1790
    for (int64_t g = 0; g < ng; g++) av[g] = V3{0, 0, 0};    // :1793
    rc = oracle_formu(info, m, p, ucont_, ucat_, nut, av_, -1.0 * coef, 2);   // :1794
    if (rc) return rc;
    reset_nonresolved(l, p, av, (const V3 *)ucont_o);         // :1795
    axpy(ng, 1.0, (const V3 *)v_, av);                        // :1796
    return TOSCA_OK;
}

// ================================================================================================
// The Crank-Nicolson momentum solve's residual (UeqnSNES / SNESFuncEval, ueqn_solvers.c:1087-1272)
// ================================================================================================

// UeqnSNES's constant right-hand side (:1099-1150, the terms of this path): bU = -dP, the bTheta
// terms (AB2 when ab2), + 0.5 Rhs_o when ab2 (Crank-Nicolson half), + farm
int oracle_cn_rhs(const tosca_dmda_info *info, const tosca_params *p, const double *dp, const double *btheta,
                  const double *btheta_o, int ab2, const double *rhs_o, const double *farm, double *bU_) {
    Lay l;
    int rc = check_layout(info, p, l);
    if (rc) return rc;
    if (!dp || !bU_ || (btheta && ab2 && !btheta_o) || (ab2 && !rhs_o)) { g_err = "cn_rhs: missing array"; return TOSCA_ERR_ARG; }
    const int64_t ng = l.mx * l.my * l.mz;
    V3 *bU = (V3 *)bU_;
    const V3 *d = (const V3 *)dp;
    for (int64_t g = 0; g < ng; g++) {                                  // :1102-1103 VecCopy + VecScale(-1)
        bU[g].x = -1.0 * d[g].x; bU[g].y = -1.0 * d[g].y; bU[g].z = -1.0 * d[g].z;
    }
    if (btheta) {                                                       // :1125-1135
        if (ab2) {
            axpy(ng, 3.0 / 2.0, (const V3 *)btheta, bU);
            axpy(ng, -1.0 / 2.0, (const V3 *)btheta_o, bU);
        } else {
            axpy(ng, 1.0, (const V3 *)btheta, bU);
        }
    }
    if (ab2) axpy(ng, 0.5, (const V3 *)rhs_o, bU);                      // :1151-1152
    if (farm) axpy(ng, 1.0, (const V3 *)farm, bU);                      // :1155-1156
    return TOSCA_OK;
}

// SNESFuncEval (:1189-1272): F(U) = dt (bU + s FormU(U)) with the non-resolved faces reset, - U +
// Ucont_o; s = 1 on the first step (backward Euler), 0.5 after (Crank-Nicolson).  u is the SNES
// iterate (global); ucont, ucat are lUcont, lUcat (overwritten, as the reference overwrites them).
int oracle_snes_func_eval(const tosca_dmda_info *info, const tosca_mesh *m, const tosca_params *p, const double *u_,
                          double *ucont_, double *ucat_, const double *nut, const double *bU, const double *ucont_o,
                          double dt, int ab2, double *F_) {
    Lay l;
    int rc = check_layout(info, p, l);
    if (rc) return rc;
    if (!u_ || !ucont_ || !ucat_ || !bU || !ucont_o || !F_) { g_err = "snes_func_eval: missing array"; return TOSCA_ERR_ARG; }
    const int64_t ng = l.mx * l.my * l.mz;
    V3 *ucont = (V3 *)ucont_, *ucat = (V3 *)ucat_, *F = (V3 *)F_;
    global_vec_to_local(l, (const V3 *)u_, ucont);                       // :1195-1199
    reset_no_penetration(l, p, ucont);                                   // :1202
    reset_face_periodic(l, ucont);                                       // :1205
    contravariant_to_cartesian(l, m, ucont, ucat);                       // :1208
    reset_cell_periodic(l, ucat);                                        // :1221
    for (int64_t g = 0; g < ng; g++) F[g] = ((const V3 *)bU)[g];        // :1224
    const double scale = ab2 ? 0.5 : 1.0;                                // :1227
    if ((rc = oracle_formu(info, m, p, ucont_, ucat_, nut, F_, scale, 0))) return rc;   // :1230
    for (int64_t g = 0; g < ng; g++) {                                   // :1264 VecScale(Rhs, dt)
        F[g].x = dt * F[g].x; F[g].y = dt * F[g].y; F[g].z = dt * F[g].z;
    }
    reset_nonresolved(l, p, F, (const V3 *)ucont_o);                     // :1267
    axpy(ng, -1.0, (const V3 *)u_, F);                                  // :1270
    axpy(ng, 1.0, (const V3 *)ucont_o, F);                              // :1271
    return TOSCA_OK;
}

// ================================================================================================
// GMRES(m) as PETSc 3.15 runs it for kspIMEX (src/ueqn.c:261-282: KSPGMRES, PCNONE, nonzero
// initial guess, KSPSetTolerances(rtol, atol, 1e30, max_it)).  PETSc is a dependency absent from
// the reference tree; this restates its published algorithm (src/ksp/ksp/impls/gmres/gmres.c:
// KSPSolve_GMRES, KSPGMRESCycle, KSPGMRESUpdateHessenberg, KSPGMRESBuildSoln, with the default
// classical Gram-Schmidt orthogonalisation without refinement; KSPConvergedDefault).  kspIMEX has a
// nonzero initial guess and no preconditioner (src/ueqn.c:264-276), so KSPConvergedDefault takes the
// relative and divergence tests against ||b|| (||r0|| when b = 0): ttol = max(rtol ||b||, atol),
// dtol 1e30 (ueqn.c:277).  Reasons are PETSc's KSPConvergedReason values: 2 rtol, 3 atol, -2 null
// rotation (KSP_DIVERGED_NULL), -3 iteration limit, -4 dtol, -5 breakdown, -9 NaN or Inf.
// ================================================================================================
static double vdot(int64_t n, const double *a, const double *b) {
    double s = 0.0;
    for (int64_t i = 0; i < n; i++) s += a[i] * b[i];
    return s;
}

static int gmres(int64_t n, const std::function<int(const double *, double *)> &A, const double *b, double *x,
                 int restart, double rtol, double atol, int maxit, int *its_out, double *res_out, int *conv) {
    const int m = restart;
    std::vector<double> V((size_t)(m + 1) * n), tmp(n);
    std::vector<double> HH((size_t)(m + 2) * (m + 1), 0.0), cc(m + 1), ss(m + 1), grs(m + 2), nrs(m + 1);
    auto H = [&](int a, int c) -> double & { return HH[(size_t)c * (m + 2) + a]; };
    int its = 0, reason = 0;
    double ttol = 0.0, res = 0.0, rnorm0 = 0.0;
    const double bnorm = std::sqrt(vdot(n, b, b)), divtol = 1e30;
    bool first = true;
    // KSPConvergedDefault: NaN / Inf, then the tolerance (atol when below it, else rtol), then dtol
    auto test = [&](double r) {
        if (std::isnan(r) || std::isinf(r)) return -9;
        if (r <= ttol) return r < atol ? 3 : 2;
        if (r >= divtol * rnorm0) return -4;
        return 0;
    };
    while (!reason) {
        // KSPInitialResidual: r = b - A x into V_0, then VecNormalize
        int rc = A(x, tmp.data());
        if (rc) return rc;
        double *v0 = V.data();
        for (int64_t i = 0; i < n; i++) v0[i] = b[i] + (-1.0) * tmp[i];
        res = std::sqrt(vdot(n, v0, v0));
        if (res > 0.0) {
            const double sc = 1.0 / res;
            for (int64_t i = 0; i < n; i++) v0[i] = v0[i] * sc;
        }
        grs[0] = res;
        if (first) {
            rnorm0 = bnorm > 0.0 ? bnorm : res;           // nonzero guess: ||b||, or ||r0|| for b = 0
            ttol = std::max(rtol * rnorm0, atol);
            first = false;
        }
        reason = test(res);
        if (reason < 0) break;
        int it = 0;
        bool hapend = false;
        while (!reason && it < m && its < maxit) {
            double *w = V.data() + (size_t)(it + 1) * n;
            rc = A(V.data() + (size_t)it * n, w);
            if (rc) return rc;
            // classical Gram-Schmidt: all dots first (VecMDot), then one update (VecMAXPY)
            std::vector<double> h(it + 1);
            for (int j = 0; j <= it; j++) h[j] = vdot(n, w, V.data() + (size_t)j * n);
            for (int j = 0; j <= it; j++) {
                const double a = -h[j];
                const double *vj = V.data() + (size_t)j * n;
                for (int64_t i = 0; i < n; i++) w[i] = w[i] + a * vj[i];
                H(j, it) = h[j];
            }
            const double tt = std::sqrt(vdot(n, w, w));
            H(it + 1, it) = tt;
            double hapbnd = std::fabs(tt / grs[it]);
            if (hapbnd > 1e-30) hapbnd = 1e-30;
            hapend = tt < hapbnd;
            if (!hapend) {
                const double sc = 1.0 / tt;
                for (int64_t i = 0; i < n; i++) w[i] = w[i] * sc;
            }
            // KSPGMRESUpdateHessenberg: previous rotations on the new column, then a new one
            for (int j = 1; j <= it; j++) {
                const double t0 = H(j - 1, it);
                H(j - 1, it) = cc[j - 1] * t0 + ss[j - 1] * H(j, it);
                H(j, it) = cc[j - 1] * H(j, it) - ss[j - 1] * t0;
            }
            if (!hapend) {
                const double t1 = std::sqrt(H(it, it) * H(it, it) + H(it + 1, it) * H(it + 1, it));
                if (t1 == 0.0) { reason = -2; break; }
                cc[it] = H(it, it) / t1;
                ss[it] = H(it + 1, it) / t1;
                grs[it + 1] = -(ss[it] * grs[it]);
                grs[it] = cc[it] * grs[it];
                H(it, it) = cc[it] * H(it, it) + ss[it] * H(it + 1, it);
                res = std::fabs(grs[it + 1]);
            } else {
                res = 0.0;
            }
            it++;
            its++;
            reason = test(res);
            if (hapend) { if (!reason) reason = -5; break; }
        }
        // KSPGMRESBuildSoln over columns 0 .. it-1
        const int last = it - 1;
        if (last >= 0) {
            if (H(last, last) != 0.0) nrs[last] = grs[last] / H(last, last);
            else nrs[last] = 0.0;
            for (int ii = 1; ii <= last; ii++) {
                const int k = last - ii;
                double tt = grs[k];
                for (int j = k + 1; j <= last; j++) tt = tt - H(k, j) * nrs[j];
                nrs[k] = tt / H(k, k);
            }
            for (int64_t i = 0; i < n; i++) tmp[i] = 0.0;
            for (int j = 0; j <= last; j++) {
                const double *vj = V.data() + (size_t)j * n;
                for (int64_t i = 0; i < n; i++) tmp[i] = tmp[i] + nrs[j] * vj[i];
            }
            for (int64_t i = 0; i < n; i++) x[i] = x[i] + 1.0 * tmp[i];
        }
        if (its >= maxit && !reason) reason = -3;
    }
    if (its_out) *its_out = its;
    if (res_out) *res_out = res;
    if (conv) *conv = reason;
    return TOSCA_OK;
}

// GMRES on a dense row-major n x n matrix (test infrastructure for the solver itself)
int oracle_gmres_dense(int64_t n, const double *Am, const double *b, double *x, int restart, double rtol, double atol,
                       int maxit, int *its, double *res, int *conv) {
    auto A = [&](const double *v, double *out) {
        for (int64_t r = 0; r < n; r++) {
            double s = 0.0;
            for (int64_t c = 0; c < n; c++) s += Am[r * n + c] * v[c];
            out[r] = s;
        }
        return 0;
    };
    return gmres(n, A, b, x, restart, rtol, atol, maxit, its, res, conv);
}

// KSPSolve(kspIMEX, b, x) with IMEXMatVec as the operator; ucont, ucat are its work arrays
int oracle_imex_solve(const tosca_dmda_info *info, const tosca_mesh *m, const tosca_params *p, const double *b,
                      double *x, double *ucont, double *ucat, const double *nut, const double *ucont_o, double coef,
                      int restart, double rtol, double atol, int maxit, int *its, double *res, int *conv) {
    Lay l;
    int rc = check_layout(info, p, l);
    if (rc) return rc;
    const int64_t n = 3 * l.mx * l.my * l.mz;
    auto A = [&](const double *v, double *out) {
        return oracle_imex_matvec(info, m, p, v, ucont, ucat, nut, ucont_o, coef, out);
    };
    return gmres(n, A, b, x, restart, rtol, atol, maxit, its, res, conv);
}

// UeqnRK3CN_W (scheme 0, ueqn_solvers.c:1508-1583) and UeqnRK3CN_SO (scheme 1, :1586-1661): three
// IMEX stages, convection and the other explicit terms from ExplicitRhsU(formMode 1), diffusion
// Crank-Nicolson-like through the IMEX solve with coef = dt gamma_i.  its / res / conv report the
// last stage's solve, as the reference prints them.
int oracle_ueqn_rk3cn(const tosca_dmda_info *info, const tosca_mesh *m, const tosca_params *p, int scheme,
                      double *ucont_, double *ucat_, const double *nut, const double *dp, const double *btheta,
                      const double *btheta_o, int ab2, const double *farm, const double *ucont_o_, double dt,
                      double *rhs_, int restart, double rtol, double atol, int maxit, int *its, double *res,
                      int *conv) {
    Lay l;
    int rc = check_layout(info, p, l);
    if (rc) return rc;
    if (scheme != 0 && scheme != 1) { g_err = "ueqn_rk3cn: scheme must be 0 (Wray) or 1 (Shu-Osher)"; return TOSCA_ERR_ARG; }
    const int64_t ng = l.mx * l.my * l.mz;
    V3 *rhs = (V3 *)rhs_;
    const V3 *uo = (const V3 *)ucont_o_;
    std::vector<V3> rhs_o(ng, V3{0, 0, 0}), visc(ng), bU(ng), utmp(ng), ucg(ng);
    const double wb[3] = {8.0 / 15.0, 5.0 / 12.0, 3.0 / 4.0}, wg[3] = {4.0 / 15.0, 1.0 / 15.0, 1.0 / 6.0},
                 wd[3] = {0.0, 17.0 / 60.0, 5.0 / 12.0};
    const double sa[3] = {1.0, 3.0 / 4.0, 1.0 / 3.0}, sb[3] = {1.0, 1.0 / 4.0, 2.0 / 3.0},
                 sg[3] = {1.0 / 2.0, 1.0 / 8.0, 1.0 / 3.0};
    for (int i = 0; i < 3; i++) {
        rc = oracle_explicit_rhs_u(info, m, p, ucont_, ucat_, nut, dp, btheta, btheta_o, ab2, farm, ucont_o_, rhs_, 1);
        if (rc) return rc;
        for (int64_t g = 0; g < ng; g++) visc[g] = V3{0, 0, 0};
        rc = oracle_formu(info, m, p, ucont_, ucat_, nut, (double *)visc.data(), 1.0, 2);
        if (rc) return rc;
        reset_nonresolved(l, p, visc.data(), uo);
        for (int64_t k = 0; k < l.mz; k++)             // the global Ucont: owned range of lUcont
            for (int64_t j = 0; j < l.my; j++)
                for (int64_t ii = 0; ii < l.mx; ii++) ucg[l.G(k, j, ii)] = ((const V3 *)ucont_)[l.L(k, j, ii)];
        double gam;
        if (scheme == 0) {
            bU = ucg;                                               // :1539
            axpy(ng, dt * wb[i], rhs, bU.data());                  // :1542
            axpy(ng, -1.0 * dt * wd[i], rhs_o.data(), bU.data());  // :1545
            axpy(ng, dt * wg[i], visc.data(), bU.data());          // :1548
            gam = wg[i];
        } else {
            for (int64_t g = 0; g < ng; g++) {                      // :1616-1617 VecCopy + VecScale
                bU[g].x = sa[i] * uo[g].x; bU[g].y = sa[i] * uo[g].y; bU[g].z = sa[i] * uo[g].z;
            }
            if (i > 0) axpy(ng, 1.0 - sa[i], ucg.data(), bU.data());   // :1620-1624
            axpy(ng, dt * sb[i], rhs, bU.data());                   // :1627
            axpy(ng, dt * sg[i], visc.data(), bU.data());           // :1630
            gam = sg[i];
        }
        utmp = bU;                                                  // initial guess
        axpy(ng, dt * gam, visc.data(), utmp.data());
        rc = oracle_imex_solve(info, m, p, (const double *)bU.data(), (double *)utmp.data(), ucont_, ucat_, nut,
                               ucont_o_, dt * gam, restart, rtol, atol, maxit, its, res, conv);
        if (rc) return rc;
        global_vec_to_local(l, utmp.data(), (V3 *)ucont_);
        if (scheme == 0) rhs_o.assign(rhs, rhs + ng);
    }
    return TOSCA_OK;
}

// UeqnABCN (ueqn_solvers.c:1665-1737): Adams-Bashforth convection, Crank-Nicolson diffusion.
// `ab2` is clock->it > clock->itStart, for the bTheta terms of ExplicitRhsU and for the scheme;
// conv_o (RhsConv_o, global) carries the convective term from the previous step and is replaced by
// this step's.  On the first step the reference adds Rhs to bU = Rhs a second time (:1694-1695);
// that is reproduced.
int oracle_ueqn_abcn(const tosca_dmda_info *info, const tosca_mesh *m, const tosca_params *p, double *ucont_,
                     double *ucat_, const double *nut, const double *dp, const double *btheta, const double *btheta_o,
                     int ab2, const double *farm, const double *ucont_o_, double dt, double *rhs_, double *conv_o_,
                     int restart, double rtol, double atol, int maxit, int *its, double *res, int *conv) {
    Lay l;
    int rc = check_layout(info, p, l);
    if (rc) return rc;
    if (!conv_o_) { g_err = "ueqn_abcn: missing conv_o"; return TOSCA_ERR_ARG; }
    const int64_t ng = l.mx * l.my * l.mz;
    V3 *rhs = (V3 *)rhs_, *conv_o = (V3 *)conv_o_;
    const V3 *uo = (const V3 *)ucont_o_;
    std::vector<V3> visc(ng, V3{0, 0, 0}), cnv(ng, V3{0, 0, 0}), bU(ng), utmp(ng);
    rc = oracle_explicit_rhs_u(info, m, p, ucont_, ucat_, nut, dp, btheta, btheta_o, ab2, farm, ucont_o_, rhs_, 1);
    if (rc) return rc;                                                   // :1675
    if ((rc = oracle_formu(info, m, p, ucont_, ucat_, nut, (double *)visc.data(), 1.0, 2))) return rc;   // :1678
    if ((rc = oracle_formu(info, m, p, ucont_, ucat_, nut, (double *)cnv.data(), 1.0, 1))) return rc;    // :1679
    reset_nonresolved(l, p, visc.data(), uo);                           // :1680-1681
    reset_nonresolved(l, p, cnv.data(), uo);
    bU.assign(rhs, rhs + ng);                                           // :1684
    if (ab2) {
        axpy(ng, 0.5, cnv.data(), bU.data());                           // :1690
        axpy(ng, -0.5, conv_o, bU.data());                              // :1691
    } else {
        axpy(ng, 1.0, rhs, bU.data());                                  // :1695
    }
    if (ab2) axpy(ng, 0.5, visc.data(), bU.data());                     // :1699-1700
    for (int64_t g = 0; g < ng; g++) {                                  // :1703 VecScale
        bU[g].x = dt * bU[g].x; bU[g].y = dt * bU[g].y; bU[g].z = dt * bU[g].z;
    }
    axpy(ng, 1.0, uo, bU.data());                                       // :1706
    for (int64_t g = 0; g < ng; g++) utmp[g] = V3{0, 0, 0};            // :1709-1713
    if (ab2) axpy(ng, 0.5 * dt, visc.data(), utmp.data());
    else axpy(ng, dt, visc.data(), utmp.data());
    axpy(ng, 1.0, bU.data(), utmp.data());
    const double coef = dt * (ab2 ? 0.5 : 1.0);                         // IMEXMatVec's scale (:1751-1759)
    rc = oracle_imex_solve(info, m, p, (const double *)bU.data(), (double *)utmp.data(), ucont_, ucat_, nut,
                           ucont_o_, coef, restart, rtol, atol, maxit, its, res, conv);
    if (rc) return rc;
    global_vec_to_local(l, utmp.data(), (V3 *)ucont_);                  // :1723-1725
    for (int64_t g = 0; g < ng; g++) conv_o[g] = cnv[g];                // :1730
    return TOSCA_OK;
}
#endif  // ORACLE_TANGENT

#ifndef ORACLE_TANGENT
// ================================================================================================
// GradP (src/peqn.c:3437-3773), meshTag = 0: the contravariant pressure-gradient fluxes
// ================================================================================================
int oracle_grad_p(const tosca_dmda_info *info, const tosca_mesh *m, const tosca_params *p_, const double *P,
                  double *dp_) {
    Lay l;
    int rc = check_layout(info, p_, l);
    if (rc) return rc;
    if (!P || !dp_) { g_err = "grad_p: missing array"; return TOSCA_ERR_ARG; }
    const int64_t mx = l.mx, my = l.my, mz = l.mz;
    const V3 *icsi = (const V3 *)m->icsi, *ieta = (const V3 *)m->ieta, *izet = (const V3 *)m->izet;
    const V3 *jcsi = (const V3 *)m->jcsi, *jeta = (const V3 *)m->jeta, *jzet = (const V3 *)m->jzet;
    const V3 *kcsi = (const V3 *)m->kcsi, *keta = (const V3 *)m->keta, *kzet = (const V3 *)m->kzet;
    V3 *dp = (V3 *)dp_;
    auto p = [&](int64_t k, int64_t j, int64_t i) { return P[l.L(k, j, i)]; };
    auto dot = [](const V3 &a, const V3 &b) { return a.x * b.x + a.y * b.y + a.z * b.z; };
    const bool iper = l.i_per() || l.ii_per(), jper = l.j_per() || l.jj_per(), kper = l.k_per() || l.kk_per();
    for (int64_t g = 0; g < mx * my * mz; g++) dp[g] = V3{0, 0, 0};          // :3468
    for (int64_t k = 1; k < mz - 1; k++)
        for (int64_t j = 1; j < my - 1; j++)
            for (int64_t i = 1; i < mx - 1; i++) {
                const int64_t c = l.L(k, j, i);
                // g_ab: face basis b dotted with the face's own normal area vector (:3492-3500)
                const double g11_i = dot(icsi[c], icsi[c]), g12_i = dot(ieta[c], icsi[c]), g13_i = dot(izet[c], icsi[c]);
                const double g21_j = dot(jcsi[c], jeta[c]), g22_j = dot(jeta[c], jeta[c]), g23_j = dot(jzet[c], jeta[c]);
                const double g31_k = dot(kcsi[c], kzet[c]), g32_k = dot(keta[c], kzet[c]), g33_k = dot(kzet[c], kzet[c]);
                double dpdc, dpde, dpdz;
                // i face (:3502-3566); the one-sided corner rows sit at i == mx-2 whatever the periodicity
                dpdc = (i == mx - 2 && l.ii_per()) ? p(k, j, mx + 1) - p(k, j, i) : p(k, j, i + 1) - p(k, j, i);
                if (j == my - 2 && i == mx - 2) dpde = (p(k, j, i) - p(k, j - 1, i) + p(k, j, i + 1) - p(k, j - 1, i + 1)) * 0.5;
                else if (j == 1 && i == mx - 2) dpde = (p(k, j + 1, i) - p(k, j, i) + p(k, j + 1, i + 1) - p(k, j, i + 1)) * 0.5;
                else dpde = (p(k, j + 1, i) - p(k, j - 1, i) + p(k, j + 1, i + 1) - p(k, j - 1, i + 1)) * 0.25;
                if (k == mz - 2 && i == mx - 2) dpdz = (p(k, j, i) - p(k - 1, j, i) + p(k, j, i + 1) - p(k - 1, j, i + 1)) * 0.5;
                else if (k == 1 && i == mx - 2) dpdz = (p(k + 1, j, i) - p(k, j, i) + p(k + 1, j, i + 1) - p(k, j, i + 1)) * 0.5;
                else dpdz = (p(k + 1, j, i) - p(k - 1, j, i) + p(k + 1, j, i + 1) - p(k - 1, j, i + 1)) * 0.25;
                dp[l.G(k, j, i)].x = (dpdc * g11_i + dpde * g12_i + dpdz * g13_i) * m->iaj[c];
                // j face (:3568-3651)
                if (i == mx - 2 && j == my - 2) dpdc = (p(k, j, i) - p(k, j, i - 1) + p(k, j + 1, i) - p(k, j + 1, i - 1)) * 0.5;
                else if (i == 1 && j == my - 2) dpdc = (p(k, j, i + 1) - p(k, j, i) + p(k, j + 1, i + 1) - p(k, j + 1, i)) * 0.5;
                else dpdc = (p(k, j, i + 1) - p(k, j, i - 1) + p(k, j + 1, i + 1) - p(k, j + 1, i - 1)) * 0.25;
                dpde = (j == my - 2 && l.jj_per()) ? p(k, my + 1, i) - p(k, j, i) : p(k, j + 1, i) - p(k, j, i);
                if (k == mz - 2 && j == my - 2) dpdz = (p(k, j, i) - p(k - 1, j, i) + p(k, j + 1, i) - p(k - 1, j + 1, i)) * 0.5;
                else if (k == 1 && j == my - 2) dpdz = (p(k + 1, j, i) - p(k, j, i) + p(k + 1, j + 1, i) - p(k, j + 1, i)) * 0.5;
                else dpdz = (p(k + 1, j, i) - p(k - 1, j, i) + p(k + 1, j + 1, i) - p(k - 1, j + 1, i)) * 0.25;
                dp[l.G(k, j, i)].y = (dpdc * g21_j + dpde * g22_j + dpdz * g23_j) * m->jaj[c];
                // k face (:3653-3734)
                if (i == mx - 2 && k == mz - 2) dpdc = (p(k, j, i) - p(k, j, i - 1) + p(k + 1, j, i) - p(k + 1, j, i - 1)) * 0.5;
                else if (i == 1 && k == mz - 2) dpdc = (p(k, j, i + 1) - p(k, j, i) + p(k + 1, j, i + 1) - p(k + 1, j, i)) * 0.5;
                else dpdc = (p(k, j, i + 1) - p(k, j, i - 1) + p(k + 1, j, i + 1) - p(k + 1, j, i - 1)) * 0.25;
                if (j == my - 2 && k == mz - 2) dpde = (p(k, j, i) - p(k, j - 1, i) + p(k + 1, j, i) - p(k + 1, j - 1, i)) * 0.5;
                else if (j == 1 && k == mz - 2) dpde = (p(k, j + 1, i) - p(k, j, i) + p(k + 1, j + 1, i) - p(k + 1, j, i)) * 0.5;
                else dpde = (p(k, j + 1, i) - p(k, j - 1, i) + p(k + 1, j + 1, i) - p(k + 1, j - 1, i)) * 0.25;
                dpdz = (k == mz - 2 && l.kk_per()) ? p(mz + 1, j, i) - p(k, j, i) : (p(k + 1, j, i) - p(k, j, i));
                dp[l.G(k, j, i)].z = (dpdc * g31_k + dpde * g32_k + dpdz * g33_k) * m->kaj[c];
                // faces whose contravariant velocity is not solved (:3736-3757)
                if (!iper && i == mx - 2) dp[l.G(k, j, i)].x = 0;
                if (!jper && j == my - 2) dp[l.G(k, j, i)].y = 0;
                if (!kper && k == mz - 2) dp[l.G(k, j, i)].z = 0;
            }
    return TOSCA_OK;
}
#endif  // ORACLE_TANGENT

#ifndef ORACLE_TANGENT
// ================================================================================================
// FormT (src/solvers/teqn_solvers.c:6-592), nvert = meshTag = 0
// ================================================================================================

// Compute_dscalar_i / _j / _k (src/include/inline.h:3210-3342): one normal difference and two
// four-point transverse averages, one-sided next to non-periodic boundary edges
static void dscalar(const Lay &l, const double *K, int dir, int64_t i, int64_t j, int64_t k, double *dc, double *de,
                    double *dz) {
    const bool iN = !l.i_per() && !l.ii_per(), jN = !l.j_per() && !l.jj_per(), kN = !l.k_per() && !l.kk_per();
    auto T = [&](int64_t kk, int64_t jj, int64_t ii) { return K[l.L(kk, jj, ii)]; };
    const int64_t mx = l.mx, my = l.my, mz = l.mz;
    if (dir == 0) {
        *dc = T(k, j, i + 1) - T(k, j, i);
        if (j == my - 2 && (i == 0 || i == mx - 2) && (jN || iN)) *de = (T(k, j, i + 1) + T(k, j, i) - T(k, j - 1, i + 1) - T(k, j - 1, i)) * 0.5;
        else if (j == 1 && (i == 0 || i == mx - 2) && (jN || iN)) *de = (T(k, j + 1, i + 1) + T(k, j + 1, i) - T(k, j, i + 1) - T(k, j, i)) * 0.5;
        else *de = (T(k, j + 1, i + 1) + T(k, j + 1, i) - T(k, j - 1, i + 1) - T(k, j - 1, i)) * 0.25;
        if (k == mz - 2 && (i == 0 || i == mx - 2) && (kN || iN)) *dz = (T(k, j, i + 1) + T(k, j, i) - T(k - 1, j, i + 1) - T(k - 1, j, i)) * 0.5;
        else if (k == 1 && (i == 0 || i == mx - 2) && (kN || iN)) *dz = (T(k + 1, j, i + 1) + T(k + 1, j, i) - T(k, j, i + 1) - T(k, j, i)) * 0.5;
        else *dz = (T(k + 1, j, i + 1) + T(k + 1, j, i) - T(k - 1, j, i + 1) - T(k - 1, j, i)) * 0.25;
    } else if (dir == 1) {
        if (i == mx - 2 && (j == 0 || j == my - 2) && (iN || jN)) *dc = (T(k, j + 1, i) + T(k, j, i) - T(k, j + 1, i - 1) - T(k, j, i - 1)) * 0.5;
        else if (i == 1 && (j == 0 || j == my - 2) && (iN || jN)) *dc = (T(k, j + 1, i + 1) + T(k, j, i + 1) - T(k, j + 1, i) - T(k, j, i)) * 0.5;
        else *dc = (T(k, j + 1, i + 1) + T(k, j, i + 1) - T(k, j + 1, i - 1) - T(k, j, i - 1)) * 0.25;
        *de = T(k, j + 1, i) - T(k, j, i);
        if (k == mz - 2 && (j == 0 || j == my - 2) && (kN || jN)) *dz = (T(k, j + 1, i) + T(k, j, i) - T(k - 1, j + 1, i) - T(k - 1, j, i)) * 0.5;
        else if (k == 1 && (j == 0 || j == my - 2) && (kN || jN)) *dz = (T(k + 1, j + 1, i) + T(k + 1, j, i) - T(k, j + 1, i) - T(k, j, i)) * 0.5;
        else *dz = (T(k + 1, j + 1, i) + T(k + 1, j, i) - T(k - 1, j + 1, i) - T(k - 1, j, i)) * 0.25;
    } else {
        if (i == mx - 2 && (k == 0 || k == mz - 2) && (iN || kN)) *dc = (T(k + 1, j, i) + T(k, j, i) - T(k + 1, j, i - 1) - T(k, j, i - 1)) * 0.5;
        else if (i == 1 && (k == 0 || k == mz - 2) && (iN || kN)) *dc = (T(k + 1, j, i + 1) + T(k, j, i + 1) - T(k + 1, j, i) - T(k, j, i)) * 0.5;
        else *dc = (T(k + 1, j, i + 1) + T(k, j, i + 1) - T(k + 1, j, i - 1) - T(k, j, i - 1)) * 0.25;
        if (j == my - 2 && (k == 0 || k == mz - 2) && (jN || kN)) *de = (T(k + 1, j, i) + T(k, j, i) - T(k + 1, j - 1, i) - T(k, j - 1, i)) * 0.5;
        else if (j == 1 && (k == 0 || k == mz - 2) && (jN || kN)) *de = (T(k + 1, j + 1, i) + T(k, j + 1, i) - T(k + 1, j, i) - T(k, j, i)) * 0.5;
        else *de = (T(k + 1, j + 1, i) + T(k, j + 1, i) - T(k + 1, j - 1, i) - T(k, j - 1, i)) * 0.25;
        *dz = T(k + 1, j, i) - T(k, j, i);
    }
}

// weno3 (src/include/inline.h:5147-5179); pow(x, 2.) is the exactly rounded square
static double weno3(double f0, double f1, double f2, double f3, double wavespeed) {
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

// models whose FormT diffusivity is les->lk_t (teqn_solvers.c:180-189)
static bool kt_model(int m) {
    return m == TOSCA_LES_AMD || m == TOSCA_LES_DSM || m == TOSCA_LES_DLASI || m == TOSCA_LES_DLASD ||
           m == TOSCA_LES_DPASD || m == TOSCA_LES_BAMD;
}

int oracle_formt(const tosca_dmda_info *info, const tosca_mesh *m, const tosca_params *p, const tosca_teqn_params *tp,
                 const double *ucont_, const double *tmprt, const double *nut, const double *kt, double *rhs,
                 double scale, int formMode) {
    Lay l;
    int rc = check_layout(info, p, l);
    if (rc) return rc;
    if (!tp || formMode < 0 || formMode > 2) { g_err = "formt: bad arguments"; return TOSCA_ERR_ARG; }
    const bool les = p->les_active != 0;
    if (les && (p->les_model == TOSCA_LES_STABILITY || p->les_model == TOSCA_LES_BV || p->les_model == TOSCA_LES_BAMD)) {
        g_err = "formt: stability-based and structural LES models are not covered";
        return TOSCA_ERR_UNSUPPORTED;
    }
    const int64_t mx = l.mx, my = l.my, mz = l.mz, n = l.nloc();
    const V3 *ucont = (const V3 *)ucont_;
    const V3 *icsi = (const V3 *)m->icsi, *ieta = (const V3 *)m->ieta, *izet = (const V3 *)m->izet;
    const V3 *jcsi = (const V3 *)m->jcsi, *jeta = (const V3 *)m->jeta, *jzet = (const V3 *)m->jzet;
    const V3 *kcsi = (const V3 *)m->kcsi, *keta = (const V3 *)m->keta, *kzet = (const V3 *)m->kzet;
    const double *iaj = m->iaj, *jaj = m->jaj, *kaj = m->kaj, *aj = m->aj;
    const double Pr = tp->pr, tRef = p->t_ref;
    const bool ktm = les && kt_model(p->les_model);
    std::vector<V3> div(n, V3{0, 0, 0}), visc(n, V3{0, 0, 0});
    // kappa_eff of one face (teqn_solvers.c:168-240 and the j / k copies); ajc the face Jacobian
    auto kappa = [&](int64_t c, int64_t cn, double ajc, double dtde) {
        double nu = p->nu;
        if (!les) return nu / Pr;
        nu = 0;
        const double nutf = 0.5 * (nut[c] + nut[cn]);
        if (ktm) {
            const double diff = 0.5 * (kt[c] + kt[cn]);
            return (nu / Pr) + diff;
        }
        const double gradTdotG = dtde * (-9.81);
        double l_, delta = pow(1. / ajc, 1. / 3.);
        if (gradTdotG < 0.) l_ = pmin(delta, 7.6 * nutf / delta * sqrt(tRef / fabs(gradTdotG)));
        else l_ = delta;
        const double Prt = 1.0 / (1.0 + (2.0 * l_ / delta));
        return (nu / Pr) + (nutf / Prt);
    };
    for (int64_t k = 0; k < mz - 1; k++)
        for (int64_t j = 0; j < my - 1; j++)
            for (int64_t i = 0; i < mx - 1; i++) {
                const int64_t c = l.L(k, j, i);
                double dtdc, dtde, dtdz;
                if (j != 0 && k != 0) {                      // i face (:122-246)
                    const double ajc = iaj[c];
                    const V3 cs = icsi[c], et = ieta[c], ze = izet[c];
                    const double g11 = cs.x * cs.x + cs.y * cs.y + cs.z * cs.z;
                    const double g21 = et.x * cs.x + et.y * cs.y + et.z * cs.z;
                    const double g31 = ze.x * cs.x + ze.y * cs.y + ze.z * cs.z;
                    dscalar(l, tmprt, 0, i, j, k, &dtdc, &dtde, &dtdz);
                    int64_t iL, iR;
                    face2cell4(l.per[0], i, mx, &iL, &iR);
                    div[c].x = -ucont[c].x * weno3(tmprt[l.L(k, j, iL)], tmprt[c], tmprt[l.L(k, j, i + 1)],
                                                   tmprt[l.L(k, j, iR)], ucont[c].x);
                    double kap = kappa(c, l.L(k, j, i + 1), ajc, dtde);
                    if (les && ((tp->wall_t[0] && i == 0) || (tp->wall_t[1] && i == mx - 2))) {
                        const double signQ = (i == 0) ? -1.0 : 1.0;
                        const double *q = tp->q_wall_i;
                        const int64_t pl = mz * my, o = k * my + j;
                        visc[c].x = signQ * (q[o] * cs.x + q[pl + o] * cs.y + q[2 * pl + o] * cs.z);
                        kap = 0.0;
                    }
                    visc[c].x += (g11 * dtdc + g21 * dtde + g31 * dtdz) * ajc * kap;
                }
                if (i != 0 && k != 0) {                      // j face (:248-370)
                    const double ajc = jaj[c];
                    const V3 cs = jcsi[c], et = jeta[c], ze = jzet[c];
                    const double g11 = cs.x * et.x + cs.y * et.y + cs.z * et.z;
                    const double g21 = et.x * et.x + et.y * et.y + et.z * et.z;
                    const double g31 = ze.x * et.x + ze.y * et.y + ze.z * et.z;
                    dscalar(l, tmprt, 1, i, j, k, &dtdc, &dtde, &dtdz);
                    int64_t jL, jR;
                    face2cell4(l.per[1], j, my, &jL, &jR);
                    div[c].y = -ucont[c].y * weno3(tmprt[l.L(k, jL, i)], tmprt[c], tmprt[l.L(k, j + 1, i)],
                                                   tmprt[l.L(k, jR, i)], ucont[c].y);
                    double kap = kappa(c, l.L(k, j + 1, i), ajc, dtde);
                    if (les && ((tp->wall_t[2] && j == 0) || (tp->wall_t[3] && j == my - 2))) {
                        const double signQ = (j == 0) ? -1.0 : 1.0;
                        const double *q = tp->q_wall_j;
                        const int64_t pl = mz * mx, o = k * mx + i;
                        visc[c].y = signQ * (q[o] * et.x + q[pl + o] * et.y + q[2 * pl + o] * et.z);
                        kap = 0.0;
                    }
                    visc[c].y += (g11 * dtdc + g21 * dtde + g31 * dtdz) * ajc * kap;
                }
                if (i != 0 && j != 0) {                      // k face (:372-478)
                    const double ajc = kaj[c];
                    const V3 cs = kcsi[c], et = keta[c], ze = kzet[c];
                    const double g11 = cs.x * ze.x + cs.y * ze.y + cs.z * ze.z;
                    const double g21 = et.x * ze.x + et.y * ze.y + et.z * ze.z;
                    const double g31 = ze.x * ze.x + ze.y * ze.y + ze.z * ze.z;
                    dscalar(l, tmprt, 2, i, j, k, &dtdc, &dtde, &dtdz);
                    int64_t kL, kR;
                    face2cell4(l.per[2], k, mz, &kL, &kR);
                    div[c].z = -ucont[c].z * weno3(tmprt[l.L(kL, j, i)], tmprt[c], tmprt[l.L(k + 1, j, i)],
                                                   tmprt[l.L(kR, j, i)], ucont[c].z);
                    const double kap = kappa(c, l.L(k + 1, j, i), ajc, dtde);
                    visc[c].z += (g11 * dtdc + g21 * dtde + g31 * dtdz) * ajc * kap;
                }
            }
    local_to_local(l, div.data());                            // :487-490
    local_to_local(l, visc.data());
    reset_face_periodic(l, div.data());                       // :502-503
    reset_face_periodic(l, visc.data());
    for (int64_t k = 1; k < mz - 1; k++)                      // :511-537
        for (int64_t j = 1; j < my - 1; j++)
            for (int64_t i = 1; i < mx - 1; i++) {
                const int64_t c = l.L(k, j, i);
                const double divBudget = div[c].x - div[l.L(k, j, i - 1)].x + div[c].y - div[l.L(k, j - 1, i)].y +
                                         div[c].z - div[l.L(k - 1, j, i)].z;
                const double viscBudget = visc[c].x - visc[l.L(k, j, i - 1)].x + visc[c].y - visc[l.L(k, j - 1, i)].y +
                                          visc[c].z - visc[l.L(k - 1, j, i)].z;
                rhs[l.G(k, j, i)] = scale * ((formMode != 2 ? divBudget : 0.0) + (formMode != 1 ? viscBudget : 0.0)) * aj[c];
            }
    return TOSCA_OK;
}
#endif  // ORACLE_TANGENT

#ifndef ORACLE_TANGENT
// ================================================================================================
// Boundary-condition ghost fills (src/boundary.c).  Each routine reads the local array as it was on
// entry and writes a global working copy (the owned range), which the GlobalToLocal scatter then
// copies back and wraps; on one rank the copy is the owned part of the local array.
// ================================================================================================
static void global_to_local(const Lay &l, const std::vector<V3> &g, V3 *loc) {
    for (int64_t k = 0; k < l.mz; k++)
        for (int64_t j = 0; j < l.my; j++)
            for (int64_t i = 0; i < l.mx; i++) loc[l.L(k, j, i)] = g[l.L(k, j, i)];
    local_to_local(l, loc);
}

// PatchType of checkBCsAndSetPatchTypes (src/boundary.c:318-348): 1 for noSlip, slip,
// velocityWallFunction, 0 otherwise
static int patch_type(int bc) {
    return (bc == TOSCA_BCU_NOSLIP || bc == TOSCA_BCU_SLIP || bc == TOSCA_BCU_WALLFUNCTION) ? 1 : 0;
}

int oracle_contravariant_bcs(const tosca_dmda_info *info, const tosca_mesh *m, const tosca_params *p, double *ucont_,
                             const double *ucat_) {
    Lay l;
    int rc = check_layout(info, p, l);
    if (rc) return rc;
    V3 *lucont = (V3 *)ucont_;
    const V3 *lucat = (const V3 *)ucat_;
    const V3 *icsi = (const V3 *)m->icsi, *jeta = (const V3 *)m->jeta, *kzet = (const V3 *)m->kzet;
    const int *bc = p->bc_u;
    const int64_t mx = l.mx, my = l.my, mz = l.mz;
    std::vector<V3> ucont(lucont, lucont + l.nloc());
    for (int64_t k = 0; k < mz - 1; k++)                     // src/boundary.c:646-807
        for (int64_t j = 0; j < my - 1; j++)
            for (int64_t i = 0; i < mx - 1; i++) {
                const int64_t c = l.L(k, j, i);
                if ((patch_type(bc[4]) == 0 && k == 0) || (patch_type(bc[5]) == 0 && k == mz - 2)) {
                    const V3 a = lucat[l.L(k + 1, j, i)], b = lucat[c];
                    ucont[c].z = 0.5 * (a.x + b.x) * kzet[c].x + 0.5 * (a.y + b.y) * kzet[c].y + 0.5 * (a.z + b.z) * kzet[c].z;
                }
                if ((patch_type(bc[2]) == 0 && j == 0) || (patch_type(bc[3]) == 0 && j == my - 2)) {
                    const V3 a = lucat[l.L(k, j + 1, i)], b = lucat[c];
                    ucont[c].y = 0.5 * (a.x + b.x) * jeta[c].x + 0.5 * (a.y + b.y) * jeta[c].y + 0.5 * (a.z + b.z) * jeta[c].z;
                }
                if ((patch_type(bc[0]) == 0 && i == 0) || (patch_type(bc[1]) == 0 && i == mx - 2)) {
                    const V3 a = lucat[l.L(k, j, i + 1)], b = lucat[c];
                    ucont[c].x = 0.5 * (a.x + b.x) * icsi[c].x + 0.5 * (a.y + b.y) * icsi[c].y + 0.5 * (a.z + b.z) * icsi[c].z;
                }
                if ((patch_type(bc[0]) == 1 && i == 0) || (patch_type(bc[1]) == 1 && i == mx - 2)) ucont[c].x = 0.0;
                if ((patch_type(bc[2]) == 1 && j == 0) || (patch_type(bc[3]) == 1 && j == my - 2)) ucont[c].y = 0.0;
                if ((patch_type(bc[4]) == 1 && k == 0) || (patch_type(bc[5]) == 1 && k == mz - 2)) ucont[c].z = 0.0;
                if (bc[0] == TOSCA_BCU_ZEROGRADIENT && i == 0 && ucont[c].x > 0.0) ucont[c].x = 0.0;
                if (bc[1] == TOSCA_BCU_ZEROGRADIENT && i == mx - 2 && ucont[c].x < 0.0) ucont[c].x = 0.0;
                if (bc[2] == TOSCA_BCU_ZEROGRADIENT && j == 0 && ucont[c].y > 0.0) ucont[c].y = 0.0;
                if (bc[3] == TOSCA_BCU_ZEROGRADIENT && j == my - 2 && ucont[c].y < 0.0) ucont[c].y = 0.0;
                if (bc[4] == TOSCA_BCU_ZEROGRADIENT && k == 0 && ucont[c].z > 0.0) ucont[c].z = 0.0;
                if (bc[5] == TOSCA_BCU_ZEROGRADIENT && k == mz - 2 && ucont[c].z < 0.0) ucont[c].z = 0.0;
                if (bc[0] == TOSCA_BCU_PERIODIC && i == 0) {
                    if (l.i_per()) ucont[c].x = lucont[l.L(k, j, mx - 2)].x;
                    else if (l.ii_per()) ucont[c].x = lucont[l.L(k, j, -2)].x;
                }
                if (bc[2] == TOSCA_BCU_PERIODIC && j == 0) {
                    if (l.j_per()) ucont[c].y = lucont[l.L(k, my - 2, i)].y;
                    else if (l.jj_per()) ucont[c].y = lucont[l.L(k, -2, i)].y;
                }
                if (bc[4] == TOSCA_BCU_PERIODIC && k == 0) {
                    if (l.k_per()) ucont[c].z = lucont[l.L(mz - 2, j, i)].z;
                    else if (l.kk_per()) ucont[c].z = lucont[l.L(-2, j, i)].z;
                }
            }
    global_to_local(l, ucont, lucont);                       // :816-817
    return TOSCA_OK;
}

// calculateNormal (src/include/inline.h:1231-1262): unit covariant base vector of each direction
static void unit_normals(V3 c, V3 e, V3 z, double ni[3], double nj[3], double nk[3]) {
    const double g[3][3] = {{c.x, c.y, c.z}, {e.x, e.y, e.z}, {z.x, z.y, z.z}};
    double G[3][3];
    covariant(g, G);
    double v[3][3] = {{G[0][0], G[1][0], G[2][0]}, {G[0][1], G[1][1], G[2][1]}, {G[0][2], G[1][2], G[2][2]}};
    double *o[3] = {ni, nj, nk};
    for (int a = 0; a < 3; a++) {
        const double sum = sqrt(v[a][0] * v[a][0] + v[a][1] * v[a][1] + v[a][2] * v[a][2]);
        o[a][0] = v[a][0] / sum; o[a][1] = v[a][1] / sum; o[a][2] = v[a][2] / sum;
    }
}

// ---- Cabot wall law (src/wallfunctions.c) ----------------------------------------------------
// near_wall_eddy_viscosity (wallfunctions.c:280-284)
static double nw_nut(double yp) { return 0.41 * yp * pow(1. - exp(-yp / 19.), 2.0); }

// integration_buffer of pre_integrate (wallfunctions.c:296-360): entry i is the integral of
// 1 / (1 + nu_t+) over y+ in [0, 2 i], summed interval by interval in index order.  The table is
// grown on demand: an entry depends only on the ones before it, so a partial table holds the same
// values as the reference's full one (n_yp = 5e6).
static const int64_t CAB_NYP = 5000000;
static std::vector<double> g_cabot{0.0};
static double cabot_entry(int64_t i) {
    if ((int64_t)g_cabot.size() <= i) {
        for (int64_t e = (int64_t)g_cabot.size(); e <= i; e++) {
            const int N = 24;
            const double ya_p = (double)(e - 1) * 2, yb_p = (double)(e) * 2, dy_p = (yb_p - ya_p) / (double)N;
            std::vector<double> E(N + 1);
            double val = 0, ybegin_p = ya_p;
            for (int k = 0; k <= N; k++) E[k] = 1. / (1. + nw_nut(ybegin_p + dy_p * k));
            for (int k = 0; k < N; k++) {
                const double F0 = E[k];
                const double F1 = 1. / (1. + nw_nut(ybegin_p + dy_p * 1. / 3.));
                const double F2 = 1. / (1. + nw_nut(ybegin_p + dy_p * 2. / 3.));
                const double F3 = E[k + 1];
                val += dy_p / 3. * (3 * F0 + 9 * F1 + 9 * F2 + 3 * F3) / 8.;
                ybegin_p += dy_p;
            }
            g_cabot.push_back(g_cabot.back() + val);
        }
    }
    return g_cabot[i];
}

// integrate_F (wallfunctions.c:364-420) for a smooth wall.  Past y+ = 1e7 the reference starts
// from integration_buffer[n_yp] with its local n_yp = 0 (pre_integrate sets it on the process's
// first call only), i.e. from 0.  Indices outside the table (y+ < 0, or y+ at its last point) are
// clamped to it; the reference reads outside its buffer there.
static double cabot_integral(double nu, double utau, double yb) {
    double val = 0;
    const double ya_plus = 0 * utau / nu, yb_plus = yb * utau / nu;
    (void)ya_plus;
    if (yb_plus <= 1e7) {
        const double qb = yb_plus / (double)2;
        int64_t ib = qb > 0.0 ? (int64_t)qb : 0;
        ib = std::min<int64_t>(ib, CAB_NYP - 1);
        const double lo = cabot_entry(ib), hi = cabot_entry(ib + 1);
        const double int_a = 0;
        const double int_b = (hi - lo) / (double)(2) * (yb_plus - (double)(ib) * 2) + lo;
        return (int_b - int_a) / utau;
    }
    const int N = 4;
    const double ya = 1e7, ydiff = yb_plus - ya, dy = ydiff / (double)N;
    std::vector<double> E(N + 1);
    double ybegin = ya;
    for (int i = 0; i <= N; i++) E[i] = 1. / (1. + nw_nut(ybegin + dy * i));
    for (int i = 0; i < N; i++) {
        const double F0 = E[i], F1 = 1. / (1. + nw_nut(ybegin + dy * 1. / 3.)), F2 = 1. / (1. + nw_nut(ybegin + dy * 2. / 3.)),
                     F3 = E[i + 1];
        val += dy / 3. * (3 * F0 + 9 * F1 + 9 * F2 + 3 * F3) / 8.;
        ybegin += dy;
    }
    return val / utau;
}

static double f_cabot(double nu, double u, double y, double utau) { return utau * utau * cabot_integral(nu, utau, y) - u; }

// uTauCabot (wallfunctions.c:519-543)
static double utau_cabot(double nu, double u, double y) {
    double x = 0, x0 = 0.01;
    const double eps = 1.e-7;
    for (int i = 0; i < 30; i++) {
        const double df = (f_cabot(nu, u, y, x0 + eps) - f_cabot(nu, u, y, x0 - eps)) / (2.0 * eps);
        x = x0 - f_cabot(nu, u, y, x0) / df;
        if (fabs(x0 - x) < 1.e-10) break;
        x0 = x;
    }
    return x;
}

// utau_wf (wallfunctions.c:568-585)
static double utau_rough(double nu, double ks, double sb, double ut) {
    const double k_s = ks * 0.033, kappa = 0.41;
    if (k_s > 1.e-13) return ut * kappa / log(sb / k_s);
    const double A = 8.3, B = 1.0 / 7.0;
    return pow(ut * pow(nu, B) / (A * pow(sb, B)), 1.0 / (1.0 + B));
}

extern "C" double oracle_cabot_utau(double nu, double u, double y) { return utau_cabot(nu, u, y); }
extern "C" double oracle_cabot_table(int64_t i) { return cabot_entry(i); }

// ---- Nieuwstadt inflow (src/inflow.c:2732-2811) and its special functions --------------------------
// The complex gamma, digamma and Gauss hypergeometric functions of src/include/inline.h:372-525,
// restated step by step, with the reference's coefficient tables (published Lanczos coefficients:
// g = 9 with 11 terms for gamma, g = 607/128 with 15 terms for digamma), so that the profile
// matches the reference's to round-off.
namespace {
using cplx = std::complex<double>;
const double kLanczosG9[11] = {1.000000000000000174663,     5716.400188274341379136,    -14815.30426768413909044,
                               14291.49277657478554025,     -6348.160217641458813289,   1301.608286058321874105,
                               -108.1767053514369634679,    2.605696505611755827729,    -0.7423452510201416151527e-2,
                               0.5384136432509564062961e-7, -0.4023533141268236372067e-8};
const double kLanczosG607[15] = {0.99999999999999709182,     57.156235665862923517,      -59.597960355475491248,
                                 14.136097974741747174,      -0.49191381609762019978,    .33994649984811888699e-4,
                                 .46523628927048575665e-4,   -.98374475304879564677e-4,  .15808870322491248884e-3,
                                 -.21026444172410488319e-3,  .21743961811521264320e-3,   -.16431810653676389022e-3,
                                 .84418223983852743293e-4,   -.26190838401581408670e-4,  .36899182659531622704e-5};

bool nonpositive_integer(cplx x) { return std::round(x.real()) == x.real() && x.imag() == 0.0 && x.real() <= 0.0; }

// gamma (inline.h:372-427): Lanczos for Re x >= 0 (of -x otherwise, then the reflection formula);
// exact products for positive integers, 1e20 at the poles
cplx cgamma(cplx x) {
    cplx z = x;
    if (x.real() < 0.0) z = -1.0 * z;
    const cplx g(9.0, 0.0);
    cplx t = z + g, s(0.0, 0.0);
    for (int k = 10; k > 0; k--) {
        s = s + (cplx)kLanczosG9[k] / t;
        t = t - 1.0;
    }
    s = s + (cplx)kLanczosG9[0];
    const cplx ss = z + g - 0.5;
    s = std::log(s * std::sqrt(2.0 * M_PI)) + (z - 0.5) * std::log(ss) - ss;
    cplx f = std::exp(s);
    if (x.real() < 0.0) f = -1.0 * M_PI / (x * f * std::sin(M_PI * x));
    if (nonpositive_integer(x)) f = 1e20;
    if (std::round(x.real()) == x.real() && x.imag() == 0.0 && x.real() > 0.0) {
        cplx pp(1.0, 0.0);
        for (int64_t m = 1; m < (int64_t)x.real(); m++) pp = pp * cplx((double)m, 0.0);
        f = pp;
    }
    return f;
}

// digamma (inline.h:429-476): the logarithmic derivative of the Lanczos sum, the reflection
// formula for Re x < 1/2, 1e20 at the poles
cplx cdigamma(cplx x) {
    cplx z = x;
    if (x.real() < 0.5) z = 1.0 - z;
    const cplx g(607.0 / 128.0, 0.0);
    cplx n(0.0, 0.0), d = n;
    for (int k = 14; k > 0; k--) {
        const cplx dz = 1.0 / (z + (double)(k + 1) - 2.0);
        const cplx dd = kLanczosG607[k] * dz;
        d = d + dd;
        n = n - dd * dz;
    }
    d = d + kLanczosG607[0];
    const cplx gg = z + g - 0.5;
    cplx f = std::log(gg) + (n / d - g / gg);
    if (x.real() < 0.5) f = f - M_PI * (1.0 / std::tan(M_PI * x));
    if (nonpositive_integer(x)) f = 1e20;
    return f;
}

// 2F1(a, b; c; z) for real z (inline.h:478-525): the power series, left after the third
// consecutive term below 1e-15 of the sum (30000 terms at most); NaN when the sum is NaN
cplx hypergeom(cplx a, cplx b, cplx c, double z) {
    if (nonpositive_integer(c)) return cplx(1e20, 0.0);
    const double epsilon = 1e-15;
    cplx term(1.0, 0.0), result(1.0, 0.0);
    int termination = 0;
    for (int64_t k = 0; k < 30000; k++) {
        term = term * ((a + (double)k) * (b + (double)k) * z / ((c + (double)k) * ((double)k + 1.0)));
        if (std::abs(term) < std::abs(result * epsilon)) {
            if (++termination >= 3) return result;
        } else {
            termination = 0;
        }
        if (term == 0.0) break;
        result = result + term;
        if (std::isnan(result.real()) || std::isnan(result.imag())) return cplx(std::nan(""), std::nan(""));
    }
    return result;
}

// W(eta) = Wg - dW(eta), the complex velocity of the capped Ekman layer over u_tau
cplx nieuwstadt_w(const tosca_inlet *in, double C, double eta) {
    const double vkConstant = 0.4;
    const cplx alpha = 0.5 + 0.5 * std::sqrt(cplx(1.0, 4.0 * C));
    const cplx Const = cdigamma(alpha + 1.0) + cdigamma(alpha - 1.0) - 2.0 * cdigamma(1.0);
    const cplx Wg = 1.0 / vkConstant * (std::log(in->h_inv / in->roughness) - Const);
    const cplx deltaW = cplx(0.0, -1.0 / vkConstant) * alpha * alpha * cgamma(alpha) * cgamma(alpha) /
                        (C * cgamma(2.0 * alpha)) * std::pow(1.0 - eta, alpha - 1.0) *
                        hypergeom(alpha + 1.0, alpha - 1.0, 2.0 * alpha, 1.0 - eta);
    return Wg - deltaW;
}

// NieuwstadtInflowEvaluate (inflow.c:2732-2811): the velocity at height h, eta = min(h / hInv, 1),
// rotated by the angle between udir and the profile's direction at hReference (whose normalisation
// divides V_r by the norm formed with the already normalised U_r, as the reference does)
V3 nieuwstadt(const tosca_inlet *in, double h) {
    const double vkConstant = 0.4;
    const double C = in->fc * in->h_inv / (vkConstant * in->u_tau);
    const cplx Wp = nieuwstadt_w(in, C, h < in->h_inv ? h / in->h_inv : 1.0);
    const double U_p = Wp.real() * in->u_tau, V_p = Wp.imag() * in->u_tau;
    const cplx Wr = nieuwstadt_w(in, C, in->h_ref / in->h_inv);
    double U_r = Wr.real() * in->u_tau, V_r = Wr.imag() * in->u_tau;
    U_r = U_r / sqrt(U_r * U_r + V_r * V_r);
    V_r = V_r / sqrt(U_r * U_r + V_r * V_r);
    const double theta = std::acos(in->udir[0] * U_r + in->udir[1] * V_r);
    return V3{std::cos(theta) * U_p - std::sin(theta) * V_p, std::sin(theta) * U_p + std::cos(theta) * V_p, 0.0};
}
}  // namespace

extern "C" void oracle_cgamma(double re, double im, double *out) {
    const cplx f = cgamma(cplx(re, im));
    out[0] = f.real(); out[1] = f.imag();
}
extern "C" void oracle_cdigamma(double re, double im, double *out) {
    const cplx f = cdigamma(cplx(re, im));
    out[0] = f.real(); out[1] = f.imag();
}
extern "C" void oracle_hypergeom(const double *a, const double *b, const double *c, double z, double *out) {
    const cplx f = hypergeom(cplx(a[0], a[1]), cplx(b[0], b[1]), cplx(c[0], c[1]), z);
    out[0] = f.real(); out[1] = f.imag();
}
extern "C" void oracle_nieuwstadt(const tosca_inlet *in, double h, double *out) {
    const V3 u = nieuwstadt(in, h);
    out[0] = u.x; out[1] = u.y; out[2] = u.z;
}

// ---- mapped inflow (typeU 3 and 4) helpers ------------------------------------------------------
// findInterpolationWeights (src/include/inline.h:5873-5938): the two closest of npts points, a
// point passed over when its distance equals the current runner-up's (or leader's) to 1e-5
// relative; lower point first; linear weights
void interp_weights(const double *pvec, int npts, double pval, double *weights, int *labels) {
    const double eqTol = 1e-5;
    std::vector<double> diff(npts);
    int idx_1 = 0, idx_2 = 1;
    for (int i = 0; i < npts; i++) diff[i] = std::fabs(pvec[i] - pval);
    for (int i = 0; i < npts; i++) {
        const double tol1 = std::fabs(diff[i] - diff[idx_2]) / diff[i];
        if (diff[i] < diff[idx_1] && tol1 > eqTol) { idx_2 = idx_1; idx_1 = i; }
        const double tol2 = std::fabs(diff[i] - diff[idx_1]) / diff[i];
        if (diff[i] < diff[idx_2] && i != idx_1 && tol2 > eqTol) idx_2 = i;
    }
    if (pvec[idx_2] < pvec[idx_1]) std::swap(idx_1, idx_2);
    labels[0] = idx_1; labels[1] = idx_2;
    const double idx = (idx_2 - idx_1) / (pvec[idx_2] - pvec[idx_1]) * (pval - pvec[idx_1]) + idx_1;
    weights[0] = (idx_2 - idx) / (idx_2 - idx_1);
    weights[1] = (idx - idx_1) / (idx_2 - idx_1);
}
// findInterpolationWeightsWithExtrap (inline.h:5943-5966): the end point, weights 1/2 each, outside
void interp_weights_extrap(const double *pvec, int npts, double pval, double *weights, int *labels) {
    if (pval >= pvec[npts - 1]) { weights[0] = weights[1] = 0.5; labels[0] = labels[1] = npts - 1; }
    else if (pval <= pvec[0]) { weights[0] = weights[1] = 0.5; labels[0] = labels[1] = 0; }
    else interp_weights(pvec, npts, pval, weights, labels);
}
extern "C" void oracle_interp_weights_extrap(const double *pvec, int npts, double pval, double *weights, int *labels) {
    interp_weights_extrap(pvec, npts, pval, weights, labels);
}
// scaleHyperTangBot / scaleHyperTangTop (inline.h:5971-5999)
double hyper_tang_bot(double h, double H, double delta) { return 0.5 * (1.0 - std::tanh(7.0 * (h - (H - 0.5 * delta)) / delta)); }
double hyper_tang_top(double h, double H, double delta) { return 0.5 * (1.0 + std::tanh(7.0 * (h - (H - 0.5 * delta)) / delta)); }

// the n1Merge blend at height h (boundary.c:1318-1343, 1580-1617)
V3 merge_top(const tosca_inlet *in, double height, V3 u) {
    double Wg[2];
    int IDs[2];
    interp_weights_extrap(in->avg_top_coords, 10, height, Wg, IDs);
    const double bot = hyper_tang_bot(height, in->avg_top_length, in->avg_top_delta);
    const double top = hyper_tang_top(height, in->avg_top_length, in->avg_top_delta);
    const double *a = in->avg_top_u[IDs[0]], *b = in->avg_top_u[IDs[1]];
    return V3{bot * u.x + top * (a[0] * Wg[0] + b[0] * Wg[1]), bot * u.y + top * (a[1] * Wg[0] + b[1] * Wg[1]),
              bot * u.z + top * (a[2] * Wg[0] + b[2] * Wg[1])};
}

// one component of a typeU-4 stencil at table entry (j, i) of a [my][mx][np] table
double stencil(const tosca_inlet *in, const int *cells, const double *w, int np, int64_t jm, int comp) {
    double s = 0.0;
    for (int q = 0; q < np; q++) {
        const int64_t e = jm * np + q;
        const int64_t pj = cells[2 * e], pi = cells[2 * e + 1];
        s += w[e] * in->plane[3 * (pj * (in->n2 + 2) + pi) + comp];
    }
    return s;
}
V3 interp_point(const tosca_inlet *in, int64_t mx, int64_t j, int64_t i) {
    const int64_t jm = j * mx + i;
    V3 u;
    u.x = stencil(in, in->cells, in->weights, 4, jm, 0);
    u.y = in->spline ? stencil(in, in->cells_2, in->weights_2, 6, jm, 1) : stencil(in, in->cells, in->weights, 4, jm, 1);
    u.z = in->spline ? stencil(in, in->cells_1, in->weights_1, 6, jm, 2) : stencil(in, in->cells, in->weights, 4, jm, 2);
    return u;
}

// the kLeft ghost of a mapped inflow (typeU 3: boundary.c:1302-1371; typeU 4: :1372-1622)
V3 mapped_inflow(const tosca_inlet *in, int64_t mx, int64_t j, int64_t i, double height) {
    auto plane = [&](int64_t pj, int64_t pi) {
        const double *v = in->plane + 3 * (pj * (in->n2 + 2) + pi);
        return V3{v[0], v[1], v[2]};
    };
    if (in->type == 3) {
        int64_t jif = j % in->n1 == 0 ? in->n1 : j % in->n1;
        int64_t iif = i % in->n2 == 0 ? in->n2 : i % in->n2;
        if (j <= in->n1 * in->prds1 && i <= in->n2 * in->prds2)
            return in->merge1 ? merge_top(in, height, plane(jif, iif)) : plane(jif, iif);
        if (in->merge1) return V3{in->avg_top_u[9][0], in->avg_top_u[9][1], in->avg_top_u[9][2]};
        if (j > in->n1 * in->prds1) jif = in->n1;
        if (i > in->n2 * in->prds2) iif = in->n2;
        return plane(jif, iif);
    }
    V3 g;
    if (in->shift2) {
        const int64_t iL = in->y_ids[i] - 1, iR = in->y_ids[i];
        const V3 uL = interp_point(in, mx, j, iL), uR = interp_point(in, mx, j, iR);
        const double w = in->y_weights[i];
        g = V3{(1.0 - w) * uL.x + w * uR.x, (1.0 - w) * uL.y + w * uR.y, (1.0 - w) * uL.z + w * uR.z};
    } else {
        g = interp_point(in, mx, j, i);
    }
    return in->merge1 ? merge_top(in, height, g) : g;
}

int oracle_cartesian_bcs(const tosca_dmda_info *info, const tosca_mesh *m, const tosca_params *p, double *ucat_,
                         const double *ufix, double *ustar, const tosca_inlet *inlet) {
    Lay l;
    int rc = check_layout(info, p, l);
    if (rc) return rc;
    const int *bc = p->bc_u;
    for (int q = 0; q < 6; q++) {
        if ((bc[q] == TOSCA_BCU_GENERIC && !(q == 4 && inlet)) || (bc[q] == TOSCA_BCU_WALLFUNCTION && q >= 4)) {
            g_err = "cartesian BCs: inlet functions other than on kLeft and k wall functions are not covered";
            return TOSCA_ERR_UNSUPPORTED;
        }
        if (bc[q] == TOSCA_BCU_WALLFUNCTION && p->wall_u[q] != -1 && p->wall_u[q] != -3) {
            g_err = "cartesian BCs: wall function without a wall model";
            return TOSCA_ERR_ARG;
        }
        if (bc[q] == TOSCA_BCU_WALLFUNCTION && p->wall_u[q] == -1 && !ustar) { g_err = "Cabot patch without ustar"; return TOSCA_ERR_ARG; }
        if (bc[q] == TOSCA_BCU_FIXEDVALUE && !ufix) { g_err = "fixedValue patch without values"; return TOSCA_ERR_ARG; }
    }
    const int *wf = p->wall_u;
    const double *aj = m->aj;
    const bool inletK = inlet && bc[4] == TOSCA_BCU_GENERIC;
    if (inletK && (inlet->type < 1 || inlet->type > 6)) {
        g_err = "inlet function type not covered";
        return TOSCA_ERR_UNSUPPORTED;
    }
    if (inletK) {
        const int t = inlet->type;
        const bool bad = !inlet->cent || (t == 1 && !inlet->noise) ||
                         ((t == 3 || t == 4) && (!inlet->plane || inlet->n1 < 1 || inlet->n2 < 1)) ||
                         (t == 3 && (inlet->prds1 < 1 || inlet->prds2 < 1)) ||
                         (t == 4 && (!inlet->cells || !inlet->weights)) ||
                         (t == 4 && inlet->spline && (!inlet->cells_1 || !inlet->cells_2 || !inlet->weights_1 || !inlet->weights_2)) ||
                         (t == 4 && inlet->shift2 && (!inlet->y_ids || !inlet->y_weights));
        if (bad) { g_err = "inlet function: missing table or bad plane size"; return TOSCA_ERR_ARG; }
    }
    V3 *lucat = (V3 *)ucat_;
    const V3 *csi = (const V3 *)m->csi, *eta = (const V3 *)m->eta, *zet = (const V3 *)m->zet;
    const int64_t mx = l.mx, my = l.my, mz = l.mz;
    std::vector<V3> ucat(lucat, lucat + l.nloc());
    auto slip = [&](int64_t c, int dir, int64_t gdst) {     // boundary.c:1684-1830
        double ni[3], nj[3], nk[3];
        unit_normals(csi[c], eta[c], zet[c], ni, nj, nk);
        const double *nn = dir == 0 ? ni : dir == 1 ? nj : nk;
        const V3 n{nn[0], nn[1], nn[2]}, U = lucat[c];
        const double d = U.x * n.x + U.y * n.y + U.z * n.z;
        const V3 N2{2.0 * d * n.x, 2.0 * d * n.y, 2.0 * d * n.z};
        ucat[gdst] = V3{U.x - N2.x, U.y - N2.y, U.z - N2.z};
    };
    auto fix = [&](int q, int64_t gdst) { ucat[gdst] = V3{ufix[3 * q], ufix[3 * q + 1], ufix[3 * q + 2]}; };
    auto zg = [&](int64_t c, int64_t c2, int64_t gdst) {
        const V3 a = lucat[c], b = lucat[c2];
        ucat[gdst] = V3{2 * a.x - b.x, 2 * a.y - b.y, 2 * a.z - b.z};
    };
    auto neg = [&](int64_t c, int64_t gdst) { const V3 a = lucat[c]; ucat[gdst] = V3{-1.0 * a.x, -1.0 * a.y, -1.0 * a.z}; };
    // Shumann wall function on a j patch (boundary.c:1153-1224): the wall-parallel part of the
    // boundary cell b is extrapolated linearly through the ghost from the next cell c, and the
    // wall-normal part of b is reflected; the slopes dudn0/dudn1 the reference also forms are
    // not used
    // Cabot wall function on an i (dir 0) or j (dir 1) patch (boundary.c:968-1150)
    auto cabot = [&](int64_t c, int dir, int64_t gdst, double roughness) {
        const V3 A = dir == 0 ? csi[c] : eta[c];
        const double area = sqrt(A.x * A.x + A.y * A.y + A.z * A.z);
        const double sb = 0.5 / aj[c] / area;
        const V3 Ub = lucat[c];
        double ni[3], nj[3], nk[3];
        unit_normals(csi[c], eta[c], zet[c], ni, nj, nk);
        const double *nn = dir == 0 ? ni : nj;
        const V3 n{nn[0], nn[1], nn[2]};
        const double d = Ub.x * n.x + Ub.y * n.y + Ub.z * n.z;
        const V3 UbN{d * n.x, d * n.y, d * n.z};
        const V3 UbP{Ub.x - UbN.x, Ub.y - UbN.y, Ub.z - UbN.z};
        const double mag = sqrt(UbP.x * UbP.x + UbP.y * UbP.y + UbP.z * UbP.z);
        const double nu = p->nu;
        if (mag < 1.e-20) {           // u* is solved for and then reset (its value is not used)
            ustar[c] = 0.0;
            ucat[gdst] = V3{0.0, 0.0, 0.0};
            return;
        }
        ustar[c] = roughness > 1.e-19 ? utau_rough(nu, roughness, sb, mag) : utau_cabot(nu, mag, sb);
        const double tau_w = ustar[c] * ustar[c];
        const double a1 = mag - 2.0 * sb * tau_w / nu, a2 = -1.0 * mag;
        const double gm = (a1 < a2) ? a2 : a1;
        const double vm = sqrt(UbP.x * UbP.x + UbP.y * UbP.y + UbP.z * UbP.z);
        const V3 dir_{UbP.x / (vm + 1.e-12), UbP.y / (vm + 1.e-12), UbP.z / (vm + 1.e-12)};
        const V3 G{gm * dir_.x, gm * dir_.y, gm * dir_.z};
        ucat[gdst] = V3{G.x - UbN.x, G.y - UbN.y, G.z - UbN.z};
    };
    const double roughI = wf[0] == -1 ? p->wall_rough[0] : p->wall_rough[1];   // boundary.c:978-979
    const double roughJ = wf[2] == -1 ? p->wall_rough[2] : p->wall_rough[3];   // :1072-1073
    auto shumann = [&](int64_t c, int64_t cn, int64_t gdst) {
        double ni[3], nj[3], nk[3];
        unit_normals(csi[c], eta[c], zet[c], ni, nj, nk);
        const V3 n{nj[0], nj[1], nj[2]}, Ub = lucat[c], Uc = lucat[cn];
        const double db = Ub.x * n.x + Ub.y * n.y + Ub.z * n.z, dc = Uc.x * n.x + Uc.y * n.y + Uc.z * n.z;
        const V3 bN{db * n.x, db * n.y, db * n.z}, cN{dc * n.x, dc * n.y, dc * n.z};
        const V3 bP{Ub.x - bN.x, Ub.y - bN.y, Ub.z - bN.z}, cP{Uc.x - cN.x, Uc.y - cN.y, Uc.z - cN.z};
        V3 gh{2. * bP.x - cP.x, 2. * bP.y - cP.y, 2. * bP.z - cP.z};
        ucat[gdst] = V3{gh.x - bN.x, gh.y - bN.y, gh.z - bN.z};
    };
    for (int64_t k = 1; k < mz - 1; k++)                     // src/boundary.c:919-2029
        for (int64_t j = 1; j < my - 1; j++)
            for (int64_t i = 1; i < mx - 1; i++) {
                const int64_t c = l.L(k, j, i);
                const int64_t gi0 = l.L(k, j, i - 1), gi1 = l.L(k, j, i + 1), gj0 = l.L(k, j - 1, i),
                              gj1 = l.L(k, j + 1, i), gk0 = l.L(k - 1, j, i), gk1 = l.L(k + 1, j, i);
                if (k == 1 && inletK) {                          // boundary.c:1228-1300
                    const V3 cc = ((const V3 *)inlet->cent)[c];
                    double uMag = 0;
                    if (inlet->type == 5) {                      // Nieuwstadt model
                        ucat[gk0] = nieuwstadt(inlet, cc.z - inlet->grnd_level);
                    } else if (inlet->type == 1) {               // power law with the caller's draws
                        const double h = cc.z - inlet->grnd_level, alpha = 0.107027;
                        const double pw = std::pow(h / inlet->href, alpha);
                        V3 u{pw * inlet->uref[0], pw * inlet->uref[1], pw * inlet->uref[2]};
                        const double f = inlet->uprime_rms;
                        const int n1 = inlet->noise[j * mx + i];
                        const double turbulence = (1 + ((double)n1) / 10000. * f);
                        ucat[gk0] = V3{turbulence * u.x, turbulence * u.y, turbulence * u.z};
                    } else if (inlet->type == 3 || inlet->type == 4) {   // mapped inflow databases
                        ucat[gk0] = mapped_inflow(inlet, mx, j, i, cc.z - inlet->grnd_level);
                    } else if (inlet->type == 2) {               // log-law profile
                        const double h = cc.z - inlet->grnd_level, vkConstant = 0.4;
                        if (h <= inlet->h_inv) {
                            const double v = (inlet->u_tau / vkConstant) * std::log(h / inlet->roughness);
                            uMag = (v < 1e-5) ? 1e-5 : v;
                        } else {
                            uMag = (inlet->u_tau / vkConstant) * std::log(inlet->h_inv / inlet->roughness);
                        }
                    } else {                                     // sinusoidal inflow
                        const double fraction = (cc.y - inlet->ymin) / inlet->ly;
                        const double *u = inlet->uref;
                        uMag = (1.0 + inlet->amplitude * std::cos(inlet->periods * 2.0 * M_PI * fraction)) *
                               sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
                    }
                    if (inlet->type == 2 || inlet->type == 6)
                        ucat[gk0] = V3{uMag * inlet->udir[0], uMag * inlet->udir[1], uMag * inlet->udir[2]};
                }
                if (bc[0] == TOSCA_BCU_WALLFUNCTION && wf[0] == -1 && i == 1) cabot(c, 0, gi0, roughI);
                else if (bc[1] == TOSCA_BCU_WALLFUNCTION && wf[1] == -1 && i == mx - 2) cabot(c, 0, gi1, roughI);
                if (bc[2] == TOSCA_BCU_WALLFUNCTION && wf[2] == -1 && j == 1) cabot(c, 1, gj0, roughJ);
                else if (bc[3] == TOSCA_BCU_WALLFUNCTION && wf[3] == -1 && j == my - 2) cabot(c, 1, gj1, roughJ);
                if (bc[2] == TOSCA_BCU_WALLFUNCTION && wf[2] == -3 && j == 1) shumann(c, gj1, gj0);
                if (bc[3] == TOSCA_BCU_WALLFUNCTION && wf[3] == -3 && j == my - 2) shumann(c, gj0, gj1);
                if (bc[0] == TOSCA_BCU_FIXEDVALUE && i == 1) fix(0, gi0);
                if (bc[1] == TOSCA_BCU_FIXEDVALUE && i == mx - 2) fix(1, gi1);
                if (bc[2] == TOSCA_BCU_FIXEDVALUE && j == 1) fix(2, gj0);
                if (bc[3] == TOSCA_BCU_FIXEDVALUE && j == my - 2) fix(3, gj1);
                if (bc[4] == TOSCA_BCU_FIXEDVALUE && k == 1) fix(4, gk0);
                if (bc[5] == TOSCA_BCU_FIXEDVALUE && k == mz - 2) fix(5, gk1);
                if (bc[0] == TOSCA_BCU_SLIP && i == 1) slip(c, 0, gi0);
                if (bc[1] == TOSCA_BCU_SLIP && i == mx - 2) slip(c, 0, gi1);
                if (bc[2] == TOSCA_BCU_SLIP && j == 1) slip(c, 1, gj0);
                if (bc[3] == TOSCA_BCU_SLIP && j == my - 2) slip(c, 1, gj1);
                if (bc[4] == TOSCA_BCU_SLIP && k == 1) slip(c, 2, gk0);
                if (bc[5] == TOSCA_BCU_SLIP && k == mz - 2) slip(c, 2, gk1);
                if (i == 1 && bc[0] == TOSCA_BCU_NOSLIP) neg(c, gi0);
                if (i == mx - 2 && bc[1] == TOSCA_BCU_NOSLIP) neg(c, gi1);
                if (j == 1 && bc[2] == TOSCA_BCU_NOSLIP) neg(c, gj0);
                if (j == my - 2 && bc[3] == TOSCA_BCU_NOSLIP) neg(c, gj1);
                if (k == 1 && bc[4] == TOSCA_BCU_NOSLIP) neg(c, gk0);
                if (k == mz - 2 && bc[5] == TOSCA_BCU_NOSLIP) neg(c, gk1);
                if (bc[0] == TOSCA_BCU_ZEROGRADIENT && i == 1) zg(c, gi1, gi0);
                if (bc[1] == TOSCA_BCU_ZEROGRADIENT && i == mx - 2) zg(c, gi0, gi1);
                if (bc[2] == TOSCA_BCU_ZEROGRADIENT && j == 1) zg(c, gj1, gj0);
                if (bc[3] == TOSCA_BCU_ZEROGRADIENT && j == my - 2) zg(c, gj0, gj1);
                if (bc[4] == TOSCA_BCU_ZEROGRADIENT && k == 1) zg(c, gk1, gk0);
                if (bc[5] == TOSCA_BCU_ZEROGRADIENT && k == mz - 2) zg(c, gk0, gk1);
                if (bc[0] == TOSCA_BCU_PERIODIC && i == 1) {
                    if (l.i_per()) ucat[gi0] = lucat[l.L(k, j, mx - 2)];
                    else if (l.ii_per()) ucat[gi0] = lucat[l.L(k, j, -2)];
                }
                if (bc[1] == TOSCA_BCU_PERIODIC && i == mx - 2) {
                    if (l.i_per()) ucat[gi1] = lucat[l.L(k, j, 1)];
                    else if (l.ii_per()) ucat[gi1] = lucat[l.L(k, j, mx + 1)];
                }
                if (bc[2] == TOSCA_BCU_PERIODIC && j == 1) {
                    if (l.j_per()) ucat[gj0] = lucat[l.L(k, my - 2, i)];
                    else if (l.jj_per()) ucat[gj0] = lucat[l.L(k, -2, i)];
                }
                if (bc[3] == TOSCA_BCU_PERIODIC && j == my - 2) {
                    if (l.j_per()) ucat[gj1] = lucat[l.L(k, 1, i)];
                    else if (l.jj_per()) ucat[gj1] = lucat[l.L(k, my + 1, i)];
                }
                if (bc[4] == TOSCA_BCU_PERIODIC && k == 1) {
                    if (l.k_per()) ucat[gk0] = lucat[l.L(mz - 2, j, i)];
                    else if (l.kk_per()) ucat[gk0] = lucat[l.L(-2, j, i)];
                }
                if (bc[5] == TOSCA_BCU_PERIODIC && k == mz - 2) {
                    if (l.k_per()) ucat[gk1] = lucat[l.L(1, j, i)];
                    else if (l.kk_per()) ucat[gk1] = lucat[l.L(mz + 1, j, i)];
                }
            }
    global_to_local(l, ucat, lucat);                         // :2042-2043
    return TOSCA_OK;
}

int oracle_temperature_bcs(const tosca_dmda_info *info, const tosca_mesh *m, const tosca_params *p, const int *bct,
                           const double *tval, double *lt) {
    Lay l;
    int rc = check_layout(info, p, l);
    if (rc) return rc;
    for (int q = 0; q < 6; q++) {
        if (bct[q] < TOSCA_BCT_ZEROGRADIENT || bct[q] > TOSCA_BCT_SURFACETEMP) { g_err = "bad temperature BC kind"; return TOSCA_ERR_ARG; }
        if (bct[q] == TOSCA_BCT_SURFACETEMP && q != 2) { g_err = "surface temperature on a patch other than jLeft"; return TOSCA_ERR_ARG; }
    }
    const V3 *csi = (const V3 *)m->csi, *eta = (const V3 *)m->eta, *zet = (const V3 *)m->zet;
    const double *aj = m->aj;
    const int64_t mx = l.mx, my = l.my, mz = l.mz;
    std::vector<double> t(lt, lt + l.nloc());
    auto area = [](V3 v) { return sqrt(v.x * v.x + v.y * v.y + v.z * v.z); };
    for (int64_t k = 1; k < mz - 1; k++)                     // src/boundary.c:2173-2565
        for (int64_t j = 1; j < my - 1; j++)
            for (int64_t i = 1; i < mx - 1; i++) {
                const int64_t c = l.L(k, j, i);
                const int64_t g[6] = {l.L(k, j, i - 1), l.L(k, j, i + 1), l.L(k, j - 1, i), l.L(k, j + 1, i),
                                      l.L(k - 1, j, i), l.L(k + 1, j, i)};
                const bool on[6] = {i == 1, i == mx - 2, j == 1, j == my - 2, k == 1, k == mz - 2};
                for (int q = 0; q < 6; q++)
                    if (on[q] && bct[q] == TOSCA_BCT_ZEROGRADIENT) t[g[q]] = lt[c];
                if (on[2] && bct[2] == TOSCA_BCT_SURFACETEMP) t[g[2]] = 2.0 * tval[2] - lt[c];   // :2392-2410
                for (int q = 0; q < 6; q++)
                    if (on[q] && bct[q] == TOSCA_BCT_FIXEDVALUE) t[g[q]] = tval[q];
                for (int q = 0; q < 6; q++)
                    if (on[q] && bct[q] == TOSCA_BCT_FIXEDGRADIENT) {
                        const V3 b = (q < 2) ? csi[c] : (q < 4) ? eta[c] : zet[c];
                        const double d = (1.0 / aj[c]) / area(b);
                        t[g[q]] = d * tval[q] + lt[c];
                    }
                if (bct[0] == TOSCA_BCT_PERIODIC && on[0]) {
                    if (l.i_per()) t[g[0]] = lt[l.L(k, j, mx - 2)];
                    else if (l.ii_per()) t[g[0]] = lt[l.L(k, j, -2)];
                }
                if (bct[1] == TOSCA_BCT_PERIODIC && on[1]) {
                    if (l.i_per()) t[g[1]] = lt[l.L(k, j, 1)];
                    else if (l.ii_per()) t[g[1]] = lt[l.L(k, j, mx + 1)];
                }
                if (bct[2] == TOSCA_BCT_PERIODIC && on[2]) {
                    if (l.j_per()) t[g[2]] = lt[l.L(k, my - 2, i)];
                    else if (l.jj_per()) t[g[2]] = lt[l.L(k, -2, i)];
                }
                if (bct[3] == TOSCA_BCT_PERIODIC && on[3]) {
                    if (l.j_per()) t[g[3]] = lt[l.L(k, 1, i)];
                    else if (l.jj_per()) t[g[3]] = lt[l.L(k, my + 1, i)];
                }
                if (bct[4] == TOSCA_BCT_PERIODIC && on[4]) {
                    if (l.k_per()) t[g[4]] = lt[l.L(mz - 2, j, i)];
                    else if (l.kk_per()) t[g[4]] = lt[l.L(-2, j, i)];
                }
                if (bct[5] == TOSCA_BCT_PERIODIC && on[5]) {
                    if (l.k_per()) t[g[5]] = lt[l.L(1, j, i)];
                    else if (l.kk_per()) t[g[5]] = lt[l.L(mz + 1, j, i)];
                }
            }
    for (int64_t k = 0; k < mz; k++)                          // DMGlobalToLocal (:2622-2623)
        for (int64_t j = 0; j < my; j++)
            for (int64_t i = 0; i < mx; i++) lt[l.L(k, j, i)] = t[l.L(k, j, i)];
    local_to_local(l, lt);
    return TOSCA_OK;
}

// ================================================================================================
// Wind-farm body force of uniform actuator disks
// ================================================================================================

// projectBladeForce, uniformADM branch (src/turbines/turbines_force.c:717-932): VecSet(lsourceFarmCat,
// 0); for each rotor t, each disk point p, each controlled cell c of t (interior cells, k-j-i order,
// with |cent - rotCenter| <= rTip + eps prjNSigma, turbines_parallel.c:57-121): r = |x_p - cent|,
// and where r < eps prjNSigma, sCat += (exp(-(r/eps)^2) / (pow(eps, 3) pow(pi, 1.5))) B_p; then
// DMLocalToLocal (:1184-1185).
int oracle_project_uadm(const tosca_dmda_info *info, const tosca_params *p, const double *cent_, int nrot,
                        int npts, double nsigma, const double *center, const double *sphere, const double *eps_,
                        const double *points, const double *force, double *scat_) {
    Lay l;
    int rc = check_layout(info, p, l);
    if (rc) return rc;
    const V3 *cent = (const V3 *)cent_;
    V3 *scat = (V3 *)scat_;
    for (int64_t q = 0; q < l.nloc(); q++) scat[q] = V3{0, 0, 0};
    for (int t = 0; t < nrot; t++) {
        const double rPrj = eps_[t] * nsigma, eps = eps_[t];
        std::vector<int64_t> cells;                                   // the sphere's controlled cells
        for (int64_t k = 1; k < l.mz - 1; k++)
            for (int64_t j = 1; j < l.my - 1; j++)
                for (int64_t i = 1; i < l.mx - 1; i++) {
                    const V3 c = cent[l.L(k, j, i)];
                    const V3 d{c.x - center[3 * t], c.y - center[3 * t + 1], c.z - center[3 * t + 2]};
                    if (sqrt(d.x * d.x + d.y * d.y + d.z * d.z) <= sphere[t]) cells.push_back(l.L(k, j, i));
                }
        for (int pp = 0; pp < npts; pp++) {
            const double *xp = points + 3 * ((int64_t)t * npts + pp), *B = force + 3 * ((int64_t)t * npts + pp);
            for (int64_t c : cells) {
                const V3 r_c{xp[0] - cent[c].x, xp[1] - cent[c].y, xp[2] - cent[c].z};
                const double r = sqrt(r_c.x * r_c.x + r_c.y * r_c.y + r_c.z * r_c.z);
                if (r < rPrj) {
                    const double pf = std::exp(-(r / eps) * (r / eps)) / (pow(eps, 3) * pow(M_PI, 1.5));
                    scat[c].x += pf * B[0];
                    scat[c].y += pf * B[1];
                    scat[c].z += pf * B[2];
                }
            }
        }
    }
    local_to_local(l, scat);
    return TOSCA_OK;
}

// bodyForceCartesian2Contravariant (src/turbines/turbines_force.c:1741-1835): VecSet(sourceFarmCont,
// 0); at the interior cells the Cartesian force averaged to the +i, +j, +k faces, dotted with
// icsi, jeta, kzet.
int oracle_body_force_to_cont(const tosca_dmda_info *info, const tosca_mesh *m, const tosca_params *p,
                              const double *scat_, double *scont_) {
    Lay l;
    int rc = check_layout(info, p, l);
    if (rc) return rc;
    const V3 *s = (const V3 *)scat_, *icsi = (const V3 *)m->icsi, *jeta = (const V3 *)m->jeta,
             *kzet = (const V3 *)m->kzet;
    V3 *sc = (V3 *)scont_;
    for (int64_t g = 0; g < l.mx * l.my * l.mz; g++) sc[g] = V3{0, 0, 0};
    for (int64_t k = 1; k < l.mz - 1; k++)
        for (int64_t j = 1; j < l.my - 1; j++)
            for (int64_t i = 1; i < l.mx - 1; i++) {
                const V3 a = s[l.L(k, j, i)], bi = s[l.L(k, j, i + 1)], bj = s[l.L(k, j + 1, i)], bk = s[l.L(k + 1, j, i)];
                const V3 fI{0.5 * (a.x + bi.x), 0.5 * (a.y + bi.y), 0.5 * (a.z + bi.z)};
                const V3 fJ{0.5 * (a.x + bj.x), 0.5 * (a.y + bj.y), 0.5 * (a.z + bj.z)};
                const V3 fK{0.5 * (a.x + bk.x), 0.5 * (a.y + bk.y), 0.5 * (a.z + bk.z)};
                const V3 n1 = icsi[l.L(k, j, i)], n2 = jeta[l.L(k, j, i)], n3 = kzet[l.L(k, j, i)];
                V3 &o = sc[l.G(k, j, i)];
                o.x = (fI.x * n1.x + fI.y * n1.y + fI.z * n1.z);
                o.y = (fJ.x * n2.x + fJ.y * n2.y + fJ.z * n2.z);
                o.z = (fK.x * n3.x + fK.y * n3.y + fK.z * n3.z);
            }
    return TOSCA_OK;
}

// tosca_rhs_les_step with the farm source: VecAXPY(Rhs, 1, sourceFarmCont), then the step
int oracle_rhs_les_step_farm(const tosca_dmda_info *info, const tosca_mesh *m, const tosca_params *p,
                             const double *ucont, const double *ucat, const double *theta, const double *farm,
                             double *rhs, double scale, double *cs, double *nut, double *btheta, double buoy_scale) {
    Lay l;
    int rc = check_layout(info, p, l);
    if (rc) return rc;
    axpy(l.mx * l.my * l.mz, 1.0, (const V3 *)farm, (V3 *)rhs);
    return oracle_rhs_les_step(info, m, p, ucont, ucat, theta, rhs, scale, cs, nut, btheta, buoy_scale);
}
#endif  // ORACLE_TANGENT

#ifndef ORACLE_TANGENT
}  // extern "C"
#else
}  // namespace
#endif
