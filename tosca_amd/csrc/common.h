// common.h -- device-side building blocks shared by the gfx950 kernels of the momentum
// right-hand side + LES library.  Index conventions are those of the reference's PETSc DMDA
// local arrays: [k][j][i], i fastest, ghost width 3, vectors AoS {x,y,z}.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/tosca_rhs.h"

namespace tosca {

constexpr int GW = 3;   // DMDA ghost width (src/mesh.c:42)

// vectors of the flow-field type T (double on the FP64 path, float for the FP32 flow fields of the
// mixed path); metrics are always double
template <class T> struct Vec3 { T x, y, z; };
using V3 = Vec3<double>;
using V3f = Vec3<float>;
template <class T, class S> __host__ __device__ __forceinline__ Vec3<T> cvt3(Vec3<S> a) {
    return Vec3<T>{static_cast<T>(a.x), static_cast<T>(a.y), static_cast<T>(a.z)};
}

// Grid geometry passed by value to every kernel.
struct Geo {
    int mx, my, mz;          // DMDA sizes
    int gxm, gym, gzm;       // ghosted sizes (m + 6)
    int per[3];              // periodic type on i, j, k
    // local (ghosted) linear index of DMDA point (k, j, i)
    __host__ __device__ __forceinline__ int L(int k, int j, int i) const {
        return ((k + GW) * gym + (j + GW)) * gxm + (i + GW);
    }
    // global (owned range) linear index
    __host__ __device__ __forceinline__ int G(int k, int j, int i) const { return (k * my + j) * mx + i; }
    __host__ __device__ __forceinline__ bool nonper(int ax) const { return per[ax] == TOSCA_PERIODIC_NONE; }
    __host__ __device__ __forceinline__ int nloc() const { return gxm * gym * gzm; }
};

template <class T> __device__ __forceinline__ Vec3<T> ld3(const T *p, int c) {
    return Vec3<T>{p[3 * c], p[3 * c + 1], p[3 * c + 2]};
}
__device__ __forceinline__ void st3(double *p, int c, V3 v) { p[3 * c] = v.x; p[3 * c + 1] = v.y; p[3 * c + 2] = v.z; }

template <class T> struct GradT { T dudc, dvdc, dwdc, dude, dvde, dwde, dudz, dvdz, dwdz; };
template <class T> struct CartT { T du_dx, du_dy, du_dz, dv_dx, dv_dy, dv_dz, dw_dx, dw_dy, dw_dz; };
using Grad = GradT<double>;
using Cart = CartT<double>;

// Velocity derivatives along the curvilinear directions at an i-face (Compute_du_i,
// src/include/inline.h:2748-2806) with no IBM / overset cells.
template <class T>
__device__ __forceinline__ GradT<T> du_i(const Geo &g, const T *u, int i, int j, int k) {
    const bool iN = g.nonper(0), jN = g.nonper(1), kN = g.nonper(2);
    GradT<T> r;
    Vec3<T> a = ld3(u, g.L(k, j, i + 1)), b = ld3(u, g.L(k, j, i));
    r.dudc = a.x - b.x; r.dvdc = a.y - b.y; r.dwdc = a.z - b.z;
    const bool ib = (i == 0 || i == g.mx - 2);
    if (j == g.my - 2 && ib && (jN || iN)) {
        Vec3<T> p0 = ld3(u, g.L(k, j, i + 1)), p1 = ld3(u, g.L(k, j, i)), m0 = ld3(u, g.L(k, j - 1, i + 1)), m1 = ld3(u, g.L(k, j - 1, i));
        r.dude = (p0.x + p1.x - m0.x - m1.x) * T(0.5); r.dvde = (p0.y + p1.y - m0.y - m1.y) * T(0.5); r.dwde = (p0.z + p1.z - m0.z - m1.z) * T(0.5);
    } else if (j == 1 && ib && (jN || iN)) {
        Vec3<T> p0 = ld3(u, g.L(k, j + 1, i + 1)), p1 = ld3(u, g.L(k, j + 1, i)), m0 = ld3(u, g.L(k, j, i + 1)), m1 = ld3(u, g.L(k, j, i));
        r.dude = (p0.x + p1.x - m0.x - m1.x) * T(0.5); r.dvde = (p0.y + p1.y - m0.y - m1.y) * T(0.5); r.dwde = (p0.z + p1.z - m0.z - m1.z) * T(0.5);
    } else {
        Vec3<T> p0 = ld3(u, g.L(k, j + 1, i + 1)), p1 = ld3(u, g.L(k, j + 1, i)), m0 = ld3(u, g.L(k, j - 1, i + 1)), m1 = ld3(u, g.L(k, j - 1, i));
        r.dude = (p0.x + p1.x - m0.x - m1.x) * T(0.25); r.dvde = (p0.y + p1.y - m0.y - m1.y) * T(0.25); r.dwde = (p0.z + p1.z - m0.z - m1.z) * T(0.25);
    }
    if (k == g.mz - 2 && ib && (kN || iN)) {
        Vec3<T> p0 = ld3(u, g.L(k, j, i + 1)), p1 = ld3(u, g.L(k, j, i)), m0 = ld3(u, g.L(k - 1, j, i + 1)), m1 = ld3(u, g.L(k - 1, j, i));
        r.dudz = (p0.x + p1.x - m0.x - m1.x) * T(0.5); r.dvdz = (p0.y + p1.y - m0.y - m1.y) * T(0.5); r.dwdz = (p0.z + p1.z - m0.z - m1.z) * T(0.5);
    } else if (k == 1 && ib && (kN || iN)) {
        Vec3<T> p0 = ld3(u, g.L(k + 1, j, i + 1)), p1 = ld3(u, g.L(k + 1, j, i)), m0 = ld3(u, g.L(k, j, i + 1)), m1 = ld3(u, g.L(k, j, i));
        r.dudz = (p0.x + p1.x - m0.x - m1.x) * T(0.5); r.dvdz = (p0.y + p1.y - m0.y - m1.y) * T(0.5); r.dwdz = (p0.z + p1.z - m0.z - m1.z) * T(0.5);
    } else {
        Vec3<T> p0 = ld3(u, g.L(k + 1, j, i + 1)), p1 = ld3(u, g.L(k + 1, j, i)), m0 = ld3(u, g.L(k - 1, j, i + 1)), m1 = ld3(u, g.L(k - 1, j, i));
        r.dudz = (p0.x + p1.x - m0.x - m1.x) * T(0.25); r.dvdz = (p0.y + p1.y - m0.y - m1.y) * T(0.25); r.dwdz = (p0.z + p1.z - m0.z - m1.z) * T(0.25);
    }
    return r;
}

// Compute_du_j (src/include/inline.h:2810-2868)
template <class T>
__device__ __forceinline__ GradT<T> du_j(const Geo &g, const T *u, int i, int j, int k) {
    const bool iN = g.nonper(0), jN = g.nonper(1), kN = g.nonper(2);
    GradT<T> r;
    const bool jb = (j == 0 || j == g.my - 2);
    if (i == g.mx - 2 && jb && (iN || jN)) {
        Vec3<T> p0 = ld3(u, g.L(k, j + 1, i)), p1 = ld3(u, g.L(k, j, i)), m0 = ld3(u, g.L(k, j + 1, i - 1)), m1 = ld3(u, g.L(k, j, i - 1));
        r.dudc = (p0.x + p1.x - m0.x - m1.x) * T(0.5); r.dvdc = (p0.y + p1.y - m0.y - m1.y) * T(0.5); r.dwdc = (p0.z + p1.z - m0.z - m1.z) * T(0.5);
    } else if (i == 1 && jb && (iN || jN)) {
        Vec3<T> p0 = ld3(u, g.L(k, j + 1, i + 1)), p1 = ld3(u, g.L(k, j, i + 1)), m0 = ld3(u, g.L(k, j + 1, i)), m1 = ld3(u, g.L(k, j, i));
        r.dudc = (p0.x + p1.x - m0.x - m1.x) * T(0.5); r.dvdc = (p0.y + p1.y - m0.y - m1.y) * T(0.5); r.dwdc = (p0.z + p1.z - m0.z - m1.z) * T(0.5);
    } else {
        Vec3<T> p0 = ld3(u, g.L(k, j + 1, i + 1)), p1 = ld3(u, g.L(k, j, i + 1)), m0 = ld3(u, g.L(k, j + 1, i - 1)), m1 = ld3(u, g.L(k, j, i - 1));
        r.dudc = (p0.x + p1.x - m0.x - m1.x) * T(0.25); r.dvdc = (p0.y + p1.y - m0.y - m1.y) * T(0.25); r.dwdc = (p0.z + p1.z - m0.z - m1.z) * T(0.25);
    }
    Vec3<T> a = ld3(u, g.L(k, j + 1, i)), b = ld3(u, g.L(k, j, i));
    r.dude = a.x - b.x; r.dvde = a.y - b.y; r.dwde = a.z - b.z;
    if (k == g.mz - 2 && jb && (kN || jN)) {
        Vec3<T> p0 = ld3(u, g.L(k, j + 1, i)), p1 = ld3(u, g.L(k, j, i)), m0 = ld3(u, g.L(k - 1, j + 1, i)), m1 = ld3(u, g.L(k - 1, j, i));
        r.dudz = (p0.x + p1.x - m0.x - m1.x) * T(0.5); r.dvdz = (p0.y + p1.y - m0.y - m1.y) * T(0.5); r.dwdz = (p0.z + p1.z - m0.z - m1.z) * T(0.5);
    } else if (k == 1 && jb && (kN || jN)) {
        Vec3<T> p0 = ld3(u, g.L(k + 1, j + 1, i)), p1 = ld3(u, g.L(k + 1, j, i)), m0 = ld3(u, g.L(k, j + 1, i)), m1 = ld3(u, g.L(k, j, i));
        r.dudz = (p0.x + p1.x - m0.x - m1.x) * T(0.5); r.dvdz = (p0.y + p1.y - m0.y - m1.y) * T(0.5); r.dwdz = (p0.z + p1.z - m0.z - m1.z) * T(0.5);
    } else {
        Vec3<T> p0 = ld3(u, g.L(k + 1, j + 1, i)), p1 = ld3(u, g.L(k + 1, j, i)), m0 = ld3(u, g.L(k - 1, j + 1, i)), m1 = ld3(u, g.L(k - 1, j, i));
        r.dudz = (p0.x + p1.x - m0.x - m1.x) * T(0.25); r.dvdz = (p0.y + p1.y - m0.y - m1.y) * T(0.25); r.dwdz = (p0.z + p1.z - m0.z - m1.z) * T(0.25);
    }
    return r;
}

// Compute_du_k (src/include/inline.h:2872-2931)
template <class T>
__device__ __forceinline__ GradT<T> du_k(const Geo &g, const T *u, int i, int j, int k) {
    const bool iN = g.nonper(0), jN = g.nonper(1), kN = g.nonper(2);
    GradT<T> r;
    const bool kb = (k == 0 || k == g.mz - 2);
    if (i == g.mx - 2 && kb && (iN || kN)) {
        Vec3<T> p0 = ld3(u, g.L(k + 1, j, i)), p1 = ld3(u, g.L(k, j, i)), m0 = ld3(u, g.L(k + 1, j, i - 1)), m1 = ld3(u, g.L(k, j, i - 1));
        r.dudc = (p0.x + p1.x - m0.x - m1.x) * T(0.5); r.dvdc = (p0.y + p1.y - m0.y - m1.y) * T(0.5); r.dwdc = (p0.z + p1.z - m0.z - m1.z) * T(0.5);
    } else if (i == 1 && kb && (iN || kN)) {
        Vec3<T> p0 = ld3(u, g.L(k + 1, j, i + 1)), p1 = ld3(u, g.L(k, j, i + 1)), m0 = ld3(u, g.L(k + 1, j, i)), m1 = ld3(u, g.L(k, j, i));
        r.dudc = (p0.x + p1.x - m0.x - m1.x) * T(0.5); r.dvdc = (p0.y + p1.y - m0.y - m1.y) * T(0.5); r.dwdc = (p0.z + p1.z - m0.z - m1.z) * T(0.5);
    } else {
        Vec3<T> p0 = ld3(u, g.L(k + 1, j, i + 1)), p1 = ld3(u, g.L(k, j, i + 1)), m0 = ld3(u, g.L(k + 1, j, i - 1)), m1 = ld3(u, g.L(k, j, i - 1));
        r.dudc = (p0.x + p1.x - m0.x - m1.x) * T(0.25); r.dvdc = (p0.y + p1.y - m0.y - m1.y) * T(0.25); r.dwdc = (p0.z + p1.z - m0.z - m1.z) * T(0.25);
    }
    if (j == g.my - 2 && kb && (jN || kN)) {
        Vec3<T> p0 = ld3(u, g.L(k + 1, j, i)), p1 = ld3(u, g.L(k, j, i)), m0 = ld3(u, g.L(k + 1, j - 1, i)), m1 = ld3(u, g.L(k, j - 1, i));
        r.dude = (p0.x + p1.x - m0.x - m1.x) * T(0.5); r.dvde = (p0.y + p1.y - m0.y - m1.y) * T(0.5); r.dwde = (p0.z + p1.z - m0.z - m1.z) * T(0.5);
    } else if (j == 1 && kb && (jN || kN)) {
        Vec3<T> p0 = ld3(u, g.L(k + 1, j + 1, i)), p1 = ld3(u, g.L(k, j + 1, i)), m0 = ld3(u, g.L(k + 1, j, i)), m1 = ld3(u, g.L(k, j, i));
        r.dude = (p0.x + p1.x - m0.x - m1.x) * T(0.5); r.dvde = (p0.y + p1.y - m0.y - m1.y) * T(0.5); r.dwde = (p0.z + p1.z - m0.z - m1.z) * T(0.5);
    } else {
        Vec3<T> p0 = ld3(u, g.L(k + 1, j + 1, i)), p1 = ld3(u, g.L(k, j + 1, i)), m0 = ld3(u, g.L(k + 1, j - 1, i)), m1 = ld3(u, g.L(k, j - 1, i));
        r.dude = (p0.x + p1.x - m0.x - m1.x) * T(0.25); r.dvde = (p0.y + p1.y - m0.y - m1.y) * T(0.25); r.dwde = (p0.z + p1.z - m0.z - m1.z) * T(0.25);
    }
    Vec3<T> a = ld3(u, g.L(k + 1, j, i)), b = ld3(u, g.L(k, j, i));
    r.dudz = a.x - b.x; r.dvdz = a.y - b.y; r.dwdz = a.z - b.z;
    return r;
}

// Cell-centred velocity derivatives (Compute_du_center, src/include/inline.h:3344-3418), from the
// cell's value c and its six neighbours: central differences, one-sided at the first and last
// interior cell of a non-periodic direction.
template <class T>
__device__ __forceinline__ GradT<T> du_center_v(const Geo &g, int i, int j, int k, Vec3<T> c, Vec3<T> im, Vec3<T> ip,
                                                Vec3<T> jm, Vec3<T> jp, Vec3<T> km, Vec3<T> kp) {
    GradT<T> r;
    if (g.nonper(0) && i == g.mx - 2) { r.dudc = c.x - im.x; r.dvdc = c.y - im.y; r.dwdc = c.z - im.z; }
    else if (g.nonper(0) && i == 1) { r.dudc = ip.x - c.x; r.dvdc = ip.y - c.y; r.dwdc = ip.z - c.z; }
    else { r.dudc = (ip.x - im.x) * T(0.5); r.dvdc = (ip.y - im.y) * T(0.5); r.dwdc = (ip.z - im.z) * T(0.5); }
    if (g.nonper(1) && j == g.my - 2) { r.dude = c.x - jm.x; r.dvde = c.y - jm.y; r.dwde = c.z - jm.z; }
    else if (g.nonper(1) && j == 1) { r.dude = jp.x - c.x; r.dvde = jp.y - c.y; r.dwde = jp.z - c.z; }
    else { r.dude = (jp.x - jm.x) * T(0.5); r.dvde = (jp.y - jm.y) * T(0.5); r.dwde = (jp.z - jm.z) * T(0.5); }
    if (g.nonper(2) && k == g.mz - 2) { r.dudz = c.x - km.x; r.dvdz = c.y - km.y; r.dwdz = c.z - km.z; }
    else if (g.nonper(2) && k == 1) { r.dudz = kp.x - c.x; r.dvdz = kp.y - c.y; r.dwdz = kp.z - c.z; }
    else { r.dudz = (kp.x - km.x) * T(0.5); r.dvdz = (kp.y - km.y) * T(0.5); r.dwdz = (kp.z - km.z) * T(0.5); }
    return r;
}

template <class T>
__device__ __forceinline__ GradT<T> du_center(const Geo &g, const T *u, int i, int j, int k) {
    return du_center_v(g, i, j, k, ld3(u, g.L(k, j, i)), ld3(u, g.L(k, j, i - 1)), ld3(u, g.L(k, j, i + 1)),
                       ld3(u, g.L(k, j - 1, i)), ld3(u, g.L(k, j + 1, i)), ld3(u, g.L(k - 1, j, i)),
                       ld3(u, g.L(k + 1, j, i)));
}

// Compute_du_dxyz (src/include/inline.h:3635-3663)
template <class T>
__device__ __forceinline__ CartT<T> du_dxyz(Vec3<T> c, Vec3<T> e, Vec3<T> z, T ajc, const GradT<T> &g) {
    CartT<T> r;
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

// |S| = sqrt(2 S_ij S_ij) (src/les.c:342-354)
template <class T>
__device__ __forceinline__ T strain_mag(const CartT<T> &d) {
    const T h = T(0.5);
    T Sxx = h * (d.du_dx + d.du_dx), Sxy = h * (d.du_dy + d.dv_dx), Sxz = h * (d.du_dz + d.dw_dx);
    T Syy = h * (d.dv_dy + d.dv_dy), Syz = h * (d.dv_dz + d.dw_dy), Szz = h * (d.dw_dz + d.dw_dz);
    return sqrt(T(2.0) * (Sxx * Sxx + Sxy * Sxy + Sxz * Sxz + Sxy * Sxy + Syy * Syy + Syz * Syz +
                          Sxz * Sxz + Syz * Syz + Szz * Szz));
}

// The viscous flux of one face (ueqn_solvers.c:196-215 + 386-392 and their j/k twins): dir is the
// face direction (0 = i, 1 = j, 2 = k); csi/eta/zet are the face's contravariant basis.
template <class T>
__device__ __forceinline__ Vec3<T> visc_flux(const GradT<T> &g, Vec3<T> c, Vec3<T> e, Vec3<T> z, int dir, T ajc, T nuEff) {
    T g11, g21, g31, n0, n1, n2;
    if (dir == 0) {
        g11 = c.x * c.x + c.y * c.y + c.z * c.z; g21 = e.x * c.x + e.y * c.y + e.z * c.z; g31 = z.x * c.x + z.y * c.y + z.z * c.z;
        n0 = c.x; n1 = c.y; n2 = c.z;
    } else if (dir == 1) {
        g11 = c.x * e.x + c.y * e.y + c.z * e.z; g21 = e.x * e.x + e.y * e.y + e.z * e.z; g31 = z.x * e.x + z.y * e.y + z.z * e.z;
        n0 = e.x; n1 = e.y; n2 = e.z;
    } else {
        g11 = c.x * z.x + c.y * z.y + c.z * z.z; g21 = e.x * z.x + e.y * z.y + e.z * z.z; g31 = z.x * z.x + z.y * z.y + z.z * z.z;
        n0 = z.x; n1 = z.y; n2 = z.z;
    }
    T r11 = g.dudc * c.x + g.dude * e.x + g.dudz * z.x;
    T r21 = g.dvdc * c.x + g.dvde * e.x + g.dvdz * z.x;
    T r31 = g.dwdc * c.x + g.dwde * e.x + g.dwdz * z.x;
    T r12 = g.dudc * c.y + g.dude * e.y + g.dudz * z.y;
    T r22 = g.dvdc * c.y + g.dvde * e.y + g.dvdz * z.y;
    T r32 = g.dwdc * c.y + g.dwde * e.y + g.dwdz * z.y;
    T r13 = g.dudc * c.z + g.dude * e.z + g.dudz * z.z;
    T r23 = g.dvdc * c.z + g.dvde * e.z + g.dvdz * z.z;
    T r33 = g.dwdc * c.z + g.dwde * e.z + g.dwdz * z.z;
    Vec3<T> o;
    o.x = (g11 * g.dudc + g21 * g.dude + g31 * g.dudz + r11 * n0 + r21 * n1 + r31 * n2) * ajc * (nuEff);
    o.y = (g11 * g.dvdc + g21 * g.dvde + g31 * g.dvdz + r12 * n0 + r22 * n1 + r32 * n2) * ajc * (nuEff);
    o.z = (g11 * g.dwdc + g21 * g.dwde + g31 * g.dwdz + r13 * n0 + r23 * n1 + r33 * n2) * ajc * (nuEff);
    return o;
}

// QUICK (quadraticUpwindVec, src/include/inline.h:5089-5109): upwind on the sign of the flux.
template <class T>
__device__ __forceinline__ Vec3<T> quick(Vec3<T> f0, Vec3<T> f1, Vec3<T> f2, Vec3<T> f3, T ws) {
    Vec3<T> UU, U, D;
    if (ws > 0) { UU = f0; U = f1; D = f2; } else { UU = f3; U = f2; D = f1; }
    const T h = T(0.5), e = T(0.125), two = T(2.0);
    Vec3<T> r;
    r.x = h * (U.x + D.x) + e * (two * U.x - (D.x + UU.x));
    r.y = h * (U.y + D.y) + e * (two * U.y - (D.y + UU.y));
    r.z = h * (U.z + D.z) + e * (two * U.z - (D.z + UU.z));
    return r;
}

// Outer cells of the 4-cell QUICK stencil of face f (getFace2Cell4Stencil*, inline.h:2190-2380)
__device__ __forceinline__ void face2cell4(int ptype, int f, int m, int &L, int &R) {
    L = f - 1; R = f + 2;
    if (f == 0 || f == m - 2) {
        if (ptype == TOSCA_PERIODIC_COPY) { L = m - 3; R = 2; }
        else if (ptype == TOSCA_PERIODIC_WRAP) { if (f == m - 2) { R = m + 2; L = f - 1; } else { L = -3; R = f + 2; } }
        else { if (f == 0) { L = f; R = f + 2; } else { L = f - 1; R = f + 1; } }
    }
}

// The Germano contraction of les.c:715-799 in three stages, so that M and G can be formed by
// different threads of a workgroup: M from the filtered strain terms, G from the cell metrics,
// then num = L_ba M_aq G_bq and den = M_nm M_nl G_ml.  pow(x, 2.) of the reference is x * x (exactly rounded); the
// inverse of the metric matrix uses one reciprocal of its determinant.
__device__ __forceinline__ void germano_M(const double Sh[3][3], const double SSh[3][3], double filter,
                                          double test_filter, V3 c0, V3 e0, V3 z0, double M[3][3]) {
    double S_hat = 0;
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) S_hat += Sh[a][b] * Sh[a][b];
    S_hat = sqrt(2 * S_hat);
    const double tf2 = test_filter * test_filter, f2 = filter * filter;
    for (int a = 0; a < 3; a++) {
        double Mc0 = (-tf2 * S_hat * Sh[a][0] + f2 * SSh[a][0]);
        double Mc1 = (-tf2 * S_hat * Sh[a][1] + f2 * SSh[a][1]);
        double Mc2 = (-tf2 * S_hat * Sh[a][2] + f2 * SSh[a][2]);
        M[a][0] = Mc0 * c0.x + Mc1 * c0.y + Mc2 * c0.z;
        M[a][1] = Mc0 * e0.x + Mc1 * e0.y + Mc2 * e0.z;
        M[a][2] = Mc0 * z0.x + Mc1 * z0.y + Mc2 * z0.z;
    }
}

__device__ __forceinline__ void germano_G(V3 c0, V3 e0, V3 z0, double G[3][3]) {
    const double a11 = c0.x, a12 = c0.y, a13 = c0.z, a21 = e0.x, a22 = e0.y, a23 = e0.z, a31 = z0.x, a32 = z0.y, a33 = z0.z;
    double det = a11 * (a33 * a22 - a32 * a23) - a21 * (a33 * a12 - a32 * a13) + a31 * (a23 * a12 - a22 * a13);
    const double rd = 1.0 / det;
    double gc[3][3];
    gc[0][0] = (a33 * a22 - a32 * a23) * rd, gc[0][1] = -(a33 * a12 - a32 * a13) * rd, gc[0][2] = (a23 * a12 - a22 * a13) * rd;
    gc[1][0] = -(a33 * a21 - a31 * a23) * rd, gc[1][1] = (a33 * a11 - a31 * a13) * rd, gc[1][2] = -(a23 * a11 - a21 * a13) * rd;
    gc[2][0] = (a32 * a21 - a31 * a22) * rd, gc[2][1] = -(a32 * a11 - a31 * a12) * rd, gc[2][2] = (a22 * a11 - a21 * a12) * rd;
    double xcsi = gc[0][0], xeta = gc[0][1], xzet = gc[0][2];
    double ycsi = gc[1][0], yeta = gc[1][1], yzet = gc[1][2];
    double zcsi = gc[2][0], zeta = gc[2][1], zzet = gc[2][2];
    G[0][0] = xcsi * xcsi + ycsi * ycsi + zcsi * zcsi;
    G[1][1] = xeta * xeta + yeta * yeta + zeta * zeta;
    G[2][2] = xzet * xzet + yzet * yzet + zzet * zzet;
    G[0][1] = G[1][0] = xeta * xcsi + yeta * ycsi + zeta * zcsi;
    G[0][2] = G[2][0] = xzet * xcsi + yzet * ycsi + zzet * zcsi;
    G[1][2] = G[2][1] = xeta * xzet + yeta * yzet + zeta * zzet;
}

__device__ __forceinline__ void germano_contract(const double L[3][3], const double M[3][3], const double G[3][3],
                                                 double &num_out, double &den_out) {
    double num = 0, den = 0;
    for (int q = 0; q < 3; q++)
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) num += L[b][a] * M[a][q] * G[b][q];
    for (int m = 0; m < 3; m++)
        for (int n = 0; n < 3; n++)
            for (int l = 0; l < 3; l++) den += M[n][m] * M[n][l] * G[m][l];
    num_out = num;
    den_out = den;
}

__device__ __forceinline__ void germano(const double L[3][3], const double Sh[3][3], const double SSh[3][3], double filter,
                        double test_filter, V3 c0, V3 e0, V3 z0, double &num_out, double &den_out) {
    double M[3][3], G[3][3];
    germano_M(Sh, SSh, filter, test_filter, c0, e0, z0, M);
    germano_G(c0, e0, z0, G);
    germano_contract(L, M, G, num_out, den_out);
}

// x^(-1/3) for normal x > 0 without a division: x = y 2^(3k) with y in [1/8, 8), then Newton's
// iteration for r^-3 = y, r <- r (4 - y r^3) / 3, which converges quadratically, from the FP32
// estimate 2^(-log2(y) / 3) (relative error ~1e-7); two steps leave a few FP64 ulp; the result is
// r 2^(-k).
__device__ __forceinline__ double rcbrt_pos(double x) {
    const int e = __builtin_amdgcn_frexp_exp(x);       // x = m 2^e, m in [1/2, 1)
    const int k = (e >= 0 ? e : e - 2) / 3;            // floor(e / 3)
    const double y = __builtin_amdgcn_ldexp(x, -3 * k);
    double r = double(__builtin_amdgcn_exp2f(-__builtin_amdgcn_logf(float(y)) * (1.0f / 3.0f)));
#pragma unroll
    for (int it = 0; it < 2; it++) {
        const double r3 = r * r * r;
        r = r * (4.0 - y * r3) * (1.0 / 3.0);
    }
    return __builtin_amdgcn_ldexp(r, -k);
}

// The Germano contraction with the filter widths given squared (f2 = Delta^2, tf2 = Delta~^2), for
// the fused kernels: M as in germano_M, then num = sum_ab L_ba P_ab with P = M G^T, and
// den = sum_n M_n. (G M_n.), the same sums as germano_contract grouped so that they take 72 instead
// of 108 multiply-adds.
__device__ __forceinline__ void germano_sq(const double L[3][3], const double Sh[3][3], const double SSh[3][3],
                                           double f2, double tf2, V3 c0, V3 e0, V3 z0, double &num_out,
                                           double &den_out) {
    double S_hat = 0;
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) S_hat += Sh[a][b] * Sh[a][b];
    S_hat = sqrt(2 * S_hat);
    double M[3][3], G[3][3];
    for (int a = 0; a < 3; a++) {
        const double Mc0 = (-tf2 * S_hat * Sh[a][0] + f2 * SSh[a][0]);
        const double Mc1 = (-tf2 * S_hat * Sh[a][1] + f2 * SSh[a][1]);
        const double Mc2 = (-tf2 * S_hat * Sh[a][2] + f2 * SSh[a][2]);
        M[a][0] = Mc0 * c0.x + Mc1 * c0.y + Mc2 * c0.z;
        M[a][1] = Mc0 * e0.x + Mc1 * e0.y + Mc2 * e0.z;
        M[a][2] = Mc0 * z0.x + Mc1 * z0.y + Mc2 * z0.z;
    }
    germano_G(c0, e0, z0, G);
    double num = 0, den = 0;
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) {
            const double P = M[a][0] * G[b][0] + M[a][1] * G[b][1] + M[a][2] * G[b][2];
            num += L[b][a] * P;
        }
    for (int n = 0; n < 3; n++)
        for (int m = 0; m < 3; m++) {
            const double v = G[m][0] * M[n][0] + G[m][1] * M[n][1] + G[m][2] * M[n][2];
            den += M[n][m] * v;
        }
    num_out = num;
    den_out = den;
}

// An index the compiler must treat as unknown: kernels that march over planes form each plane's
// element index from a per-lane base and a uniform offset and pass it through here once per plane,
// so that the per-array addresses are recomputed from (kernel-argument base + index) at each use
// instead of being strength-reduced into loop-carried 64-bit pointers, one register pair per array.
__device__ __forceinline__ int opaque_index(int x) {
    __asm__ volatile("" : "+v"(x));
    return x;
}

// Neighbour values across the lanes of a 64-wide wave: shup gives lane l the value of lane l-1,
// shdn that of lane l+1.  Lane 0 (shup) and lane 63 (shdn) have no neighbour and receive 0; the
// tiled kernels use both lanes as halo lanes whose own results are discarded.  The DPP wavefront
// shifts (wave_shr:1, wave_shl:1) are VALU register moves; __shfl_up/down would go through the
// LDS crossbar (ds_bpermute) with its latency.
__device__ __forceinline__ int dpp_shr(int x) { return __builtin_amdgcn_mov_dpp(x, 0x138, 0xf, 0xf, true); }
__device__ __forceinline__ int dpp_shl(int x) { return __builtin_amdgcn_mov_dpp(x, 0x130, 0xf, 0xf, true); }
__device__ __forceinline__ double shup(double x) {
    return __hiloint2double(dpp_shr(__double2hiint(x)), dpp_shr(__double2loint(x)));
}
__device__ __forceinline__ double shdn(double x) {
    return __hiloint2double(dpp_shl(__double2hiint(x)), dpp_shl(__double2loint(x)));
}
__device__ __forceinline__ float shup(float x) { return __int_as_float(dpp_shr(__float_as_int(x))); }
__device__ __forceinline__ float shdn(float x) { return __int_as_float(dpp_shl(__float_as_int(x))); }
template <class T> __device__ __forceinline__ Vec3<T> shup3(Vec3<T> a) { return Vec3<T>{shup(a.x), shup(a.y), shup(a.z)}; }
template <class T> __device__ __forceinline__ Vec3<T> shdn3(Vec3<T> a) { return Vec3<T>{shdn(a.x), shdn(a.y), shdn(a.z)}; }

// Sum of v over the 64 lanes of a wave by a butterfly of cross-lane exchanges; every lane
// receives the same sum (each step adds the same two operands in every lane pair).
__device__ __forceinline__ double wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// The filter widths Delta = (1/aj)^(1/3) and Delta~ = (8 sum)^(1/3) of les.c:657-660, 2510 are
// written pow(x, 1./3.) in the reference.  The fused kernels take cbrt: the two differ by a few
// ulp (pow's exponent 1./3. is itself rounded), and on gfx950 the FP64 pow expands to about six
// times the instructions of cbrt.
__device__ __forceinline__ double cube_root(double x) { return cbrt(x); }

__device__ __forceinline__ double pmax(double a, double b) { return (a < b) ? b : a; }
__device__ __forceinline__ double pmin(double a, double b) { return (a < b) ? a : b; }

// MI355X dispatches the workgroups of a launch to its 8 XCDs round-robin (linear id % 8), and each
// XCD has its own L2.  xcd_tile maps the linear workgroup id to a tile id such that every XCD
// receives one contiguous range of tiles: neighbouring tiles, which read each other's halos, then
// run on the same XCD at about the same time and share those halos in its L2.
#ifndef TOSCA_XCD_TILES
#define TOSCA_XCD_TILES 1
#endif
constexpr int NUM_XCD = 8;
__device__ __forceinline__ int xcd_tile() {
    const int n = gridDim.x * gridDim.y * gridDim.z;
    const int w = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
    if (!TOSCA_XCD_TILES) return w;
    const int q = n / NUM_XCD, r = n % NUM_XCD, x = w % NUM_XCD, slot = w / NUM_XCD;
    return (x < r) ? x * (q + 1) + slot : r * (q + 1) + (x - r) * q + slot;
}

}  // namespace tosca
