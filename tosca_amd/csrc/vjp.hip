// vjp.hip -- hand-written adjoint (vector-Jacobian product) of the momentum right-hand side.
//
// Function (no counterpart in the reference; it is the adjoint of the reference's forward path):
//   R(ucont, ucat, theta, Cs) = FormU(ucont, ucat, nu_t; scale, formMode 0) + Buoyancy(theta; b)
//   nu_t = UpdateNut(ucat, Cs): Cs Delta^2 |S(ucat)| at interior cells, UpdateNutBCs ghosts
// with every input a full DMDA local array (ghost cells are inputs too) and R on the owned faces.
// Given the cotangent Rbar, the kernels accumulate ucont_bar, ucat_bar, theta_bar, Cs_bar, the
// transpose of each forward stage in reverse order:
//   k_vjp_fp      face dot (ueqn_solvers.c:1006-1035) -> fpbar per interior cell (a gather: the
//                 cell's own faces, the faces of its lower neighbours, and the ghost m-1 aliases:
//                 cell 1 on a periodic axis, the copied cell m-2 otherwise)
//   k_vjp_faces   (the i, j and k faces of each index point) cell divergence (ueqn_solvers.c:871-993) -> face-flux cotangent, then the
//                 convective flux (central / QUICK, inline.h:4314-4324, 5089-5109) -> ucont_bar and
//                 a packet of ucat cotangents, the viscous flux (ueqn_solvers.c:184-216...) -> the
//                 cotangents of its nine derivatives (Compute_du_i/j/k, inline.h:2748-2931) into the
//                 same packet, and the nu_eff cotangent.  Only the boundary faces scatter (atomics).
//   k_vjp_nut     nu_t = Cs Delta^2 |S| (les.c:2510-2514), with the UpdateNutBCs ghosts
//                 (boundary.c:2812-2917) transposed -> Cs_bar, and a packet for the |S| path
//                 (Compute_du_dxyz / Compute_du_center, inline.h:3344-3418, 3635-3663)
//   k_vjp_gather  per owned point: ucat_bar from the packets of the faces and cells whose regular
//                 stencils reach it, and theta_bar from Buoyancy (ueqn_sources.c:2812-2870)
// The forward intermediates the adjoint needs (nu_t, the face gradients, the interpolated face
// velocities) are recomputed from the inputs instead of being stored by the forward pass.
#include "fused.h"

namespace tosca {

namespace {

__device__ __forceinline__ void aadd(double *p, double v) { unsafeAtomicAdd(p, v); }
__device__ __forceinline__ void aadd(float *p, double v) { unsafeAtomicAdd(p, float(v)); }
template <class T>
__device__ __forceinline__ void aadd3(T *p, int c, V3 v) {
    aadd(p + 3 * c, v.x);
    aadd(p + 3 * c + 1, v.y);
    aadd(p + 3 * c + 2, v.z);
}
__device__ __forceinline__ V3 sc3(double s, V3 v) { return V3{s * v.x, s * v.y, s * v.z}; }
__device__ __forceinline__ V3 add3v(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ double dot3(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

// w * (p0 + p1 - m0 - m1): adjoint of the four-point differences of Compute_du_*
template <class T>
__device__ __forceinline__ void scat4(T *ub, int p0, int p1, int m0, int m1, double w, V3 gb) {
    const V3 a = sc3(w, gb), b = sc3(-w, gb);
    aadd3(ub, p0, a); aadd3(ub, p1, a); aadd3(ub, m0, b); aadd3(ub, m1, b);
}
template <class T>
__device__ __forceinline__ void scat2(T *ub, int p, int m, double w, V3 gb) {
    aadd3(ub, p, sc3(w, gb)); aadd3(ub, m, sc3(-w, gb));
}

// Cotangents of the nine curvilinear derivatives, grouped by derivative direction: gc = (dudc,
// dvdc, dwdc), ge = (dude, ...), gz = (dudz, ...).  Adjoint of Compute_du_i (same branches).
template <class T>
__device__ void du_i_adj(const Geo &g, T *ub, int i, int j, int k, V3 gc, V3 ge, V3 gz) {
    const bool iN = g.nonper(0), jN = g.nonper(1), kN = g.nonper(2);
    scat2(ub, g.L(k, j, i + 1), g.L(k, j, i), 1.0, gc);
    const bool ib = (i == 0 || i == g.mx - 2);
    if (j == g.my - 2 && ib && (jN || iN)) scat4(ub, g.L(k, j, i + 1), g.L(k, j, i), g.L(k, j - 1, i + 1), g.L(k, j - 1, i), 0.5, ge);
    else if (j == 1 && ib && (jN || iN)) scat4(ub, g.L(k, j + 1, i + 1), g.L(k, j + 1, i), g.L(k, j, i + 1), g.L(k, j, i), 0.5, ge);
    else scat4(ub, g.L(k, j + 1, i + 1), g.L(k, j + 1, i), g.L(k, j - 1, i + 1), g.L(k, j - 1, i), 0.25, ge);
    if (k == g.mz - 2 && ib && (kN || iN)) scat4(ub, g.L(k, j, i + 1), g.L(k, j, i), g.L(k - 1, j, i + 1), g.L(k - 1, j, i), 0.5, gz);
    else if (k == 1 && ib && (kN || iN)) scat4(ub, g.L(k + 1, j, i + 1), g.L(k + 1, j, i), g.L(k, j, i + 1), g.L(k, j, i), 0.5, gz);
    else scat4(ub, g.L(k + 1, j, i + 1), g.L(k + 1, j, i), g.L(k - 1, j, i + 1), g.L(k - 1, j, i), 0.25, gz);
}

// adjoint of Compute_du_j
template <class T>
__device__ void du_j_adj(const Geo &g, T *ub, int i, int j, int k, V3 gc, V3 ge, V3 gz) {
    const bool iN = g.nonper(0), jN = g.nonper(1), kN = g.nonper(2);
    const bool jb = (j == 0 || j == g.my - 2);
    if (i == g.mx - 2 && jb && (iN || jN)) scat4(ub, g.L(k, j + 1, i), g.L(k, j, i), g.L(k, j + 1, i - 1), g.L(k, j, i - 1), 0.5, gc);
    else if (i == 1 && jb && (iN || jN)) scat4(ub, g.L(k, j + 1, i + 1), g.L(k, j, i + 1), g.L(k, j + 1, i), g.L(k, j, i), 0.5, gc);
    else scat4(ub, g.L(k, j + 1, i + 1), g.L(k, j, i + 1), g.L(k, j + 1, i - 1), g.L(k, j, i - 1), 0.25, gc);
    scat2(ub, g.L(k, j + 1, i), g.L(k, j, i), 1.0, ge);
    if (k == g.mz - 2 && jb && (kN || jN)) scat4(ub, g.L(k, j + 1, i), g.L(k, j, i), g.L(k - 1, j + 1, i), g.L(k - 1, j, i), 0.5, gz);
    else if (k == 1 && jb && (kN || jN)) scat4(ub, g.L(k + 1, j + 1, i), g.L(k + 1, j, i), g.L(k, j + 1, i), g.L(k, j, i), 0.5, gz);
    else scat4(ub, g.L(k + 1, j + 1, i), g.L(k + 1, j, i), g.L(k - 1, j + 1, i), g.L(k - 1, j, i), 0.25, gz);
}

// adjoint of Compute_du_k
template <class T>
__device__ void du_k_adj(const Geo &g, T *ub, int i, int j, int k, V3 gc, V3 ge, V3 gz) {
    const bool iN = g.nonper(0), jN = g.nonper(1), kN = g.nonper(2);
    const bool kb = (k == 0 || k == g.mz - 2);
    if (i == g.mx - 2 && kb && (iN || kN)) scat4(ub, g.L(k + 1, j, i), g.L(k, j, i), g.L(k + 1, j, i - 1), g.L(k, j, i - 1), 0.5, gc);
    else if (i == 1 && kb && (iN || kN)) scat4(ub, g.L(k + 1, j, i + 1), g.L(k, j, i + 1), g.L(k + 1, j, i), g.L(k, j, i), 0.5, gc);
    else scat4(ub, g.L(k + 1, j, i + 1), g.L(k, j, i + 1), g.L(k + 1, j, i - 1), g.L(k, j, i - 1), 0.25, gc);
    if (j == g.my - 2 && kb && (jN || kN)) scat4(ub, g.L(k + 1, j, i), g.L(k, j, i), g.L(k + 1, j - 1, i), g.L(k, j - 1, i), 0.5, ge);
    else if (j == 1 && kb && (jN || kN)) scat4(ub, g.L(k + 1, j + 1, i), g.L(k, j + 1, i), g.L(k + 1, j, i), g.L(k, j, i), 0.5, ge);
    else scat4(ub, g.L(k + 1, j + 1, i), g.L(k, j + 1, i), g.L(k + 1, j - 1, i), g.L(k, j - 1, i), 0.25, ge);
    scat2(ub, g.L(k + 1, j, i), g.L(k, j, i), 1.0, gz);
}

// adjoint of Compute_du_center at cell (i, j, k)
template <class T>
__device__ void du_center_adj(const Geo &g, T *ub, int i, int j, int k, V3 gc, V3 ge, V3 gz) {
    const int c = g.L(k, j, i);
    if (g.nonper(0) && i == g.mx - 2) scat2(ub, c, g.L(k, j, i - 1), 1.0, gc);
    else if (g.nonper(0) && i == 1) scat2(ub, g.L(k, j, i + 1), c, 1.0, gc);
    else scat2(ub, g.L(k, j, i + 1), g.L(k, j, i - 1), 0.5, gc);
    if (g.nonper(1) && j == g.my - 2) scat2(ub, c, g.L(k, j - 1, i), 1.0, ge);
    else if (g.nonper(1) && j == 1) scat2(ub, g.L(k, j + 1, i), c, 1.0, ge);
    else scat2(ub, g.L(k, j + 1, i), g.L(k, j - 1, i), 0.5, ge);
    if (g.nonper(2) && k == g.mz - 2) scat2(ub, c, g.L(k - 1, j, i), 1.0, gz);
    else if (g.nonper(2) && k == 1) scat2(ub, g.L(k + 1, j, i), c, 1.0, gz);
    else scat2(ub, g.L(k + 1, j, i), g.L(k - 1, j, i), 0.5, gz);
}

struct NutBCv { int bc[6]; };

// T: type of the flow fields (ucont, ucat, theta, nu_t) and of their cotangents; P: storage type
// of the cotangent packets and intermediates (fpbar, nu_eff_bar, the face and cell packets).  The
// cotangent arithmetic is FP64 in both forms.
template <class T, class P>
struct VjpArgs {
    const T *ucont, *ucat, *theta, *nut;
    T *nut_w;           // where k_vjp_pre writes the recomputed nu_t (the array nut reads)
    const double *cs, *rbar;
    const double *csi, *eta, *zet, *icsi, *ieta, *izet, *jcsi, *jeta, *jzet, *kcsi, *keta, *kzet;
    const double *aj, *iaj, *jaj, *kaj;
    const double *tauI, *tauJ;
    P *fpbar;
    T *vbar, *ubar, *tbar;
    double *csbar;
    double *nutbar_acc; // nu_t given as an input: its cotangent is accumulated here (local)
    P *pk[3];           // per face direction: 12 SoA planes of n values (cv, gn, ga, gb)
    P *nueb[3];         // per face direction: cotangent of nu_eff at each face (n values)
    P *spk;             // per interior cell: 9 SoA planes (cotangents of the du_center derivatives)
    long n;
    double scale, bscale, tRef, nu;
    int quick, les;
    int nut_in;         // nu_t is an input array (no |S| path, no Cs)
    int bc_u[6], bc_nut[6];
};

template <class A_>
__device__ __forceinline__ V3 cellS(const A_ &A, int q, int c) {
    return ld3(q == 0 ? A.csi : (q == 1 ? A.eta : A.zet), c);
}
template <class A_>
__device__ __forceinline__ double faceA(const A_ &A, int q, int c) {
    return (q == 0 ? A.iaj : (q == 1 ? A.jaj : A.kaj))[c];
}
template <class T>
__device__ __forceinline__ Grad grad64(const GradT<T> &a) {
    return Grad{double(a.dudc), double(a.dvdc), double(a.dwdc), double(a.dude), double(a.dvde), double(a.dwde),
                double(a.dudz), double(a.dvdz), double(a.dwdz)};
}
template <class P>
__device__ __forceinline__ void stp3(P *p, int c, V3 v) { p[3 * c] = P(v.x); p[3 * c + 1] = P(v.y); p[3 * c + 2] = P(v.z); }
template <class P>
__device__ __forceinline__ V3 ldp3(const P *p, int c) { return V3{double(p[3 * c]), double(p[3 * c + 1]), double(p[3 * c + 2])}; }

template <class T>
__global__ void k_vjp_zero(T *a, long n) {
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) a[t] = T(0);
}

}  // namespace

// fpbar of interior cell (i, j, k) (a gather; no atomics); S: its csi, eta, zet
template <class T, class P>
__device__ __forceinline__ void fp_cell(const VjpArgs<T, P> &A, const Geo &g, int i, int j, int k, const V3 *S) {
    const int idx[3] = {i, j, k}, m[3] = {g.mx, g.my, g.mz};
    const int c = g.L(k, j, i);
    V3 fb{0, 0, 0};
    for (int q = 0; q < 3; q++) {
        const int dq[3] = {q == 0, q == 1, q == 2};
        const V3 Sc = S[q];
        // face q of this cell (this cell on the left)
        double w = A.scale * faceA(A, q, c) * 0.5 * A.rbar[3 * g.G(k, j, i) + q];
        fb = add3v(fb, sc3(w, Sc));
        // face q of the lower neighbour (this cell on the right)
        if (idx[q] >= 2) {
            const int kk = k - dq[2], jj = j - dq[1], ii = i - dq[0];
            const double w2 = A.scale * faceA(A, q, g.L(kk, jj, ii)) * 0.5 * A.rbar[3 * g.G(kk, jj, ii) + q];
            fb = add3v(fb, sc3(w2, Sc));
        }
        // the ghost cell m-1 to the right of face m-2 carries fp of cell 1 (periodic) or of m-2
        if (g.per[q] != TOSCA_PERIODIC_NONE && idx[q] == 1) {
            int t[3] = {i, j, k};
            t[q] = m[q] - 2;
            const int cf = g.L(t[2], t[1], t[0]);
            t[q] = m[q] - 1;
            const V3 Sg = cellS(A, q, g.L(t[2], t[1], t[0]));
            t[q] = m[q] - 2;
            const double w3 = A.scale * faceA(A, q, cf) * 0.5 * A.rbar[3 * g.G(t[2], t[1], t[0]) + q];
            fb = add3v(fb, sc3(w3, Sg));
        }
        if (g.per[q] == TOSCA_PERIODIC_NONE && idx[q] == m[q] - 2) {
            int t[3] = {i, j, k};
            t[q] = m[q] - 1;
            const V3 Sg = cellS(A, q, g.L(t[2], t[1], t[0]));
            fb = add3v(fb, sc3(w, Sg));
        }
    }
    stp3(A.fpbar, c, fb);
}

// fpbar at every interior cell
template <class T, class P>
__global__ void __launch_bounds__(256) k_vjp_fp(VjpArgs<T, P> A, Geo g) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1, j = blockIdx.y * blockDim.y + threadIdx.y + 1,
              k = blockIdx.z + 1;
    if (i > g.mx - 2 || j > g.my - 2 || k > g.mz - 2) return;
    const int c = g.L(k, j, i);
    const V3 S[3] = {ld3(A.csi, c), ld3(A.eta, c), ld3(A.zet, c)};
    fp_cell(A, g, i, j, k, S);
}

// fpbar and the recomputed nu_t = Cs Delta^2 |S| (UpdateNut's interior, as k_nut_fused forms it,
// les.c:2510-2514) of every interior cell, from one load of its csi, eta, zet
template <class T, class P>
__global__ void __launch_bounds__(256) k_vjp_pre(VjpArgs<T, P> A, Geo g) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1, j = blockIdx.y * blockDim.y + threadIdx.y + 1,
              k = blockIdx.z + 1;
    if (i > g.mx - 2 || j > g.my - 2 || k > g.mz - 2) return;
    const int c = g.L(k, j, i);
    const V3 S[3] = {ld3(A.csi, c), ld3(A.eta, c), ld3(A.zet, c)};
    const double a = A.aj[c];
    const double Sm = strain_mag(du_dxyz(cvt3<T>(S[0]), cvt3<T>(S[1]), cvt3<T>(S[2]), T(a), du_center(g, A.ucat, i, j, k)));
    const double filter = cube_root(1. / a);
    A.nut_w[c] = T(A.cs[c] * (filter * filter) * Sm);
    fp_cell(A, g, i, j, k, S);
}

// ------------------------------------------------------------------------------------------------
// Faces: the flux cotangent, ucont_bar, the nu_eff cotangent, and the ucat cotangent packets.
// Interior faces (1 <= f <= m-3 along the face direction) store a packet -- the convective
// cotangent cv = -V Fbar and the cotangents of the normal and two transverse derivatives -- that
// k_vjp_gather later folds into ucat_bar with the regular stencils; the two boundary faces per line
// (the only ones with the one-sided and clipped stencils) scatter into ucat_bar directly.
// Thread x runs along i for every face direction (coalesced).
// ------------------------------------------------------------------------------------------------
// one face of direction DIR: x = (i, j, k) its index point, f = x[DIR] in 0 .. m-2, the other two
// in 1 .. m-2
template <int DIR, class T, class P>
__device__ __forceinline__ void face_adj(const VjpArgs<T, P> &A, const Geo &g, const int *x) {
    const int m[3] = {g.mx, g.my, g.mz};
    const int f = x[DIR];
    const bool per = g.per[DIR] != TOSCA_PERIODIC_NONE;
    if (per && f == 0) return;                     // face 0 of a periodic axis is face m-2
    const int i = x[0], j = x[1], k = x[2];
    const int c = g.L(k, j, i);
    const int up = (DIR == 0) ? g.L(k, j, i + 1) : (DIR == 1) ? g.L(k, j + 1, i) : g.L(k + 1, j, i);
    const bool interior = f >= 1 && f <= m[DIR] - 3;
    // cotangent of the face flux: + fp of the left cell, - fp of the right cell (interior cells)
    V3 Fb{0, 0, 0};
    if (f >= 1) Fb = add3v(Fb, ldp3(A.fpbar, c));
    if (f + 1 <= m[DIR] - 2) Fb = add3v(Fb, sc3(-1.0, ldp3(A.fpbar, up)));
    if (per && f == m[DIR] - 2) {                  // face 0 aliases this face: it is cell 1's left face
        int t[3] = {i, j, k};
        t[DIR] = 1;
        Fb = add3v(Fb, sc3(-1.0, ldp3(A.fpbar, g.L(t[2], t[1], t[0]))));
    }
    const T *u = A.ucat;
    // ---- convective: div = -V * u_face ---------------------------------------------------------
    const T VT = A.ucont[3 * c + DIR];
    const double V = double(VT);
    const V3 cv = sc3(-V, Fb);
    if (A.quick) {
        int L, R;
        face2cell4(g.per[DIR], f, m[DIR], L, R);
        int xl[3] = {i, j, k}, xr[3] = {i, j, k};
        xl[DIR] = L; xr[DIR] = R;
        const int cLL = g.L(xl[2], xl[1], xl[0]), cRR = g.L(xr[2], xr[1], xr[0]);
        const V3 uf = cvt3<double>(quick(ld3(u, cLL), ld3(u, c), ld3(u, up), ld3(u, cRR), VT));
        A.vbar[3 * c + DIR] = T(-dot3(Fb, uf));
        if (!interior) {
            // d/dfU = 1/2 + 2/8, d/dfD = 1/2 - 1/8, d/dfUU = -1/8
            if (V > 0) { aadd3(A.ubar, c, sc3(0.75, cv)); aadd3(A.ubar, up, sc3(0.375, cv)); aadd3(A.ubar, cLL, sc3(-0.125, cv)); }
            else { aadd3(A.ubar, up, sc3(0.75, cv)); aadd3(A.ubar, c, sc3(0.375, cv)); aadd3(A.ubar, cRR, sc3(-0.125, cv)); }
        }
    } else {
        const Vec3<T> a0 = ld3(u, c), a1 = ld3(u, up);
        const V3 uf{double(T(0.5) * (a0.x + a1.x)), double(T(0.5) * (a0.y + a1.y)), double(T(0.5) * (a0.z + a1.z))};
        A.vbar[3 * c + DIR] = T(-dot3(Fb, uf));
        if (!interior) { aadd3(A.ubar, c, sc3(0.5, cv)); aadd3(A.ubar, up, sc3(0.5, cv)); }
    }
    // ---- viscous ------------------------------------------------------------------------------------
    double nuEff = A.nu;
    bool dep = false;                               // nu_eff depends on nu_t
    if (A.les) {
        const int bl = A.bc_u[2 * DIR], br = A.bc_u[2 * DIR + 1];
        if (DIR < 2 && ((bl == TOSCA_BCU_WALLFUNCTION && f == 0) || (br == TOSCA_BCU_WALLFUNCTION && f == m[DIR] - 2))) {
            nuEff = A.nu;
        } else if ((bl == TOSCA_BCU_SLIP && f == 0) || (br == TOSCA_BCU_SLIP && f == m[DIR] - 2)) {
            nuEff = 0.0;
        } else {
            nuEff = A.nu + 0.5 * (double(A.nut[c]) + double(A.nut[up]));
            dep = true;
        }
    }
    const double *fcp = (DIR == 0) ? A.icsi : (DIR == 1) ? A.jcsi : A.kcsi;
    const double *fep = (DIR == 0) ? A.ieta : (DIR == 1) ? A.jeta : A.keta;
    const double *fzp = (DIR == 0) ? A.izet : (DIR == 1) ? A.jzet : A.kzet;
    const double *fap = (DIR == 0) ? A.iaj : (DIR == 1) ? A.jaj : A.kaj;
    const V3 fc = ld3(fcp, c), fe = ld3(fep, c), fz = ld3(fzp, c);
    const double ajc = fap[c];
    double nb = 0.0;
    if (dep) {
        const Grad gr = grad64((DIR == 0) ? du_i(g, u, i, j, k) : (DIR == 1) ? du_j(g, u, i, j, k) : du_k(g, u, i, j, k));
        nb = dot3(Fb, visc_flux(gr, fc, fe, fz, DIR, ajc, 1.0));
    }
    A.nueb[DIR][c] = P(nb);
    // visc_m = a nu (g_X1 d_X u_m + (M_X . d_X u) n_m) summed over X = csi, eta, zet
    const V3 nrm = (DIR == 0) ? fc : (DIR == 1) ? fe : fz;
    const double g11 = dot3(fc, nrm), g21 = dot3(fe, nrm), g31 = dot3(fz, nrm);
    const V3 Fa = sc3(ajc * nuEff, Fb);
    auto gx = [&](double gX1, V3 MX) {
        const double d = dot3(MX, Fa);
        return V3{gX1 * Fa.x + d * nrm.x, gX1 * Fa.y + d * nrm.y, gX1 * Fa.z + d * nrm.z};
    };
    const V3 gc = gx(g11, fc), ge = gx(g21, fe), gz = gx(g31, fz);
    if (interior) {
        // packet: cv, normal, transverse a, transverse b (a/b: the lower / higher other axis)
        const V3 gn = (DIR == 0) ? gc : (DIR == 1) ? ge : gz;
        const V3 ga = (DIR == 0) ? ge : gc;
        const V3 gb = (DIR == 2) ? ge : gz;
        const V3 pv[4] = {cv, gn, ga, gb};
        P *pk = A.pk[DIR];
        for (int w = 0; w < 4; w++) {
            pk[(3 * w + 0) * A.n + c] = P(pv[w].x);
            pk[(3 * w + 1) * A.n + c] = P(pv[w].y);
            pk[(3 * w + 2) * A.n + c] = P(pv[w].z);
        }
    } else if (nuEff != 0.0) {
        if (DIR == 0) du_i_adj(g, A.ubar, i, j, k, gc, ge, gz);
        else if (DIR == 1) du_j_adj(g, A.ubar, i, j, k, gc, ge, gz);
        else du_k_adj(g, A.ubar, i, j, k, gc, ge, gz);
    }
}

// the i, j and k faces of index point (i, j, k), i in 0 .. mx-2, j in 0 .. my-2, k in 0 .. mz-2
// (each face where it exists): the three share the point's fpbar, ucont, nu_t and ucat loads
template <class T, class P>
__global__ void __launch_bounds__(256) k_vjp_faces(VjpArgs<T, P> A, Geo g) {
    const int x[3] = {(int)(blockIdx.x * blockDim.x + threadIdx.x), (int)blockIdx.y, (int)blockIdx.z};
    if (x[0] > g.mx - 2) return;
    const bool in0 = x[0] >= 1, in1 = x[1] >= 1, in2 = x[2] >= 1;
    if (in1 && in2) face_adj<0>(A, g, x);
    if (in0 && in2) face_adj<1>(A, g, x);
    if (in0 && in1) face_adj<2>(A, g, x);
}

// ------------------------------------------------------------------------------------------------
// Interior cells: nu_t_bar (gathered from the nu_eff cotangents of the cell's six faces, plus the
// ghosts UpdateNutBCs fills from this cell), Cs_bar, and the packet of the |S| path.
// ------------------------------------------------------------------------------------------------
template <class T, class P>
__global__ void __launch_bounds__(256) k_vjp_nut(VjpArgs<T, P> A, Geo g) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1, j = blockIdx.y * blockDim.y + threadIdx.y + 1,
              k = blockIdx.z + 1;
    if (i > g.mx - 2 || j > g.my - 2 || k > g.mz - 2) return;
    const int c = g.L(k, j, i);
    const int x[3] = {i, j, k}, m[3] = {g.mx, g.my, g.mz};
    double nb = 0.0;
    for (int q = 0; q < 3; q++) {
        int t[3] = {i, j, k};
        nb += 0.5 * double(A.nueb[q][c]);                              // face q of the cell
        t[q] = x[q] - 1;
        if (!(x[q] == 1 && g.per[q] != TOSCA_PERIODIC_NONE))          // periodic face 0 is not evaluated
            nb += 0.5 * double(A.nueb[q][g.L(t[2], t[1], t[0])]);
        // ghosts filled from this cell: cotangent of nu_t there is 1/2 nu_eff_bar of the boundary face
        const int bl = A.bc_nut[2 * q], br = A.bc_nut[2 * q + 1];
        const bool pq = g.per[q] != TOSCA_PERIODIC_NONE;
        t[q] = 0;                                                      // face 0 (its left cell is ghost 0)
        const int f0 = g.L(t[2], t[1], t[0]);
        t[q] = m[q] - 2;                                               // face m-2 (right cell ghost m-1)
        const int fm = g.L(t[2], t[1], t[0]);
        if (bl == TOSCA_BCNUT_ZEROGRADIENT && x[q] == 1) nb += 0.5 * double(A.nueb[q][f0]);
        if (bl == TOSCA_BCNUT_PERIODIC && pq && x[q] == m[q] - 2) nb += 0.5 * double(A.nueb[q][f0]);
        if (br == TOSCA_BCNUT_ZEROGRADIENT && x[q] == m[q] - 2) nb += 0.5 * double(A.nueb[q][fm]);
        if (br == TOSCA_BCNUT_PERIODIC && pq && x[q] == 1) nb += 0.5 * double(A.nueb[q][fm]);
    }
    const V3 cc = ld3(A.csi, c), ee = ld3(A.eta, c), zz = ld3(A.zet, c);
    const double a = A.aj[c];
    // the strain as the forward computes it (in T), the rest in FP64
    const CartT<T> dT = du_dxyz(cvt3<T>(cc), cvt3<T>(ee), cvt3<T>(zz), T(a), du_center(g, A.ucat, i, j, k));
    const Cart d{double(dT.du_dx), double(dT.du_dy), double(dT.du_dz), double(dT.dv_dx), double(dT.dv_dy),
                 double(dT.dv_dz), double(dT.dw_dx), double(dT.dw_dy), double(dT.dw_dz)};
    const double S = double(strain_mag(dT));
    const double filter = cube_root(1. / a), d2 = filter * filter;
    A.csbar[c] = nb * d2 * S;
    V3 gc{0, 0, 0}, ge{0, 0, 0}, gz{0, 0, 0};
    if (S != 0.0 && nb != 0.0) {                   // |S| is not differentiable at 0: subgradient 0
        // d|S| / dD_mn = 2 S_mn / |S|, S = sym(D)
        const double sb = nb * A.cs[c] * d2 * 2.0 / S;
        const double Sxx = d.du_dx, Sxy = 0.5 * (d.du_dy + d.dv_dx), Sxz = 0.5 * (d.du_dz + d.dw_dx);
        const double Syy = d.dv_dy, Syz = 0.5 * (d.dv_dz + d.dw_dy), Szz = d.dw_dz;
        const V3 Du{sb * Sxx, sb * Sxy, sb * Sxz}, Dv{sb * Sxy, sb * Syy, sb * Syz}, Dw{sb * Sxz, sb * Syz, sb * Szz};
        // D_m. = aj (d_c u_m csi + d_e u_m eta + d_z u_m zet)  ->  d_X u_m bar = aj (Dbar_m . M_X)
        gc = V3{a * dot3(Du, cc), a * dot3(Dv, cc), a * dot3(Dw, cc)};
        ge = V3{a * dot3(Du, ee), a * dot3(Dv, ee), a * dot3(Dw, ee)};
        gz = V3{a * dot3(Du, zz), a * dot3(Dv, zz), a * dot3(Dw, zz)};
    }
    const V3 pv[3] = {gc, ge, gz};
    for (int w = 0; w < 3; w++) {
        A.spk[(3 * w + 0) * A.n + c] = P(pv[w].x);
        A.spk[(3 * w + 1) * A.n + c] = P(pv[w].y);
        A.spk[(3 * w + 2) * A.n + c] = P(pv[w].z);
    }
}

namespace {
template <class P>
__device__ __forceinline__ V3 pk3(const P *pk, long n, int w, int c) {
    return V3{double(pk[(3 * w + 0) * n + c]), double(pk[(3 * w + 1) * n + c]), double(pk[(3 * w + 2) * n + c])};
}
}  // namespace

// ------------------------------------------------------------------------------------------------
// ucat_bar and theta_bar of every owned point (boundary ghosts included): the regular stencils of
// the interior faces and interior cells, transposed.  Adds to the boundary-face scatter.
// ------------------------------------------------------------------------------------------------
template <class T, class P>
__global__ void __launch_bounds__(256) k_vjp_gather(VjpArgs<T, P> A, Geo g) {
    // a workgroup is 64 points along i on 4 consecutive k planes (which read each other's k
    // neighbours' packets); consecutive workgroups on one XCD take consecutive j rows
    const int tile = xcd_tile();
    const int bx = tile % gridDim.x, j = (tile / gridDim.x) % g.my, bk = tile / (gridDim.x * g.my);
    const int i = bx * 64 + threadIdx.x, k = bk * 4 + threadIdx.y;
    if (i > g.mx - 1 || k > g.mz - 1) return;
    const int x[3] = {i, j, k}, m[3] = {g.mx, g.my, g.mz};
    const int c = g.L(k, j, i);
    auto Lof = [&](const int *t) { return g.L(t[2], t[1], t[0]); };
    V3 acc = cvt3<double>(ld3(A.ubar, c));
    double tb = 0.0;
    for (int q = 0; q < 3; q++) {
        const int a = (q == 0) ? 1 : 0, b = (q == 2) ? 1 : 2;          // transverse axes, lower first
        const P *pk = A.pk[q];
        // an interior face of direction q at index point t
        auto faceOK = [&](const int *t) {
            return t[q] >= 1 && t[q] <= m[q] - 3 && t[a] >= 1 && t[a] <= m[a] - 2 && t[b] >= 1 && t[b] <= m[b] - 2;
        };
        int t[3];
        // normal derivative and convective interpolation: x is the left cell of face x, the right
        // cell of face x - q; QUICK also reaches x from faces x + q (V > 0) and x - 2q (V <= 0)
        for (int s = -2; s <= 1; s++) {
            t[0] = i; t[1] = j; t[2] = k;
            t[q] = x[q] + s;
            if (!faceOK(t)) continue;
            const int fc = Lof(t);
            const V3 cv = pk3(pk, A.n, 0, fc);
            if (s == 0 || s == -1) {
                const V3 gn = pk3(pk, A.n, 1, fc);
                acc = add3v(acc, sc3(s == 0 ? -1.0 : 1.0, gn));
            }
            double w;
            if (!A.quick) w = (s == 0 || s == -1) ? 0.5 : 0.0;
            else {
                const bool pos = A.ucont[3 * fc + q] > T(0);
                if (s == 0) w = pos ? 0.75 : 0.375;
                else if (s == -1) w = pos ? 0.375 : 0.75;
                else if (s == 1) w = pos ? -0.125 : 0.0;
                else w = pos ? 0.0 : -0.125;
            }
            if (w != 0.0) acc = add3v(acc, sc3(w, cv));
        }
        // transverse four-point averages: faces x - e and x - e - q add, x + e and x + e - q subtract
        for (int e = 0; e < 2; e++) {
            const int ax = e == 0 ? a : b;
            for (int sgn = -1; sgn <= 1; sgn += 2) {
                for (int sq = 0; sq <= 1; sq++) {
                    t[0] = i; t[1] = j; t[2] = k;
                    t[ax] = x[ax] + sgn;
                    t[q] = x[q] - sq;
                    if (!faceOK(t)) continue;
                    const V3 gt = pk3(pk, A.n, 2 + e, Lof(t));
                    acc = add3v(acc, sc3(-0.25 * sgn, gt));
                }
            }
        }
        // |S| path: du_center of the interior cells x - q, x, x + q along q
        if (A.les && !A.nut_in) {
            const bool np = g.per[q] == TOSCA_PERIODIC_NONE;
            auto cellOK = [&](const int *t) {
                return t[0] >= 1 && t[0] <= m[0] - 2 && t[1] >= 1 && t[1] <= m[1] - 2 && t[2] >= 1 && t[2] <= m[2] - 2;
            };
            for (int s = -1; s <= 1; s++) {
                t[0] = i; t[1] = j; t[2] = k;
                t[q] = x[q] + s;
                if (!cellOK(t)) continue;
                const int cq = t[q];
                double w = 0.0;
                if (s == 1) w = (np && cq == m[q] - 2) ? -1.0 : (np && cq == 1) ? 0.0 : -0.5;
                else if (s == -1) w = (np && cq == m[q] - 2) ? 0.0 : (np && cq == 1) ? 1.0 : 0.5;
                else w = (np && cq == m[q] - 2) ? 1.0 : (np && cq == 1) ? -1.0 : 0.0;
                if (w != 0.0) acc = add3v(acc, sc3(w, pk3(A.spk, A.n, q, Lof(t))));
            }
        }
        // buoyancy: the + face q of this cell and of the cell below along q
        if (!A.tbar) continue;
        for (int s = -1; s <= 0; s++) {
            t[0] = i; t[1] = j; t[2] = k;
            t[q] = x[q] + s;
            if (t[0] < 1 || t[0] > m[0] - 2 || t[1] < 1 || t[1] > m[1] - 2 || t[2] < 1 || t[2] > m[2] - 2) continue;
            const int cb = Lof(t);
            const double *face = (q == 0) ? A.icsi : (q == 1) ? A.jeta : A.kzet;
            tb += -A.bscale * (-9.81 * face[3 * cb + 2]) * 0.5 / A.tRef * A.rbar[3 * g.G(t[2], t[1], t[0]) + q];
        }
    }
    stp3(A.ubar, c, acc);
    if (A.tbar) A.tbar[c] = T(tb);
}

// ------------------------------------------------------------------------------------------------
// nu_t as an input array: its cotangent at every owned point, accumulated (+=).  The face of
// direction q at index point c averages nu_t of cells c and c + e_q, so point c collects half the
// nu_eff cotangent of faces c and c - e_q (zero where a face is not evaluated).
// ------------------------------------------------------------------------------------------------
template <class T, class P>
__global__ void __launch_bounds__(256) k_vjp_nutbar(VjpArgs<T, P> A, Geo g) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, k = blockIdx.z;
    if (i > g.mx - 1) return;
    const int c = g.L(k, j, i);
    const double s = double(A.nueb[0][c]) + double(A.nueb[0][g.L(k, j, i - 1)]) + double(A.nueb[1][c]) +
                     double(A.nueb[1][g.L(k, j - 1, i)]) + double(A.nueb[2][c]) + double(A.nueb[2][g.L(k - 1, j, i)]);
    A.nutbar_acc[c] += 0.5 * s;
}

// The transposed stages on the stream.  cs != nullptr: nu_t = UpdateNut(ucat, cs) is recomputed
// and differentiated (Cs_bar, |S| path); nut_in != nullptr: nu_t is an input and its cotangent is
// added to nut_bar.  theta_bar == nullptr: no Buoyancy term.  ws: n values of T for nu_t, then
// 51 n values of P for the intermediates (52 n doubles cover both forms).
template <class T, class P>
void vjp_rhs_impl(const Geo &g, const tosca_mesh *mesh, const tosca_params *p, const T *ucont, const T *ucat,
                  const double *cs, const T *nut_in, const double *rbar, double scale, double buoy_scale, T *ucont_bar,
                  T *ucat_bar, T *theta_bar, double *cs_bar, double *nut_bar, double *ws, hipStream_t s) {
    const long n = g.nloc();
    T *nut = reinterpret_cast<T *>(ws);
    P *base = reinterpret_cast<P *>(ws + n);
    P *fpbar = base, *nueb = base + 3 * n, *spk = base + 6 * n, *pk = base + 15 * n;   // 51 n values of P
    const dim3 zb(256), zg(2048);
    TOSCA_LAUNCH("k_vjp_zero", k_vjp_zero<T>, zg, zb, 0, s, ucont_bar, 3 * n);
    TOSCA_LAUNCH("k_vjp_zero", k_vjp_zero<T>, zg, zb, 0, s, ucat_bar, 3 * n);
    if (theta_bar) TOSCA_LAUNCH("k_vjp_zero", k_vjp_zero<T>, zg, zb, 0, s, theta_bar, n);
    if (cs_bar) launch_fill(cs_bar, n, 0.0, s);
    TOSCA_LAUNCH("k_vjp_zero", k_vjp_zero<P>, zg, zb, 0, s, nueb, 3 * n);
    const bool les = p->les_active != 0;
    const bool renut = les && !nut_in;      // nu_t recomputed as the forward step does (k_vjp_pre)
    VjpArgs<T, P> A;
    A.ucont = ucont; A.ucat = ucat; A.theta = nullptr; A.cs = cs; A.rbar = rbar;
    A.nut = les ? (nut_in ? nut_in : nut) : ucat;
    A.nut_w = nut;
    A.csi = mesh->csi; A.eta = mesh->eta; A.zet = mesh->zet;
    A.icsi = mesh->icsi; A.ieta = mesh->ieta; A.izet = mesh->izet;
    A.jcsi = mesh->jcsi; A.jeta = mesh->jeta; A.jzet = mesh->jzet;
    A.kcsi = mesh->kcsi; A.keta = mesh->keta; A.kzet = mesh->kzet;
    A.aj = mesh->aj; A.iaj = mesh->iaj; A.jaj = mesh->jaj; A.kaj = mesh->kaj;
    A.tauI = p->tau_wall_i; A.tauJ = p->tau_wall_j;
    A.fpbar = fpbar; A.nutbar_acc = nut_bar;
    A.vbar = ucont_bar; A.ubar = ucat_bar; A.tbar = theta_bar; A.csbar = cs_bar;
    for (int q = 0; q < 3; q++) { A.pk[q] = pk + 12 * q * n; A.nueb[q] = nueb + q * n; }
    A.spk = spk;
    A.n = n;
    A.scale = scale; A.bscale = buoy_scale; A.tRef = p->t_ref; A.nu = p->nu;
    A.quick = p->div_scheme == TOSCA_DIV_QUICK;
    A.les = les;
    A.nut_in = nut_in != nullptr;
    for (int q = 0; q < 6; q++) { A.bc_u[q] = p->bc_u[q]; A.bc_nut[q] = p->bc_nut[q]; }
    const bool f32 = sizeof(T) == 4;
    const dim3 cb(64, 4), cg((g.mx - 2 + 63) / 64, (g.my - 2 + 3) / 4, g.mz - 2);
    if (renut) {
        TOSCA_LAUNCH(f32 ? "k_vjp_pre_f32" : "k_vjp_pre", (k_vjp_pre<T, P>), cg, cb, 0, s, A, g);
        zero_outside_interior(nut, g, s);
        nut_bcs(nut, p, g, s);
    } else {
        TOSCA_LAUNCH(f32 ? "k_vjp_fp_f32" : "k_vjp_fp", (k_vjp_fp<T, P>), cg, cb, 0, s, A, g);
    }
    // faces: x along i (0 .. mx-2 for i faces, 1 .. mx-2 otherwise), y / z the other two axes
    TOSCA_LAUNCH(f32 ? "k_vjp_faces_f32" : "k_vjp_faces", (k_vjp_faces<T, P>), dim3((g.mx - 1 + 255) / 256, g.my - 1, g.mz - 1),
                 dim3(256), 0, s, A, g);
    if (les && !nut_in) TOSCA_LAUNCH(f32 ? "k_vjp_nut_f32" : "k_vjp_nut", (k_vjp_nut<T, P>), cg, cb, 0, s, A, g);
    if (les && nut_in)
        TOSCA_LAUNCH("k_vjp_nutbar", (k_vjp_nutbar<T, P>), dim3((g.mx + 255) / 256, g.my, g.mz), dim3(256), 0, s, A, g);
    TOSCA_LAUNCH(f32 ? "k_vjp_gather_f32" : "k_vjp_gather", (k_vjp_gather<T, P>), dim3((g.mx + 63) / 64, (g.mz + 3) / 4, g.my),
                 dim3(64, 4), 0, s, A, g);
}

void vjp_rhs(const Geo &g, const tosca_mesh *mesh, const tosca_params *p, const double *ucont, const double *ucat,
             const double *theta, const double *cs, const double *rbar, double scale, double buoy_scale,
             double *ucont_bar, double *ucat_bar, double *theta_bar, double *cs_bar, double *ws, hipStream_t s) {
    vjp_rhs_impl<double, double>(g, mesh, p, ucont, ucat, cs, nullptr, rbar, scale, buoy_scale, ucont_bar, ucat_bar,
                                 theta_bar, cs_bar, nullptr, ws, s);
}

void vjp_rhs_mixed(const Geo &g, const tosca_mesh *mesh, const tosca_params *p, const float *ucont, const float *ucat,
                   const double *cs, const double *rbar, double scale, double buoy_scale, float *ucont_bar,
                   float *ucat_bar, float *theta_bar, double *cs_bar, double *ws, hipStream_t s) {
    vjp_rhs_impl<float, float>(g, mesh, p, ucont, ucat, cs, nullptr, rbar, scale, buoy_scale, ucont_bar, ucat_bar,
                               theta_bar, cs_bar, nullptr, ws, s);
}

// FormU(formMode 0, scale) as a function of (ucont, ucat, nu_t): ucont_bar and ucat_bar
// overwritten, the nu_t cotangent added to nut_bar (when LES is active)
void vjp_formu_nut(const Geo &g, const tosca_mesh *mesh, const tosca_params *p, const double *ucont, const double *ucat,
                   const double *nut, const double *rbar, double scale, double *ucont_bar, double *ucat_bar,
                   double *nut_bar, double *ws, hipStream_t s) {
    vjp_rhs_impl<double, double>(g, mesh, p, ucont, ucat, nullptr, p->les_active ? nut : nullptr, rbar, scale, 0.0,
                                 ucont_bar, ucat_bar, nullptr, nullptr, nut_bar, ws, s);
}

}  // namespace tosca
