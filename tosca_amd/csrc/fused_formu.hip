// fused_formu.hip -- FormU (formMode 0/1/2) + Buoyancy in one k-marching kernel for gfx950.
//
// The reference evaluates FormU in three sweeps over HBM (face fluxes into six workspace vectors,
// their divergence into lFp, the face-dot into Rhs; src/solvers/ueqn_solvers.c:173-1035) plus a
// separate Buoyancy sweep (src/sources/ueqn_sources.c:2726-2880).  Here one workgroup owns a
// 62 x (NR-1) tile of (i, j) cells and marches along k; per plane it
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
#include <type_traits>

#include "fused.h"

namespace tosca {

namespace {

template <class T> struct FluxT { Vec3<T> d, v; };
using Flux = FluxT<double>;

template <class T> __device__ __forceinline__ Vec3<T> add3(Vec3<T> a, Vec3<T> b) { return Vec3<T>{a.x + b.x, a.y + b.y, a.z + b.z}; }

// convective (div) and viscous (visc) flux of one face.  dir: 0 = i face (k,j,i)|(k,j,i+1),
// 1 = j face, 2 = k face.  (i,j,k) is the face index after the periodic remap.
template <int DIR, class T, class MT>
__device__ __forceinline__ FluxT<T> face_flux(const FormUArgsT<T, MT> &A, const Geo &g, int i, int j, int k) {
    const int c = g.L(k, j, i);
    const T *u = A.ucat;
    FluxT<T> r;
    r.d = Vec3<T>{0, 0, 0};
    r.v = Vec3<T>{0, 0, 0};
    const int up = (DIR == 0) ? g.L(k, j, i + 1) : (DIR == 1) ? g.L(k, j + 1, i) : g.L(k + 1, j, i);
    if (A.mode != 2) {
        const T V = A.ucont[3 * c + DIR];
        Vec3<T> f;
        if (A.quick) {
            int L, R;
            if (DIR == 0) { face2cell4(g.per[0], i, g.mx, L, R); f = quick(ld3(u, g.L(k, j, L)), ld3(u, c), ld3(u, up), ld3(u, g.L(k, j, R)), V); }
            else if (DIR == 1) { face2cell4(g.per[1], j, g.my, L, R); f = quick(ld3(u, g.L(k, L, i)), ld3(u, c), ld3(u, up), ld3(u, g.L(k, R, i)), V); }
            else { face2cell4(g.per[2], k, g.mz, L, R); f = quick(ld3(u, g.L(L, j, i)), ld3(u, c), ld3(u, up), ld3(u, g.L(R, j, i)), V); }
        } else {
            Vec3<T> a = ld3(u, c), b = ld3(u, up);
            const T h = T(0.5);
            f = Vec3<T>{h * (a.x + b.x), h * (a.y + b.y), h * (a.z + b.z)};
        }
        r.d = Vec3<T>{-V * f.x, -V * f.y, -V * f.z};
    }
    if (A.mode != 1) {
        GradT<T> gr = (DIR == 0) ? du_i(g, u, i, j, k) : (DIR == 1) ? du_j(g, u, i, j, k) : du_k(g, u, i, j, k);
        T nuEff = T(A.nu);
        Vec3<T> base{0, 0, 0};
        if (A.les) {
            const T nt = T(0.5) * (A.nut[c] + A.nut[up]);
            const int f = (DIR == 0) ? i : (DIR == 1) ? j : k;
            const int m = (DIR == 0) ? g.mx : (DIR == 1) ? g.my : g.mz;
            const int bl = A.bc_u[2 * DIR], br = A.bc_u[2 * DIR + 1];
            if (DIR < 2 && ((bl == TOSCA_BCU_WALLFUNCTION && f == 0) || (br == TOSCA_BCU_WALLFUNCTION && f == m - 2))) {
                if (DIR == 0) { int t = k * g.my + j, P = g.mz * g.my; base = Vec3<T>{T(-A.tauI[t]), T(-A.tauI[P + t]), T(-A.tauI[2 * P + t])}; }
                else { int t = k * g.mx + i, P = g.mz * g.mx; base = Vec3<T>{T(-A.tauJ[t]), T(-A.tauJ[P + t]), T(-A.tauJ[2 * P + t])}; }
                nuEff = T(A.nu);
            } else if ((bl == TOSCA_BCU_SLIP && f == 0) || (br == TOSCA_BCU_SLIP && f == m - 2)) {
                nuEff = T(0.0);
            } else {
                nuEff = T(A.nu) + nt;
            }
        }
        const MT *fc = (DIR == 0) ? A.icsi : (DIR == 1) ? A.jcsi : A.kcsi;
        const MT *fe = (DIR == 0) ? A.ieta : (DIR == 1) ? A.jeta : A.keta;
        const MT *fz = (DIR == 0) ? A.izet : (DIR == 1) ? A.jzet : A.kzet;
        const MT *fa = (DIR == 0) ? A.iaj : (DIR == 1) ? A.jaj : A.kaj;
        r.v = add3(base, visc_flux<T>(gr, cvt3<T>(ld3(fc, c)), cvt3<T>(ld3(fe, c)), cvt3<T>(ld3(fz, c)), DIR, T(fa[c]), nuEff));
    }
    return r;
}

template <class MT>
__device__ __forceinline__ V3 ldnt3(const MT *p, int c) {
    return V3{double(__builtin_nontemporal_load(p + 3 * c)), double(__builtin_nontemporal_load(p + 3 * c + 1)),
              double(__builtin_nontemporal_load(p + 3 * c + 2))};
}
// the same, kept in the storage type
template <class MT>
__device__ __forceinline__ Vec3<MT> ldnt3s(const MT *p, int c) {
    return Vec3<MT>{__builtin_nontemporal_load(p + 3 * c), __builtin_nontemporal_load(p + 3 * c + 1),
                    __builtin_nontemporal_load(p + 3 * c + 2)};
}

// LDS ring of ucat planes: SoA [slot][component][row][col], rows j0-2 .. j0+NR, cols i0-2 .. i0+65
template <int NR, class T>
struct Ring {
    static constexpr int RI = 68, RJ = NR + 3, NS = 4;
    T v[NS][3][RJ][RI];
    __device__ __forceinline__ Vec3<T> at(int slot, int row, int col) const {
        return Vec3<T>{v[slot][0][row][col], v[slot][1][row][col], v[slot][2][row][col]};
    }
};


// fp differences in the reference's order: ((((a - b) + c) - d) + e) - f, div and visc apart
__device__ __forceinline__ double fpsum(double d1, double d1m, double d2, double d2m, double d3, double d3m,
                                        double v1, double v1m, double v2, double v2m, double v3, double v3m, int mode) {
    double f = 0.0;
    if (mode != 2) f += (d1 - d1m + d2 - d2m + d3 - d3m);
    if (mode != 1) f += (v1 - v1m + v2 - v2m + v3 - v3m);
    return f;
}

}  // namespace

// ------------------------------------------------------------------------------------------------
// boundary faces and periodic seams (general evaluation, ~2 % of the faces)
// ------------------------------------------------------------------------------------------------
// Buffers (doubles): bF[d] holds, for the two boundary faces 0 and m-2 of direction d, the div
// and visc flux (6 values) at every transverse position; on a periodic axis face 0 holds the flux
// of face m-2 (resetFacePeriodicFluxesVector).  seam[d] holds fp at cell 1 of a periodic axis d
// for every transverse interior position: what the ghost cell m-1 carries after
// resetCellPeriodicFluxes.
struct BoundaryBufs {
    double *bF[3];
    double *seam[3];
};

__host__ __device__ inline int tr_dims(const Geo &g, int d, int &n1, int &n2) {
    // transverse sizes (fastest first) of direction d
    if (d == 0) { n1 = g.my; n2 = g.mz; }
    else if (d == 1) { n1 = g.mx; n2 = g.mz; }
    else { n1 = g.mx; n2 = g.my; }
    return n1 * n2;
}

template <int DIR, class T, class MT>
__device__ __forceinline__ FluxT<T> face_general(const FormUArgsT<T, MT> &A, const Geo &g, int f, int a, int b) {
    // face f along DIR at transverse (a, b) = (j, k) | (i, k) | (i, j); periodic face 0 is face m-2
    const int m = (DIR == 0) ? g.mx : (DIR == 1) ? g.my : g.mz;
    if (f == 0 && g.per[DIR] != TOSCA_PERIODIC_NONE) f = m - 2;
    if (DIR == 0) return face_flux<0>(A, g, f, a, b);
    if (DIR == 1) return face_flux<1>(A, g, a, f, b);
    return face_flux<2>(A, g, a, b, f);
}

template <class T, class MT>
__global__ void __launch_bounds__(128) k_formu_boundary(FormUArgsT<T, MT> A, Geo g, BoundaryBufs B) {
    const int d = blockIdx.z / 3, what = blockIdx.z % 3;      // what: 0 face 0, 1 face m-2, 2 seam fp
    int n1, n2;
    const int n = tr_dims(g, d, n1, n2);
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int a = t % n1, b = t / n1;
    const int ma = (d == 0) ? g.my : g.mx, mb = (d == 2) ? g.my : g.mz;
    if (a < 1 || a > ma - 2 || b < 1 || b > mb - 2) return;     // transverse interior only
    const int m = (d == 0) ? g.mx : (d == 1) ? g.my : g.mz;
    if (what < 2) {
        const int f = what == 0 ? 0 : m - 2;
        FluxT<T> F = (d == 0) ? face_general<0>(A, g, f, a, b) : (d == 1) ? face_general<1>(A, g, f, a, b)
                                                                        : face_general<2>(A, g, f, a, b);
        double *o = B.bF[d] + ((long)what * n + t) * 6;
        o[0] = F.d.x; o[1] = F.d.y; o[2] = F.d.z; o[3] = F.v.x; o[4] = F.v.y; o[5] = F.v.z;
        return;
    }
    if (g.per[d] == TOSCA_PERIODIC_NONE) return;
    // fp at cell 1 along d (transverse interior), from its six faces
    int i, j, k;
    if (d == 0) { i = 1; j = a; k = b; } else if (d == 1) { i = a; j = 1; k = b; } else { i = a; j = b; k = 1; }
    const FluxT<T> f1 = face_general<0>(A, g, i, j, k), f1m = face_general<0>(A, g, i - 1, j, k);
    const FluxT<T> f2 = face_general<1>(A, g, j, i, k), f2m = face_general<1>(A, g, j - 1, i, k);
    const FluxT<T> f3 = face_general<2>(A, g, k, i, j), f3m = face_general<2>(A, g, k - 1, i, j);
    double *o = B.seam[d] + (long)t * 3;
    o[0] = fpsum(f1.d.x, f1m.d.x, f2.d.x, f2m.d.x, f3.d.x, f3m.d.x, f1.v.x, f1m.v.x, f2.v.x, f2m.v.x, f3.v.x, f3m.v.x, A.mode);
    o[1] = fpsum(f1.d.y, f1m.d.y, f2.d.y, f2m.d.y, f3.d.y, f3m.d.y, f1.v.y, f1m.v.y, f2.v.y, f2m.v.y, f3.v.y, f3m.v.y, A.mode);
    o[2] = fpsum(f1.d.z, f1m.d.z, f2.d.z, f2m.d.z, f3.d.z, f3m.d.z, f1.v.z, f1m.v.z, f2.v.z, f2m.v.z, f3.v.z, f3m.v.z, A.mode);
}

// ------------------------------------------------------------------------------------------------
// main k-marching kernel: interior faces only; boundary faces and seams come from the buffers
// ------------------------------------------------------------------------------------------------
template <class T, class MT>
struct MainArgs {
    const T *ucont, *ucat, *nut;
    const MT *csi, *eta, *zet;
    const MT *icsi, *ieta, *izet, *jcsi, *jeta, *jzet, *kcsi, *keta, *kzet;
    const MT *iaj, *jaj, *kaj;
    double *rhs;
    const double *farm;    // added to rhs at the output cells before the accumulation (nullptr: none)
    const T *theta;
    double *btheta;
    const double *bF1, *bF2, *bF3, *sI, *sJ, *sK;
    double *sv1, *sv2;     // j-tile seams: the two halves of the last output row's j-face dot
    double scale, nu, bscale, tRef;
    int mode, quick, les;
};

// boundary-buffer fluxes are stored as double; on the mixed path they hold float values exactly
template <class T>
__device__ __forceinline__ FluxT<T> ldflux(const double *p) {
    return FluxT<T>{Vec3<T>{T(p[0]), T(p[1]), T(p[2])}, Vec3<T>{T(p[3]), T(p[4]), T(p[5])}};
}

// Every array of the main kernel is addressed as its kernel-argument base (uniform, in scalar
// registers) plus a 32-bit element index per lane, instead of a 64-bit pointer per array and lane:
// the column's local index c0 plus the plane offset, made opaque once per plane (opaque_index).

template <int DIR, class T, class MT>
__device__ __forceinline__ const MT *face_fc(const MainArgs<T, MT> &A) { return DIR == 0 ? A.icsi : DIR == 1 ? A.jcsi : A.kcsi; }
template <int DIR, class T, class MT>
__device__ __forceinline__ const MT *face_fe(const MainArgs<T, MT> &A) { return DIR == 0 ? A.ieta : DIR == 1 ? A.jeta : A.keta; }
template <int DIR, class T, class MT>
__device__ __forceinline__ const MT *face_fz(const MainArgs<T, MT> &A) { return DIR == 0 ? A.izet : DIR == 1 ? A.jzet : A.kzet; }
template <int DIR, class T, class MT>
__device__ __forceinline__ const MT *face_fa(const MainArgs<T, MT> &A) { return DIR == 0 ? A.iaj : DIR == 1 ? A.jaj : A.kaj; }

// metrics and fluxes of one interior face, loaded ahead of use, the metrics in their storage type
template <class T, class MT>
struct FaceIn { Vec3<MT> fc, fe, fz; MT fa; T V, nutR; };

// c: local index of the face's cell
template <int DIR, class T, class MT>
__device__ __forceinline__ FaceIn<T, MT> face_in(const MainArgs<T, MT> &A, int c, int s) {
    FaceIn<T, MT> f;
    f.fc = ldnt3s(face_fc<DIR>(A), c);
    f.fe = ldnt3s(face_fe<DIR>(A), c);
    f.fz = ldnt3s(face_fz<DIR>(A), c);
    f.fa = face_fa<DIR>(A)[c];
    f.V = __builtin_nontemporal_load(A.ucont + 3 * c + DIR);
    f.nutR = A.les ? A.nut[c + s] : T(0);      // nu_t is read only when the LES term uses it
    return f;
}

// viscous flux of an interior face from the LDS ring: the normal difference and the two four-point
// transverse averages of Compute_du_i/j/k (inline.h:2748-2931, no boundary rule applies), then the
// flux with the face's metrics converted to T
template <int DIR, int NR, class T>
__device__ __forceinline__ Vec3<T> ring_visc(const Ring<NR, T> &R, int row, int col, int t, V3 fc, V3 fe, V3 fz,
                                             double fa, T nuEff) {
    auto U = [&](int dk, int dj, int di) { return R.at((t + dk) & 3, row + dj, col + di); };
    const int ok = (DIR == 2), oj = (DIR == 1), oi = (DIR == 0);
    const int ak = 0, aj = (DIR == 0), ai = (DIR != 0);
    const int bk = (DIR != 2), bj = (DIR == 2), bi = 0;
    const Vec3<T> u0 = U(0, 0, 0), u1 = U(ok, oj, oi);
    const Vec3<T> ap = U(ak + ok, aj + oj, ai + oi), a0 = U(ak, aj, ai), am = U(ok - ak, oj - aj, oi - ai), an = U(-ak, -aj, -ai);
    const Vec3<T> bp = U(bk + ok, bj + oj, bi + oi), b0 = U(bk, bj, bi), bm = U(ok - bk, oj - bj, oi - bi), bn = U(-bk, -bj, -bi);
    const T q4 = T(0.25);
    const T nx = u1.x - u0.x, ny = u1.y - u0.y, nz = u1.z - u0.z;
    const T ax = (ap.x + a0.x - am.x - an.x) * q4, ay = (ap.y + a0.y - am.y - an.y) * q4, az = (ap.z + a0.z - am.z - an.z) * q4;
    const T bx = (bp.x + b0.x - bm.x - bn.x) * q4, by = (bp.y + b0.y - bm.y - bn.y) * q4, bz = (bp.z + b0.z - bm.z - bn.z) * q4;
    GradT<T> g;
    if (DIR == 0) {        // normal csi; transverse a = eta (j), b = zet (k)
        g.dudc = nx; g.dvdc = ny; g.dwdc = nz; g.dude = ax; g.dvde = ay; g.dwde = az; g.dudz = bx; g.dvdz = by; g.dwdz = bz;
    } else if (DIR == 1) { // normal eta; transverse a = csi (i), b = zet (k)
        g.dudc = ax; g.dvdc = ay; g.dwdc = az; g.dude = nx; g.dvde = ny; g.dwde = nz; g.dudz = bx; g.dvdz = by; g.dwdz = bz;
    } else {               // normal zet; transverse a = csi (i), b = eta (j)
        g.dudc = ax; g.dvdc = ay; g.dwdc = az; g.dude = bx; g.dvde = by; g.dwde = bz; g.dudz = nx; g.dvdz = ny; g.dwdz = nz;
    }
    return visc_flux<T>(g, cvt3<T>(fc), cvt3<T>(fe), cvt3<T>(fz), DIR, T(fa), nuEff);
}

// convective flux of an interior face from the LDS ring (central or QUICK)
template <int DIR, int NR, class T>
__device__ __forceinline__ Vec3<T> ring_conv(const Ring<NR, T> &R, int row, int col, int t, int quick_, T V) {
    auto U = [&](int dk, int dj, int di) { return R.at((t + dk) & 3, row + dj, col + di); };
    const int ok = (DIR == 2), oj = (DIR == 1), oi = (DIR == 0);
    const Vec3<T> u0 = U(0, 0, 0), u1 = U(ok, oj, oi);
    Vec3<T> f;
    const T h = T(0.5);
    if (quick_) f = quick(U(-ok, -oj, -oi), u0, u1, U(2 * ok, 2 * oj, 2 * oi), V);
    else f = Vec3<T>{h * (u0.x + u1.x), h * (u0.y + u1.y), h * (u0.z + u1.z)};
    return Vec3<T>{-V * f.x, -V * f.y, -V * f.z};
}

// interior face flux from the LDS ring and preloaded face data (nut0: nu_t of the face's cell)
template <int DIR, int NR, class T, class MT>
__device__ __forceinline__ FluxT<T> face_pre(const MainArgs<T, MT> &A, const FaceIn<T, MT> &in, T nut0, const Ring<NR, T> &R,
                                             int row, int col, int t) {
    FluxT<T> r;
    r.d = Vec3<T>{0, 0, 0};
    r.v = Vec3<T>{0, 0, 0};
    if (A.mode != 2) r.d = ring_conv<DIR, NR, T>(R, row, col, t, A.quick, in.V);
    if (A.mode != 1) {
        T nuEff = T(A.nu);
        if (A.les) nuEff = T(A.nu) + T(0.5) * (nut0 + in.nutR);
        r.v = ring_visc<DIR, NR, T>(R, row, col, t, cvt3<double>(in.fc), cvt3<double>(in.fe), cvt3<double>(in.fz),
                                    double(in.fa), nuEff);
    }
    return r;
}

// interior face flux from the LDS ring, loading its face data; c is the local index of the face's cell
template <int DIR, int NR, class T, class MT>
__device__ __forceinline__ FluxT<T> face_main(const MainArgs<T, MT> &A, const Ring<NR, T> &R, int c, int row, int col,
                                              int t, int s) {
    FluxT<T> r;
    r.d = Vec3<T>{0, 0, 0};
    r.v = Vec3<T>{0, 0, 0};
    if (A.mode != 2) r.d = ring_conv<DIR, NR, T>(R, row, col, t, A.quick, A.ucont[3 * c + DIR]);
    if (A.mode != 1) {
        T nuEff = T(A.nu);
        if (A.les) nuEff = T(A.nu) + T(0.5) * (A.nut[c] + A.nut[c + s]);
        r.v = ring_visc<DIR, NR, T>(R, row, col, t, ldnt3(face_fc<DIR>(A), c), ldnt3(face_fe<DIR>(A), c),
                                    ldnt3(face_fz<DIR>(A), c), face_fa<DIR>(A)[c], nuEff);
    }
    return r;
}

// 0.5 a + 0.5 b in double: a cell-centred basis vector from the face bases on its two faces
// (AxByC(0.5, own face, 0.5, lower face), mesh.c:938-954; 0.5 x is exact, so the FMA the compiler
// may form rounds the same)
template <class MT>
__device__ __forceinline__ V3 face_avg(Vec3<MT> a, Vec3<MT> b) {
    return V3{0.5 * double(a.x) + 0.5 * double(b.x), 0.5 * double(a.y) + 0.5 * double(b.y),
              0.5 * double(a.z) + 0.5 * double(b.z)};
}

// DER: the cell-centred csi, eta, zet of the face dot are formed from the face bases icsi, jeta,
// kzet the plane steps load anyway (valid for meshes whose cell bases are those averages; ghost
// neighbours still load theirs)
template <int NR, class T, class MT, bool DER>
__global__ void __launch_bounds__(64 * NR) k_fused_formu(MainArgs<T, MT> A, Geo g, int kchunk) {
    // Every thread evaluates its three faces and fp on every plane (threads outside the tile's
    // useful range work on clamped, in-bounds addresses and their results are discarded), so the
    // hot loop has no divergent branches; only the stores and the rare boundary-buffer loads are
    // predicated.
    constexpr int TI = 62, TJ = NR - 1;
    __shared__ T sF2[6][NR][64];          // j-face fluxes of the plane (div xyz, visc xyz)
    __shared__ double sFp[2][3][NR][64];  // fp of the previous plane (double-buffered)
    __shared__ MT sEt[2][3][NR][DER ? 64 : 1];   // jeta of the plane (DER; double-buffered)
    __shared__ MT sEc[3][NR][DER ? 64 : 1];      // icsi of the plane (DER; read by its own wave only)
    __shared__ Ring<NR, T> ring;          // ucat planes t-1 .. t+2
    using RingT = Ring<NR, T>;
    constexpr int PLANE = RingT::RJ * RingT::RI * 3;
    constexpr int PRE = (PLANE + 64 * NR - 1) / (64 * NR);
    const int tid = threadIdx.x + 64 * threadIdx.y;
    const int lane = threadIdx.x, r = threadIdx.y;
    // tile order: j fastest, then i, then the k chunk (a j column of tiles shares its j halos)
    const int tile = xcd_tile();
    const int bj = tile % gridDim.y, bi = (tile / gridDim.y) % gridDim.x, bk = tile / (gridDim.x * gridDim.y);
    const int i0 = 1 + bi * TI, j0 = 1 + bj * TJ;
    const int kb = 1 + bk * kchunk;
    const int ke = min(kb + kchunk, g.mz - 1);      // output planes [kb, ke)
    const int i = i0 - 1 + lane, j = j0 - 1 + r;
    const bool iper = g.per[0] != TOSCA_PERIODIC_NONE, jper = g.per[1] != TOSCA_PERIODIC_NONE,
               kper = g.per[2] != TOSCA_PERIODIC_NONE;
    const bool colI = i >= 1 && i <= g.mx - 2, colJ = j >= 1 && j <= g.my - 2;
    const bool cell = lane >= 1 && r >= 1 && colI && colJ;
    const bool bndI = i == 0 || i == g.mx - 2, bndJ = j == 0 || j == g.my - 2;
    const bool seamI = colJ && i == g.mx - 1 && iper;       // ghost m-1 carries fp of cell 1
    const bool seamJ = colI && j == g.my - 1 && jper;
    const bool out = cell && lane <= TI && r <= TJ;
    // The last output row has no row above it in the tile: the j-face dot of its Rhs is split at
    // the seam.  It keeps the half of its own cell (v1) in sv1 and the next tile's first output
    // row keeps the half of its cell (the same v2 this row would form) in sv2; k_formu_jseam adds
    // both.  The domain's last interior row completes its dot here (own fp, or the periodic seam).
    const bool seamRow = r == NR - 1 && j < g.my - 2;
    const bool seamNext = r == 1 && bj > 0;
    // faces this thread evaluates: the halo row 0 only supplies the j face of row 1 and the halo
    // lane 0 only the i face of lane 1; every other thread needs its three faces for fp
    const bool needF1 = r >= 1, needF2 = lane >= 1, needF3 = r >= 1 && lane >= 1;
    const bool ownI = (i == g.mx - 2) && !iper, ownJ = (j == g.my - 2) && !jper;
    // clamped column for addressing
    const int ia = min(i, g.mx - 1), ja = min(j, g.my - 1);
    const int svo = (bj * g.mz) * g.mx + ia, svn = ((bj - 1) * g.mz) * g.mx + ia;
    const int SJ = g.gxm, SK = g.gxm * g.gym;
    const int c0 = g.L(0, ja, ia);            // local index of the column at plane 0
    const int SG = g.mx * g.my;
    const int g0 = ja * g.mx + ia;            // global (owned-range) index of the column at plane 0
    const bool buoy = A.btheta != A.rhs;   // the launcher passes btheta == rhs when Buoyancy is off
    const int nI = g.my * g.mz, nJ = g.mx * g.mz, nK = g.mx * g.my;
    const int oF1 = ((i == 0) ? 0 : nI) + ja;    // boundary-face buffers, in fluxes of 6 doubles: + k * my
    const int oF2 = ((j == 0) ? 0 : nJ) + ia;    // + k * mx
    const int oF3 = ja * g.mx + ia;              // + (0 | nK)

    // ucat ring: rows j0-2 .., cols i0-2 ..
    const int ri0 = i0 - 2, rj0 = j0 - 2;
    const int rowMax = min(RingT::RJ - 1, g.my + 2 - rj0), colMax = min(RingT::RI - 1, g.mx + 2 - ri0);
    auto plane_load = [&](int k, T *buf) {
        for (int n = 0; n < PRE; n++) {
            const int idx = tid + n * 64 * NR;
            buf[n] = 0.0;
            if (idx < PLANE) {
                const int row = idx / (3 * RingT::RI), e = idx % (3 * RingT::RI), col = e / 3;
                if (row <= rowMax && col <= colMax) buf[n] = A.ucat[3 * (long)g.L(k, rj0 + row, ri0) + e];
            }
        }
    };
    auto plane_store = [&](int k, const T *buf) {
        const int slot = k & 3;
        for (int n = 0; n < PRE; n++) {
            const int idx = tid + n * 64 * NR;
            if (idx < PLANE) {
                const int row = idx / (3 * RingT::RI), e = idx % (3 * RingT::RI);
                ring.v[slot][e % 3][row][e / 3] = buf[n];
            }
        }
    };
    const int rrow = r + 1, rcol = lane + 1;
    auto kface = [&](int k) -> FluxT<T> {      // k face k of this column
        if (k == 0) return ldflux<T>(A.bF3 + 6 * oF3);
        if (k == g.mz - 2) return ldflux<T>(A.bF3 + 6 * (oF3 + nK));
        return face_main<2, NR, T>(A, ring, c0 + k * SK, rrow, rcol, k, SK);
    };
    // prime: k face kb-1 needs planes kb-2 .. kb+1, the first plane step kb-1 .. kb+2
    FluxT<T> F3p;
    {
        T buf[PRE];
        for (int k = kb - 2; k <= kb + 1; k++) { plane_load(k, buf); plane_store(k, buf); }
        __syncthreads();
        if (needF3) F3p = kface(kb - 1);
        else F3p = FluxT<T>{Vec3<T>{0, 0, 0}, Vec3<T>{0, 0, 0}};
        plane_load(kb + 2, buf);
        __syncthreads();
        plane_store(kb + 2, buf);
        __syncthreads();
    }
    V3 fpPrev{0, 0, 0};
    int buf = 0;
    // values of plane t the next plane step reuses instead of loading them again: the k+1 cell
    // metrics zet and theta of the Rhs / buoyancy step, and nu_t of the k face's upper cell
    V3 zeNext{0, 0, 0};
    double thNext = 0;
    T nutNext = T(0);
    // values of plane t-1 the Rhs step reuses: inverse face Jacobians and the z components of the
    // face bases (buoyancy), loaded with the face data of that plane
    const int cp0 = c0 + (kb - 1) * SK;
    double fa0p = A.iaj[cp0], fa1p = A.jaj[cp0], fa2p = A.kaj[cp0];
    double gzIp = A.icsi[3 * cp0 + 2], gzJp = A.jeta[3 * cp0 + 2], gzKp = A.kzet[3 * cp0 + 2];
    // DER: kzet of plane t-1 (of plane kb-1 for the first step's zet); icsi and jeta of plane t-1
    // are read back from their LDS copies
    Vec3<MT> fzP{0, 0, 0};
    if constexpr (DER) fzP = ld3(A.kzet, cp0);
    // With FP32 metric copies the face data of the next plane step fit in the registers the kernel
    // has left at its occupancy (2 waves per SIMD), so they are loaded one step ahead; with FP64
    // metrics they would spill.
    constexpr bool AHEAD = std::is_same<MT, float>::value;
    FaceIn<T, MT> nx1{}, nx2{}, nx3{};
    if constexpr (AHEAD) {
        const int cn = opaque_index(c0 + min(kb, g.mz - 2) * SK);
        if (needF1) nx1 = face_in<0>(A, cn, 1);
        if (needF2) nx2 = face_in<1>(A, cn, SJ);
        if (needF3) nx3 = face_in<2>(A, cn, SK);
    }
    for (int t = kb; t <= ke; t++) {
        // ---- every global load of this plane step, issued together ------------------------------
        T pre[PRE];
        const bool doPre = t + 3 <= ke + 2;
        if (doPre) plane_load(t + 3, pre);
        const bool ghostK = (t == g.mz - 1);
        // faces of plane t (clamped at the ghost plane): local index of the face's cell
        const int cq = opaque_index(c0 + min(t, g.mz - 2) * SK);
        FaceIn<T, MT> in1{}, in2{}, in3{};
        if constexpr (AHEAD) {
            in1 = nx1; in2 = nx2; in3 = nx3;
            if (t + 1 <= ke) {
                const int cn = opaque_index(c0 + min(t + 1, g.mz - 2) * SK);
                if (needF1) nx1 = face_in<0>(A, cn, 1);
                if (needF2) nx2 = face_in<1>(A, cn, SJ);
                if (needF3) nx3 = face_in<2>(A, cn, SK);
            }
        } else {
            if (needF1) in1 = face_in<0>(A, cq, 1);
            if (needF2) in2 = face_in<1>(A, cq, SJ);
            if (needF3) in3 = face_in<2>(A, cq, SK);
        }
        // nu_t is read only when the LES term uses it (A.les, uniform across the launch); A.nut
        // still must never be null: the launcher substitutes a valid local array when the caller
        // passes none (a null nu_t loaded here unconditionally faulted the GPU once, in formMode 1)
        const T nut0 = (needF3 && t > kb) ? nutNext : (A.les ? A.nut[cq] : T(0));
        // ---- face fluxes ---------------------------------------------------------------------------
        const Vec3<T> z3{0, 0, 0};
        FluxT<T> F1{z3, z3}, F2{z3, z3}, F3{z3, z3};
        if (!ghostK) {
            if (needF1) F1 = bndI ? ldflux<T>(A.bF1 + 6 * (oF1 + t * g.my)) : face_pre<0, NR, T>(A, in1, nut0, ring, rrow, rcol, t);
            if (needF2) F2 = bndJ ? ldflux<T>(A.bF2 + 6 * (oF2 + t * g.mx)) : face_pre<1, NR, T>(A, in2, nut0, ring, rrow, rcol, t);
            if (needF3) {
                if (t == g.mz - 2) F3 = ldflux<T>(A.bF3 + 6 * (oF3 + nK));
                else F3 = face_pre<2, NR, T>(A, in3, nut0, ring, rrow, rcol, t);
            }
        }
        // ---- loads of the Rhs / buoyancy step (cells of plane t-1), in flight across the exchange;
        // output cells only
        const int cpq = opaque_index(c0 + (t - 1) * SK);        // the cells of plane t-1
        const bool rhsStep = out && t > kb;
        V3 csA{0, 0, 0}, csB{0, 0, 0}, etA{0, 0, 0}, etB{0, 0, 0}, zeA{0, 0, 0}, zeB{0, 0, 0};
        const int ro = 3 * (g0 + (t - 1) * SG);                 // rhs, btheta, farm (owned range)
        double r0 = 0, r1 = 0, r2 = 0;
        double th = 0, thI = 0, thJ = 0, thK = 0;
        if (out) {
            // the upper cell's zet and theta: this step's k neighbours, the next step's own values
            if constexpr (DER) zeB = ghostK ? cvt3<double>(ld3(A.zet, cpq + SK)) : face_avg(in3.fz, fzP);
            else zeB = cvt3<double>(ld3(A.zet, cpq + SK));
            if (buoy) thK = A.theta[cpq + SK];
        }
        if (rhsStep) {
            if constexpr (DER) {
                // the ghost neighbours of the last interior cells keep their own (mirrored or
                // periodic) bases; the other bases are formed below
                if (i == g.mx - 2) csB = cvt3<double>(ld3(A.csi, cpq + 1));
                if (j == g.my - 2) etB = cvt3<double>(ld3(A.eta, cpq + SJ));
            } else {
                csA = cvt3<double>(ld3(A.csi, cpq)); csB = cvt3<double>(ld3(A.csi, cpq + 1));
                etA = cvt3<double>(ld3(A.eta, cpq)); etB = cvt3<double>(ld3(A.eta, cpq + SJ));
            }
            zeA = zeNext;
            r0 = A.rhs[ro]; r1 = A.rhs[ro + 1]; r2 = A.rhs[ro + 2];
            if (A.farm) {
                // VecAXPY(Rhs, 1, sourceFarmCont) before FormU's accumulation (ueqn_solvers.c:1341-1347)
                r0 = r0 + 1.0 * A.farm[ro]; r1 = r1 + 1.0 * A.farm[ro + 1]; r2 = r2 + 1.0 * A.farm[ro + 2];
            }
            // theta, like nu_t, must be a valid local array even when Buoyancy is off (the launcher
            // substitutes one); only the buoy flag decides whether it is read
            if (buoy) { th = thNext; thI = A.theta[cpq + 1]; thJ = A.theta[cpq + SJ]; }
        }
        const Vec3<T> F1dm = shup3(F1.d), F1vm = shup3(F1.v);
        sF2[0][r][lane] = F2.d.x; sF2[1][r][lane] = F2.d.y; sF2[2][r][lane] = F2.d.z;
        sF2[3][r][lane] = F2.v.x; sF2[4][r][lane] = F2.v.y; sF2[5][r][lane] = F2.v.z;
        if constexpr (DER) { sEt[buf][0][r][lane] = in2.fe.x; sEt[buf][1][r][lane] = in2.fe.y; sEt[buf][2][r][lane] = in2.fe.z; }
        __syncthreads();
        const int rm = r > 0 ? r - 1 : 0;
        const Vec3<T> F2dm{sF2[0][rm][lane], sF2[1][rm][lane], sF2[2][rm][lane]};
        const Vec3<T> F2vm{sF2[3][rm][lane], sF2[4][rm][lane], sF2[5][rm][lane]};
        V3 fp;
        fp.x = fpsum(F1.d.x, F1dm.x, F2.d.x, F2dm.x, F3.d.x, F3p.d.x, F1.v.x, F1vm.x, F2.v.x, F2vm.x, F3.v.x, F3p.v.x, A.mode);
        fp.y = fpsum(F1.d.y, F1dm.y, F2.d.y, F2dm.y, F3.d.y, F3p.d.y, F1.v.y, F1vm.y, F2.v.y, F2vm.y, F3.v.y, F3p.v.y, A.mode);
        fp.z = fpsum(F1.d.z, F1dm.z, F2.d.z, F2dm.z, F3.d.z, F3p.d.z, F1.v.z, F1vm.z, F2.v.z, F2vm.z, F3.v.z, F3p.v.z, A.mode);
        if (ghostK) {
            if (kper) { const double *qq = A.sK + 3 * (ja * g.mx + ia); fp = V3{qq[0], qq[1], qq[2]}; }
            else fp = fpPrev;                        // non-periodic top: ghost fp copies cell m-2
        } else if (seamI) {
            const double *qq = A.sI + 3 * (ja + t * g.my);
            fp = V3{qq[0], qq[1], qq[2]};
        } else if (seamJ) {
            const double *qq = A.sJ + 3 * (ia + t * g.mx);
            fp = V3{qq[0], qq[1], qq[2]};
        }
        // Rhs of the previous plane (t-1): fp(t-1) own / i+1 (shuffle) / j+1 (LDS), fp(t) own
        V3 fpI = shdn3(fpPrev);
        const int rp = r + 1 < NR ? r + 1 : r;
        V3 fJ{sFp[buf ^ 1][0][rp][lane], sFp[buf ^ 1][1][rp][lane], sFp[buf ^ 1][2][rp][lane]};
        if (ownI) fpI = fpPrev;
        if (ownJ) fJ = fpPrev;
        else if (r == NR - 1 && jper && j == g.my - 2) {
            const double *qq = A.sJ + 3 * (ia + (t - 1) * g.mx);
            fJ = V3{qq[0], qq[1], qq[2]};
        }
        Vec3<MT> fcP{0, 0, 0}, feP{0, 0, 0};
        if constexpr (DER) {
            // csi of cells i, i+1 from icsi of faces i-1, i, i+1 (lanes), eta of rows j, j+1 from
            // jeta of rows j-1, j, j+1 (the previous step's LDS copy), all of plane t-1
            fcP = Vec3<MT>{sEc[0][r][lane], sEc[1][r][lane], sEc[2][r][lane]};
            const Vec3<MT> fcL = shup3(fcP), fcU = shdn3(fcP);
            const int rmE = r > 0 ? r - 1 : 0, rpE = r + 1 < NR ? r + 1 : r;
            feP = Vec3<MT>{sEt[buf ^ 1][0][r][lane], sEt[buf ^ 1][1][r][lane], sEt[buf ^ 1][2][r][lane]};
            const Vec3<MT> feM{sEt[buf ^ 1][0][rmE][lane], sEt[buf ^ 1][1][rmE][lane], sEt[buf ^ 1][2][rmE][lane]};
            const Vec3<MT> feN{sEt[buf ^ 1][0][rpE][lane], sEt[buf ^ 1][1][rpE][lane], sEt[buf ^ 1][2][rpE][lane]};
            if (rhsStep) {
                csA = face_avg(fcP, fcL);
                if (i != g.mx - 2) csB = face_avg(fcU, fcP);
                etA = face_avg(feP, feM);
                if (j != g.my - 2) etB = face_avg(feN, feP);
            }
        }
        if (rhsStep) {
            const V3 fK = fp;
            double v1 = csA.x * fpPrev.x + csA.y * fpPrev.y + csA.z * fpPrev.z;
            double v2 = csB.x * fpI.x + csB.y * fpI.y + csB.z * fpI.z;
            const double rx = A.scale * fa0p * (0.5 * (v1 + v2));
            v1 = etA.x * fpPrev.x + etA.y * fpPrev.y + etA.z * fpPrev.z;
            v2 = etB.x * fJ.x + etB.y * fJ.y + etB.z * fJ.z;
            const double ry = A.scale * fa1p * (0.5 * (v1 + v2));
            if (seamRow) A.sv1[svo + (t - 1) * g.mx] = v1;
            if (seamNext) A.sv2[svn + (t - 1) * g.mx] = v1;
            v1 = zeA.x * fpPrev.x + zeA.y * fpPrev.y + zeA.z * fpPrev.z;
            v2 = zeB.x * fK.x + zeB.y * fK.y + zeB.z * fK.z;
            const double rz = A.scale * fa2p * (0.5 * (v1 + v2));
            A.rhs[ro] = r0 + rx;
            if (!seamRow) A.rhs[ro + 1] = r1 + ry;
            else if (A.farm) A.rhs[ro + 1] = r1;      // k_formu_jseam adds the dot to rhs + farm
            A.rhs[ro + 2] = r2 + rz;
            if (buoy) {
                // Boussinesq source on the three + faces of the cell (ueqn_sources.c:2812-2870):
                // g . S with g = (0, 0, -9.81) is -9.81 * S.z exactly
                const double tI = 0.5 * (th + thI), tJ = 0.5 * (th + thJ), tK = 0.5 * (th + thK);
                const double gzI = DER ? double(fcP.z) : gzIp, gzJ = DER ? double(feP.z) : gzJp,
                             gzK = DER ? double(fzP.z) : gzKp;
                A.btheta[ro] = A.bscale * ((-9.81 * gzI) * (A.tRef - tI) / A.tRef);
                A.btheta[ro + 1] = A.bscale * ((-9.81 * gzJ) * (A.tRef - tJ) / A.tRef);
                A.btheta[ro + 2] = A.bscale * ((-9.81 * gzK) * (A.tRef - tK) / A.tRef);
            }
        }
        sFp[buf][0][r][lane] = fp.x; sFp[buf][1][r][lane] = fp.y; sFp[buf][2][r][lane] = fp.z;
        if (doPre) plane_store(t + 3, pre);
        __syncthreads();
        buf ^= 1;
        fpPrev = fp;
        F3p = F3;
        fa0p = in1.fa; fa1p = in2.fa; fa2p = in3.fa;
        if constexpr (DER) {
            // this wave's row only: its reads of the plane's icsi above precede this write
            sEc[0][r][lane] = in1.fc.x; sEc[1][r][lane] = in1.fc.y; sEc[2][r][lane] = in1.fc.z;
            fzP = in3.fz;
        }
        else { gzIp = in1.fc.z; gzJp = in2.fe.z; gzKp = in3.fz.z; }
        zeNext = zeB;
        thNext = thK;
        nutNext = in3.nutR;
    }
}

// Rhs.y of the j-tile seam rows: rhs += scale aj_j (0.5 (v1 + v2)) with the two halves the main
// kernel kept (the same expression, so the same value, as the other rows' in-kernel update)
template <class MT>
__global__ void k_formu_jseam(double *rhs, const double *sv1, const double *sv2, const MT *jaj, double scale, Geo g,
                              int tj) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, k = blockIdx.y, sidx = blockIdx.z;
    const int jA = (sidx + 1) * tj;                 // last output row of j tile sidx
    if (i < 1 || i > g.mx - 2 || k < 1 || k > g.mz - 2 || jA >= g.my - 2) return;
    const long o = ((long)sidx * g.mz + k) * g.mx + i;
    const double fa = jaj[g.L(k, jA, i)];
    double *pr = rhs + 3 * (long)g.G(k, jA, i) + 1;
    *pr = *pr + scale * fa * (0.5 * (sv1[o] + sv2[o]));
}

constexpr int FORMU_NR = 8;      // rows of the main kernel's workgroup (one wave each)

template <class T, class MT>
void fused_formu(const Geo &g, const FormUArgsT<T, MT> &A, double *ws, hipStream_t s) {
    constexpr int NR = FORMU_NR;
    const int TI = 62, TJ = NR - 1;
    // boundary faces and seams first
    BoundaryBufs B;
    long off = 0;
    int n1, n2;
    for (int d = 0; d < 3; d++) { const long n = tr_dims(g, d, n1, n2); B.bF[d] = ws + off; off += 2 * n * 6; }
    for (int d = 0; d < 3; d++) { const long n = tr_dims(g, d, n1, n2); B.seam[d] = ws + off; off += n * 3; }
    const int nsj = (g.my - 2 + TJ - 1) / TJ;                 // j tiles; the last one has no seam
    double *sv1 = ws + off, *sv2 = ws + off + (long)nsj * g.mz * g.mx;
    const int nmax = max(g.my * g.mz, max(g.mx * g.mz, g.mx * g.my));
    TOSCA_LAUNCH(sizeof(T) == 8 ? "k_formu_boundary" : sizeof(MT) == 8 ? "k_formu_boundary_f32" : "k_formu_boundary_f32m", (k_formu_boundary<T, MT>), dim3((nmax + 127) / 128, 1, 9), dim3(128), 0, s, A, g, B);

    MainArgs<T, MT> M;
    // the main kernel loads nu_t unconditionally; without LES (or without the viscous term) the
    // caller passes no nu_t, so any valid local array stands in (its values are never used)
    M.ucont = A.ucont; M.ucat = A.ucat; M.nut = A.nut ? A.nut : A.ucat;
    M.csi = A.csi; M.eta = A.eta; M.zet = A.zet;
    M.icsi = A.icsi; M.ieta = A.ieta; M.izet = A.izet; M.jcsi = A.jcsi; M.jeta = A.jeta; M.jzet = A.jzet;
    M.kcsi = A.kcsi; M.keta = A.keta; M.kzet = A.kzet; M.iaj = A.iaj; M.jaj = A.jaj; M.kaj = A.kaj;
    M.rhs = A.rhs;
    M.farm = A.farm;
    M.theta = A.btheta ? A.theta : A.ucat;          // any valid local array when Buoyancy is off
    M.btheta = A.btheta ? A.btheta : A.rhs;
    M.bF1 = B.bF[0]; M.bF2 = B.bF[1]; M.bF3 = B.bF[2]; M.sI = B.seam[0]; M.sJ = B.seam[1]; M.sK = B.seam[2];
    M.sv1 = sv1; M.sv2 = sv2;
    M.scale = A.scale; M.nu = A.nu; M.bscale = A.bscale; M.tRef = A.btheta ? A.tRef : 1.0;
    M.mode = A.mode; M.quick = A.quick; M.les = A.les;

    const int nk = g.mz - 2;
    const int tiles = ((g.mx - 2 + TI - 1) / TI) * ((g.my - 2 + TJ - 1) / TJ);
    int chunks = (2048 + tiles - 1) / tiles;
    chunks = max(1, min(chunks, nk / 8 > 0 ? nk / 8 : 1));
    const int kchunk = (nk + chunks - 1) / chunks;
    chunks = (nk + kchunk - 1) / kchunk;
    dim3 grid((g.mx - 2 + TI - 1) / TI, (g.my - 2 + TJ - 1) / TJ, chunks);
    const char *kname = sizeof(T) == 8 ? "k_fused_formu" : sizeof(MT) == 8 ? "k_fused_formu_f32" : "k_fused_formu_f32m";
    // the cell bases formed from the face bases on FP64 metrics only (on the FP32 copies, whose
    // kernel is latency-bound, that measured slower than loading them)
    bool derive = false;
    if constexpr (std::is_same<MT, double>::value) derive = A.derive_cc != 0;
    if constexpr (std::is_same<MT, double>::value) {
        if (derive) TOSCA_LAUNCH(kname, (k_fused_formu<NR, T, MT, true>), grid, dim3(64, NR), 0, s, M, g, kchunk);
    }
    if (!derive) TOSCA_LAUNCH(kname, (k_fused_formu<NR, T, MT, false>), grid, dim3(64, NR), 0, s, M, g, kchunk);
    // the farm AXPY on the owned points outside the interior cells (the kernel covers the interior)
    if (A.farm) shell_add(A.rhs, A.farm, g.mx, g.my, g.mz, s);
    if (nsj > 1)
        TOSCA_LAUNCH("k_formu_jseam", k_formu_jseam<MT>, dim3((g.mx + 127) / 128, g.mz, nsj - 1), dim3(128), 0, s,
                     A.rhs, sv1, sv2, A.jaj, A.scale, g, TJ);
    // the reference zeroes all of bTheta; the kernel writes the interior cells, this the rest
    if (A.btheta) zero_shell(A.btheta, 3, g.mx, g.my, g.mz, 0, 0, s);
}

template void fused_formu<double, double>(const Geo &, const FormUArgsT<double, double> &, double *, hipStream_t);
template void fused_formu<float, double>(const Geo &, const FormUArgsT<float, double> &, double *, hipStream_t);
template void fused_formu<float, float>(const Geo &, const FormUArgsT<float, float> &, double *, hipStream_t);

size_t formu_ws_doubles(const Geo &g) {
    long tot = 0;
    int n1, n2;
    for (int d = 0; d < 3; d++) tot += (long)tr_dims(g, d, n1, n2) * (12 + 3);
    const int TJ = FORMU_NR - 1;                              // j tiles of the main kernel
    const int nsj = (g.my - 2 + TJ - 1) / TJ;
    tot += 2L * nsj * g.mz * g.mx;
    return (size_t)tot;
}

}  // namespace tosca
