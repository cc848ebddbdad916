// fused_les.hip -- dynamic Smagorinsky on gfx950 in two kernels.
//
// The reference's UpdateCs (src/les.c:197-1977) makes four HBM sweeps with nine workspace
// vectors: strain rows lSx/lSy/lSz/lS, the 27-point test filter into lLM/lMM, ghost scatters,
// and the 27-point box average into lCs; UpdateNut (src/les.c:2152-2514) then recomputes |S|.
//   k_dsm_lm: one workgroup per 62 x (NR-2) tile of (i, j) cells marches along k.  Per plane each
//     thread evaluates, at its point, the strain (Compute_du_center + Compute_du_dxyz), |S| and
//     the 27 Germano quantities of les.c:466-611 (u, U = u.csi, U u, S_ij, |S| S_ij) plus the
//     volume term of the test-filter width, masked to the interior cells (les.c:634-650).  The
//     (1,4,1) Simpson test filter is applied separably: along i across the lanes of the wave,
//     along j through LDS, along k in registers (two planes of partial sums).  The 27 filtered
//     quantities are contracted (les.c:715-799) into LM, MM; |S| of the cell is kept for nu_t.
//     The separable order sums the same 27 terms in another association than the reference's
//     triple loop; on the benchmark and test fields the dynamic coefficient moves by <1e-13
//     (normwise) under that reordering.
//   k_dsm_cs_nut: 3x3x3 volume-weighted Simpson box average of LM/MM with the reference's
//     periodic remap and boundary-ghost weights (les.c:1885-1950), clip to [0, maxCs], and
//     nu_t = Cs Delta^2 |S| (les.c:2510-2514) -- Cs and nu_t written in one pass.
#include <type_traits>

#include "fused.h"

namespace tosca {

namespace {
constexpr int NQ = 28;   // 27 Germano quantities + the binomial-weighted volume of the filter width
constexpr int DSM_NR = 8;
constexpr int DSM_TI = 62, DSM_TJ = DSM_NR - 2;


// sum of the 1-D Simpson weights (1, 4, 1) over the interior points c-1, c, c+1 of an axis of size m
__device__ __forceinline__ double simpson_mask_sum(int c, int m) {
    return ((c - 1 >= 1) ? 1.0 : 0.0) + 4.0 + ((c + 1 <= m - 2) ? 1.0 : 0.0);
}
}  // namespace

// Type of the filtered quantities and their sums: FP64 with FP64 flow fields.  With FP32 flow
// fields they are FP32, after the tile's reference velocity is taken out of both factors of the
// Leonard term (below), so that L = <U u> - <U><u> does not lose its digits to cancellation.
template <class T> struct DsmSum { using type = double; };
template <> struct DsmSum<float> { using type = float; };

template <class T>
__global__ void __launch_bounds__(64 * DSM_NR) k_dsm_lm(Geo g, const T *ucat, const double *csi,
                                                        const double *eta, const double *zet, const double *aj,
                                                        double *LM, double *MM, double *Sabs, int kchunk) {
    using S = typename DsmSum<T>::type;
    constexpr bool shifted = !std::is_same<S, double>::value;
    __shared__ T sU[3][DSM_NR][64];            // ucat of the plane, for the j neighbours
    __shared__ S sH[NQ][DSM_NR][64];           // i-filtered quantities of the plane
    const int lane = threadIdx.x, r = threadIdx.y;
    const int tile = xcd_tile();                // j fastest, then i, then the k chunk
    const int bj = tile % gridDim.y, bi = (tile / gridDim.y) % gridDim.x, bk = tile / (gridDim.x * gridDim.y);
    const int i = bi * DSM_TI + lane, j = bj * DSM_TJ + r;     // this thread's point column
    const int kb = 1 + bk * kchunk, ke = min(kb + kchunk, g.mz - 1);   // output planes [kb, ke)
    // L_ij = <U_i u_j> - <U_i><u_j> is unchanged when U_i and u_j are shifted by constants (the
    // filter is normalised), so with FP32 sums every point of the tile's march takes the velocity and
    // the contravariant velocity of one interior cell of the tile out of the two factors
    Vec3<T> uref{T(0), T(0), T(0)}, Uref{T(0), T(0), T(0)};
    if constexpr (shifted) {
        const int ic = min(max(bi * DSM_TI + 32, 1), g.mx - 2), jc = min(max(bj * DSM_TJ + 3, 1), g.my - 2);
        const int kc = min(kb, g.mz - 2);
        const int cr = g.L(kc, jc, ic);
        uref = ld3(ucat, cr);
        const Vec3<T> cr3 = cvt3<T>(ld3(csi, cr)), er3 = cvt3<T>(ld3(eta, cr)), zr3 = cvt3<T>(ld3(zet, cr));
        Uref = Vec3<T>{uref.x * cr3.x + uref.y * cr3.y + uref.z * cr3.z, uref.x * er3.x + uref.y * er3.y + uref.z * er3.z,
                       uref.x * zr3.x + uref.y * zr3.y + uref.z * zr3.z};
    }
    const bool mi = i >= 1 && i <= g.mx - 2, mj = j >= 1 && j <= g.my - 2;
    const bool outCol = lane >= 1 && lane <= DSM_TI && r >= 1 && r <= DSM_TJ && mi && mj;
    // addresses of columns past the array are clamped; those points are masked out
    const int ia = min(i, g.mx + 2), ja = min(j, g.my + 2);
    const int ima = max(ia - 1, -GW), ipa = min(ia + 1, g.mx + 2), jma = max(ja - 1, -GW), jpa = min(ja + 1, g.my + 2);
    const double dij = mi && mj ? simpson_mask_sum(i, g.mx) * simpson_mask_sum(j, g.my) : 1.0;
    const int SK = g.gxm * g.gym;
    // every array is addressed as its kernel-argument base plus a 32-bit index (the column's local
    // index at plane 0 plus the plane offset), not through 64-bit pointers kept per lane
    const int c0 = g.L(0, ja, ia);

    // k-ring of ucat for this column: planes K-1, K, K+1 of the current point plane K
    Vec3<T> uKm = ld3(ucat, c0 + (kb - 2) * SK), uK = ld3(ucat, c0 + (kb - 1) * SK);
    // the metrics of the previous plane's point, the cell the finished output belongs to: carried in
    // registers where the FP32 sums leave room for them, loaded again with FP64 sums
    V3 cP{0, 0, 0}, eP{0, 0, 0}, zP{0, 0, 0};
    double aP = 1.0;

    // partial sums along k: A for output K-1 (terms K-2, K-1 in), B for output K (term K-1 in)
    S A[NQ], B[NQ];
#pragma unroll
    for (int n = 0; n < NQ; n++) { A[n] = S(0); B[n] = S(0); }
    double sabsPrev = 0.0;

    for (int K = kb - 1; K <= ke; K++) {
        // loads of this plane, issued before the exchange barrier (a register prefetch of the next
        // plane spills at this register budget and measured slower)
        const int cK = opaque_index(c0 + K * SK);
        const Vec3<T> uKp = ld3(ucat, cK + SK);
        const V3 cC = ld3(csi, cK), eC = ld3(eta, cK), zC = ld3(zet, cK);
        const double aC = aj[cK];
        // ---- neighbours of the point (i, j, K) ----------------------------------------------------
        sU[0][r][lane] = uK.x; sU[1][r][lane] = uK.y; sU[2][r][lane] = uK.z;
        const bool pt = mi && mj && K >= 1 && K <= g.mz - 2;     // point inside the filter mask
        Vec3<T> uim = shup3(uK), uip = shdn3(uK);
        if (lane == 0) uim = ld3(ucat, g.L(K, ja, ima));
        if (lane == 63) uip = ld3(ucat, g.L(K, ja, ipa));
        __syncthreads();
        Vec3<T> ujm, ujp;
        if (r > 0) ujm = Vec3<T>{sU[0][r - 1][lane], sU[1][r - 1][lane], sU[2][r - 1][lane]};
        else ujm = ld3(ucat, g.L(K, jma, ia));
        if (r < DSM_NR - 1) ujp = Vec3<T>{sU[0][r + 1][lane], sU[1][r + 1][lane], sU[2][r + 1][lane]};
        else ujp = ld3(ucat, g.L(K, jpa, ia));
        // ---- point quantities (les.c:466-611), zero outside the mask -------------------------------
        // base values b: u v w, U V W, Sxx Sxy Sxz Syy Syz Szz, |S|, 1/aj; the 28 filtered quantities
        // are formed from them one at a time, right before their i filter
        S b[14];
#pragma unroll
        for (int n = 0; n < 14; n++) b[n] = S(0);
        if (pt) {
            // velocity arithmetic in T (the strain, |S| and the contravariant velocity); the
            // filtered quantities below are formed and summed in S
            const Vec3<T> cT = cvt3<T>(cC), eT = cvt3<T>(eC), zT = cvt3<T>(zC);
            const CartT<T> d = du_dxyz(cT, eT, zT, T(aC), du_center_v(g, i, j, K, uK, uim, uip, ujm, ujp, uKm, uKp));
            const T h = T(0.5);
            b[0] = uK.x; b[1] = uK.y; b[2] = uK.z;
            b[3] = uK.x * cT.x + uK.y * cT.y + uK.z * cT.z;
            b[4] = uK.x * eT.x + uK.y * eT.y + uK.z * eT.z;
            b[5] = uK.x * zT.x + uK.y * zT.y + uK.z * zT.z;
            if constexpr (shifted) {
                b[0] -= uref.x; b[1] -= uref.y; b[2] -= uref.z;
                b[3] -= Uref.x; b[4] -= Uref.y; b[5] -= Uref.z;
            }
            b[6] = h * (d.du_dx + d.du_dx); b[7] = h * (d.du_dy + d.dv_dx); b[8] = h * (d.du_dz + d.dw_dx);
            b[9] = h * (d.dv_dy + d.dv_dy); b[10] = h * (d.dv_dz + d.dw_dy); b[11] = h * (d.dw_dz + d.dw_dz);
            b[12] = strain_mag(d);
            b[13] = S(1. / aC);
        }
        const double sabs = b[12];
        // ---- filter along i (lanes), then j (LDS) ---------------------------------------------------
#pragma unroll
        for (int n = 0; n < NQ; n++) {
            S v;
            if (n < 6) v = b[n];
            else if (n < 15) v = b[3 + (n - 6) / 3] * b[(n - 6) % 3];      // U_a u_b
            else if (n < 21) v = b[6 + (n - 15)];                            // S_ij
            else if (n < 27) v = b[6 + (n - 21)] * b[12];                    // S_ij |S|
            else v = b[13];
            const S lo = shup(v), hi = shdn(v);
            sH[n][r][lane] = (n < 27) ? (lo + S(4.0) * v + hi) : (S(0.25) * lo + S(0.5) * v + S(0.25) * hi);
        }
        __syncthreads();
        const int rm = r > 0 ? r - 1 : r, rp = r < DSM_NR - 1 ? r + 1 : r;
        // ---- filter along k, and the finished output cell K-1 --------------------------------------
        const bool emit = outCol && K - 1 >= kb && K - 1 < ke;
        S q[NQ];
#pragma unroll
        for (int n = 0; n < NQ; n++) {
            const S f = (n < 27) ? (sH[n][rm][lane] + S(4.0) * sH[n][r][lane] + sH[n][rp][lane])
                                 : (S(0.25) * sH[n][rm][lane] + S(0.5) * sH[n][r][lane] + S(0.25) * sH[n][rp][lane]);
            const S wK = (n < 27) ? S(1.0) : S(0.25), wC = (n < 27) ? S(4.0) : S(0.5);
            const S done = A[n] + wK * f;            // output K-1 complete
            A[n] = B[n] + wC * f;
            B[n] = wK * f;
            q[n] = done;
        }
        if (emit) {
            const int k = K - 1;
            const long c = g.L(k, j, i);
            const double rden = 1.0 / (dij * simpson_mask_sum(k, g.mz));
            double F[27];
#pragma unroll
            for (int n = 0; n < 27; n++) F[n] = double(q[n]) * rden;
            double L[3][3], Sh[3][3], SSh[3][3];
            L[0][0] = F[6] - F[3] * F[0]; L[0][1] = F[7] - F[3] * F[1]; L[0][2] = F[8] - F[3] * F[2];
            L[1][0] = F[9] - F[4] * F[0]; L[1][1] = F[10] - F[4] * F[1]; L[1][2] = F[11] - F[4] * F[2];
            L[2][0] = F[12] - F[5] * F[0]; L[2][1] = F[13] - F[5] * F[1]; L[2][2] = F[14] - F[5] * F[2];
            Sh[0][0] = F[15]; Sh[0][1] = Sh[1][0] = F[16]; Sh[0][2] = Sh[2][0] = F[17];
            Sh[1][1] = F[18]; Sh[1][2] = Sh[2][1] = F[19]; Sh[2][2] = F[20];
            SSh[0][0] = F[21]; SSh[0][1] = SSh[1][0] = F[22]; SSh[0][2] = SSh[2][0] = F[23];
            SSh[1][1] = F[24]; SSh[1][2] = SSh[2][1] = F[25]; SSh[2][2] = F[26];
            if constexpr (!shifted) { cP = ld3(csi, c); eP = ld3(eta, c); zP = ld3(zet, c); aP = aj[c]; }
            const double filter = cube_root(1. / aP);
            const double test_filter = cube_root(8.0 * double(q[27]));
            double nm, dn;
            germano(L, Sh, SSh, filter, test_filter, cP, eP, zP, nm, dn);
            LM[c] = nm;
            MM[c] = dn;
            Sabs[c] = sabsPrev;
        }
        sabsPrev = sabs;
        if constexpr (shifted) { cP = cC; eP = eC; zP = zC; aP = aC; }
        uKm = uK;
        uK = uKp;
        // the next plane writes sH only after its sU barrier, which every thread reaches after its j
        // filter reads of this plane
    }
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

// Box average, marching along k: a workgroup of CX x CY threads owns a CX x CY column of output
// cells over a chunk of k planes.  Per plane the weights 1/aj and the remapped LM w, MM w of the
// tile plus a one-cell i/j halo are staged once per point in LDS (double-buffered); each thread
// forms the 3 x 3 Simpson-weighted sums of its point in the plane and keeps those of the last
// three planes in registers, so the k direction costs no reloads.  The products are the
// reference's (the weights are powers of two, so (wR wQ wP) w L = wR ((wQ wP) w) L exactly); the
// 27 terms are summed plane by plane.  The per-point rules do not depend on the centre cell,
// because the centre index j of les.c:1911 never equals my-1.
constexpr int CX = 64, CY = 8, HX = CX + 2, HY = CY + 2, HP = HX * HY;

template <class T>
__global__ void __launch_bounds__(CX * CY) k_dsm_cs_nut(Geo g, const double *LM, const double *MM, const double *Sabs,
                                                        const double *aj, double maxCs, double *Cs, T *nut, int kchunk) {
    __shared__ double W[2][HP], Lw[2][HP], Mw[2][HP];
    const int tid = threadIdx.x + CX * threadIdx.y;
    const int tile = xcd_tile();     // i fastest, then j, then the k chunk
    const int bi = tile % gridDim.x, bj = (tile / gridDim.x) % gridDim.y, bk = tile / (gridDim.x * gridDim.y);
    const int i0 = 1 + bi * CX, j0 = 1 + bj * CY;
    const int kb = 1 + bk * kchunk, ke = min(kb + kchunk, g.mz - 1);     // output planes [kb, ke)
    const int i = i0 + threadIdx.x, j = j0 + threadIdx.y;
    const bool own = i <= g.mx - 2 && j <= g.my - 2;
    double pw[2] = {0, 0}, pl[2] = {0, 0}, pm[2] = {0, 0};     // 2-D sums of planes K-2, K-1
    int b = 0;
    for (int K = kb - 1; K <= ke; K++) {
        for (int idx = tid; idx < HP; idx += CX * CY) {
            int I = i0 - 1 + idx % HX, J = j0 - 1 + idx / HX, KK = K;
            double w = 0.0, l = 0.0, mm = 0.0;
            if (I <= g.mx - 1 && J <= g.my - 1) {     // points past the owned range feed no output
                w = 1. / aj[g.L(KK, J, I)];
                bool ghost = false;
                bool ok = remap_axis(I, 1, g.mx, g.per[0], false, ghost);
                ok = remap_axis(J, 1, g.my, g.per[1], true, ghost) && ok;
                ok = remap_axis(KK, 1, g.mz, g.per[2], false, ghost) && ok;
                if (!ok) w = 0;
                if (!ghost && w != 0) { const int c = g.L(KK, J, I); l = LM[c]; mm = MM[c]; }
            }
            W[b][idx] = w;
            Lw[b][idx] = l;
            Mw[b][idx] = mm;
        }
        __syncthreads();
        // 3 x 3 sums of this plane at (i, j), Q (j) outer, P (i) inner, as the reference's loops
        double sw = 0, sl = 0, sm = 0;
        for (int Q = 0; Q < 3; Q++)
            for (int P = 0; P < 3; P++) {
                const int n = (threadIdx.y + Q) * HX + (threadIdx.x + P);
                const double wq = (Q == 1 ? 4. : 1.) * (P == 1 ? 4. : 1.);
                const double w = W[b][n];
                sw += wq * w;
                sl += wq * w * Lw[b][n];
                sm += wq * w * Mw[b][n];
            }
        const int k = K - 1;
        if (own && k >= kb && k < ke) {
            const double wsum = (pw[0] + 4. * pw[1]) + sw, lsum = (pl[0] + 4. * pl[1]) + sl,
                         msum = (pm[0] + 4. * pm[1]) + sm;
            const double C = 0.5 * (lsum / wsum) / ((msum / wsum) + 1e-10);
            const int c = g.L(k, j, i);
            const double cs = pmin(pmax(C, 0.0), maxCs);
            Cs[c] = cs;
            if (nut) {
                const double filter = cube_root(1. / aj[c]);
                nut[c] = T(cs * (filter * filter) * Sabs[c]);
            }
        }
        pw[0] = pw[1]; pw[1] = sw;
        pl[0] = pl[1]; pl[1] = sl;
        pm[0] = pm[1]; pm[1] = sm;
        b ^= 1;       // the next plane stages into the other buffer; the barrier above orders the reuse
    }
}

template <class T>
__global__ void __launch_bounds__(256) k_nut_fused(Geo g, const T *ucat, const double *csi, const double *eta, const double *zet,
                            const double *aj, const double *Cs, T *nut, T *ksgs) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1, j = blockIdx.y * blockDim.y + threadIdx.y + 1,
              k = blockIdx.z + 1;
    if (i > g.mx - 2 || j > g.my - 2 || k > g.mz - 2) return;
    const int c = g.L(k, j, i);
    const V3 cc = ld3(csi, c), ee = ld3(eta, c), zz = ld3(zet, c);
    const double a = aj[c];
    const double S = strain_mag(du_dxyz(cvt3<T>(cc), cvt3<T>(ee), cvt3<T>(zz), T(a), du_center(g, ucat, i, j, k)));
    const double filter = cube_root(1. / a);
    const double nt = Cs[c] * (filter * filter) * S;      // pow(x, 2.) of les.c:2510 is x * x, exactly rounded
    nut[c] = T(nt);
    if (ksgs) {
        const double V = 1.0 / a;
        const double dx = V / sqrt(zz.x * zz.x + zz.y * zz.y + zz.z * zz.z);
        const double dy = V / sqrt(cc.x * cc.x + cc.y * cc.y + cc.z * cc.z);
        const double dz = V / sqrt(ee.x * ee.x + ee.y * ee.y + ee.z * ee.z);
        const double d2 = 3.0 / (1.0 / (dx * dx) + 1.0 / (dy * dy) + 1.0 / (dz * dz));
        ksgs[c] = T((nt * nt) / ((0.094 * 0.094) * d2));
    }
}

// zero six disjoint boxes of an array with row length nx and plane size nx*ny (ncomp values per
// point) in one launch: blockIdx.y selects the box, the blocks stride over its values
struct ShellBoxes { int i0[6], j0[6], k0[6], ni[6], nj[6], nk[6]; };

template <class T>
__global__ void k_zero_box(T *a, int ncomp, int nx, int ny, ShellBoxes B) {
    const int b = blockIdx.y;
    if (B.ni[b] <= 0 || B.nj[b] <= 0 || B.nk[b] <= 0) return;
    const long row = (long)B.ni[b] * ncomp, tot = row * B.nj[b] * B.nk[b];
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < tot; t += (long)gridDim.x * blockDim.x) {
        const long e = t % row, r = t / row;
        const int jj = (int)(r % B.nj[b]), kk = (int)(r / B.nj[b]);
        const long p = ((long)(B.k0[b] + kk) * ny + (B.j0[b] + jj)) * nx + B.i0[b];
        a[p * ncomp + e] = T(0);
    }
}

template <class T>
void zero_shell(T *a, int ncomp, int nx, int ny, int nz, int lo, int hi, hipStream_t s) {
    // boxes (in array coordinates) covering everything but [lo+1 .. n-hi-2] on every axis:
    // `lo` / `hi` extra layers outside the boundary cells on the low / high side
    const int w0 = lo + 1, w1 = hi + 1;            // shell thickness per side
    ShellBoxes B{};
    auto box = [&](int q, int i0, int j0, int k0, int ni, int nj, int nk) {
        B.i0[q] = i0; B.j0[q] = j0; B.k0[q] = k0; B.ni[q] = ni; B.nj[q] = nj; B.nk[q] = nk;
    };
    box(0, 0, 0, 0, nx, ny, w0);                                   // low k slab
    box(1, 0, 0, nz - w1, nx, ny, w1);                             // high k slab
    box(2, 0, 0, w0, nx, w0, nz - w0 - w1);                        // low j slab
    box(3, 0, ny - w1, w0, nx, w1, nz - w0 - w1);                  // high j slab
    box(4, 0, w0, w0, w0, ny - w0 - w1, nz - w0 - w1);             // low i slab
    box(5, nx - w1, w0, w0, w1, ny - w0 - w1, nz - w0 - w1);       // high i slab
    TOSCA_LAUNCH("k_zero_box", k_zero_box<T>, dim3(512, 6), dim3(256), 0, s, a, ncomp, nx, ny, B);
}

// y = y + 1 x over the same six boxes (three values per point)
__global__ void k_add_box(double *y, const double *x, int nx, int ny, ShellBoxes B) {
    const int b = blockIdx.y;
    if (B.ni[b] <= 0 || B.nj[b] <= 0 || B.nk[b] <= 0) return;
    const long row = (long)B.ni[b] * 3, tot = row * B.nj[b] * B.nk[b];
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < tot; t += (long)gridDim.x * blockDim.x) {
        const long e = t % row, r = t / row;
        const int jj = (int)(r % B.nj[b]), kk = (int)(r / B.nj[b]);
        const long p = ((long)(B.k0[b] + kk) * ny + (B.j0[b] + jj)) * nx + B.i0[b];
        y[p * 3 + e] = y[p * 3 + e] + 1.0 * x[p * 3 + e];
    }
}

void shell_add(double *y, const double *x, int nx, int ny, int nz, hipStream_t s) {
    ShellBoxes B{};
    auto box = [&](int q, int i0, int j0, int k0, int ni, int nj, int nk) {
        B.i0[q] = i0; B.j0[q] = j0; B.k0[q] = k0; B.ni[q] = ni; B.nj[q] = nj; B.nk[q] = nk;
    };
    box(0, 0, 0, 0, nx, ny, 1);
    box(1, 0, 0, nz - 1, nx, ny, 1);
    box(2, 0, 0, 1, nx, 1, nz - 2);
    box(3, 0, ny - 1, 1, nx, 1, nz - 2);
    box(4, 0, 1, 1, 1, ny - 2, nz - 2);
    box(5, nx - 1, 1, 1, 1, ny - 2, nz - 2);
    TOSCA_LAUNCH("k_add_box", k_add_box, dim3(256, 6), dim3(256), 0, s, y, x, nx, ny, B);
}

template <class T>
void zero_outside_interior(T *a, const Geo &g, hipStream_t s) {
    zero_shell(a, 1, g.gxm, g.gym, g.gzm, GW, GW, s);
}

template <class T>
void fused_dsm(const Geo &g, const tosca_mesh *m, const T *ucat, double maxCs, double *cs, T *nut, double *ws,
               hipStream_t s) {
    const long n = g.nloc();
    double *LM = ws, *MM = ws + n, *Sabs = ws + 2 * n;
    const int nk = g.mz - 2;
    const int tiles = ((g.mx - 2 + DSM_TI - 1) / DSM_TI) * ((g.my - 2 + DSM_TJ - 1) / DSM_TJ);
    int chunks = (2048 + tiles - 1) / tiles;
    chunks = max(1, min(chunks, nk / 8 > 0 ? nk / 8 : 1));
    const int kchunk = (nk + chunks - 1) / chunks;
    chunks = (nk + kchunk - 1) / kchunk;
    dim3 grid((g.mx - 2 + DSM_TI - 1) / DSM_TI, (g.my - 2 + DSM_TJ - 1) / DSM_TJ, chunks);
    TOSCA_LAUNCH(sizeof(T) == 8 ? "k_dsm_lm" : "k_dsm_lm_f32", k_dsm_lm<T>, grid, dim3(64, DSM_NR), 0, s, g, ucat, m->csi, m->eta, m->zet, m->aj, LM, MM,
                 Sabs, kchunk);
    const int ctiles = ((g.mx - 2 + CX - 1) / CX) * ((g.my - 2 + CY - 1) / CY);
    int cchunks = max(1, min((2304 + ctiles - 1) / ctiles, nk / 8 > 0 ? nk / 8 : 1));
    const int ckchunk = (nk + cchunks - 1) / cchunks;
    cchunks = (nk + ckchunk - 1) / ckchunk;
    TOSCA_LAUNCH(sizeof(T) == 8 ? "k_dsm_cs_nut" : "k_dsm_cs_nut_f32", k_dsm_cs_nut<T>,
                 dim3((g.mx - 2 + CX - 1) / CX, (g.my - 2 + CY - 1) / CY, cchunks), dim3(CX, CY), 0, s, g, LM, MM, Sabs,
                 m->aj, maxCs, cs, nut, ckchunk);
}

template <class T>
void fused_nut(const Geo &g, const tosca_mesh *m, const T *ucat, const double *cs, T *nut, T *ksgs, hipStream_t s) {
    TOSCA_LAUNCH(sizeof(T) == 8 ? "k_nut_fused" : "k_nut_fused_f32", k_nut_fused<T>, dim3((g.mx - 2 + 63) / 64, (g.my - 2 + 3) / 4, g.mz - 2), dim3(64, 4), 0,
                 s, g, ucat, m->csi, m->eta, m->zet, m->aj, cs, nut, ksgs);
}

template <class T>
void nut_bcs(T *nut, const tosca_params *p, const Geo &g, hipStream_t s) { launch_nut_bcs(nut, p, g, s); }

#define TOSCA_LES_INSTANTIATE(T)                                                                            \
    template void zero_shell<T>(T *, int, int, int, int, int, int, hipStream_t);                            \
    template void zero_outside_interior<T>(T *, const Geo &, hipStream_t);                                  \
    template void fused_dsm<T>(const Geo &, const tosca_mesh *, const T *, double, double *, T *, double *,   \
                               hipStream_t);                                                                 \
    template void fused_nut<T>(const Geo &, const tosca_mesh *, const T *, const double *, T *, T *,         \
                               hipStream_t);                                                                 \
    template void nut_bcs<T>(T *, const tosca_params *, const Geo &, hipStream_t);
TOSCA_LES_INSTANTIATE(double)
TOSCA_LES_INSTANTIATE(float)

}  // namespace tosca
