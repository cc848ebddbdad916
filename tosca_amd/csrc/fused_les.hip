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

// The base values of the test-filter quantities at one point (les.c:466-611): u v w, U V W (the
// contravariant velocity u.csi, u.eta, u.zet), Sxx Sxy Sxz Syy Syz Szz, |S|, 1/aj.  The velocity
// arithmetic (strain, |S|, U) is in the flow type T; the shift (uref, Uref) is subtracted from
// u and U when the sums are FP32.
template <class S, class T, class MT>
__device__ __forceinline__ void filter_base(const Geo &g, int i, int j, int K, Vec3<T> uK, Vec3<T> uim, Vec3<T> uip,
                                            Vec3<T> ujm, Vec3<T> ujp, Vec3<T> uKm, Vec3<T> uKp, Vec3<MT> cC, Vec3<MT> eC,
                                            Vec3<MT> zC, double aC, Vec3<T> uref, Vec3<T> Uref, S b[14]) {
    const Vec3<T> cT = cvt3<T>(cC), eT = cvt3<T>(eC), zT = cvt3<T>(zC);
    const CartT<T> d = du_dxyz(cT, eT, zT, T(aC), du_center_v(g, i, j, K, uK, uim, uip, ujm, ujp, uKm, uKp));
    const T h = T(0.5);
    b[0] = uK.x; b[1] = uK.y; b[2] = uK.z;
    b[3] = uK.x * cT.x + uK.y * cT.y + uK.z * cT.z;
    b[4] = uK.x * eT.x + uK.y * eT.y + uK.z * eT.z;
    b[5] = uK.x * zT.x + uK.y * zT.y + uK.z * zT.z;
    if constexpr (!std::is_same<S, double>::value) {
        b[0] -= uref.x; b[1] -= uref.y; b[2] -= uref.z;
        b[3] -= Uref.x; b[4] -= Uref.y; b[5] -= Uref.z;
    }
    b[6] = h * (d.du_dx + d.du_dx); b[7] = h * (d.du_dy + d.dv_dx); b[8] = h * (d.du_dz + d.dw_dx);
    b[9] = h * (d.dv_dy + d.dv_dy); b[10] = h * (d.dv_dz + d.dw_dy); b[11] = h * (d.dw_dz + d.dw_dz);
    b[12] = strain_mag(d);
    b[13] = S(1. / aC);
}

// Filtered quantity n of a point from its base values: u_a (0-2), U_a (3-5), U_a u_b (6-14),
// S_ij (15-20), S_ij |S| (21-26), 1/aj (27)
template <class S>
__device__ __forceinline__ S filter_quantity(int n, const S b[14]) {
    if (n < 6) return b[n];
    if (n < 15) return b[3 + (n - 6) / 3] * b[(n - 6) % 3];
    if (n < 21) return b[6 + (n - 15)];
    if (n < 27) return b[6 + (n - 21)] * b[12];
    return b[13];
}

// The reference-velocity shift of a tile: the velocity and the contravariant velocity of cell cr
template <class T, class MT>
__device__ __forceinline__ void tile_shift(const T *ucat, const MT *csi, const MT *eta, const MT *zet, int cr,
                                           Vec3<T> &uref, Vec3<T> &Uref) {
    uref = ld3(ucat, cr);
    const Vec3<T> cr3 = cvt3<T>(ld3(csi, cr)), er3 = cvt3<T>(ld3(eta, cr)), zr3 = cvt3<T>(ld3(zet, cr));
    Uref = Vec3<T>{uref.x * cr3.x + uref.y * cr3.y + uref.z * cr3.z, uref.x * er3.x + uref.y * er3.y + uref.z * er3.z,
                   uref.x * zr3.x + uref.y * zr3.y + uref.z * zr3.z};
}

// Type of the filtered quantities and their sums: FP64 with FP64 flow fields.  With FP32 flow
// fields they are FP32, after the tile's reference velocity is taken out of both factors of the
// Leonard term (below), so that L = <U u> - <U><u> does not lose its digits to cancellation.
template <class T> struct DsmSum { using type = double; };
template <> struct DsmSum<float> { using type = float; };

// MT: storage type of csi, eta, zet (float: the FP32 metric copies of the mixed step that runs
// FormU on them; with FP32 flow fields the strain uses FP32 metrics either way)
template <class T, class MT = double>
__global__ void __launch_bounds__(64 * DSM_NR) k_dsm_lm(Geo g, const T *ucat, const MT *csi,
                                                        const MT *eta, const MT *zet, const double *aj,
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
        tile_shift(ucat, csi, eta, zet, g.L(kc, jc, ic), uref, Uref);
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
    Vec3<MT> cP{0, 0, 0}, eP{0, 0, 0}, zP{0, 0, 0};
    double aP = 1.0;

    // 1 / (the mask weight of the output cell's 3 x 3 x 3 neighbourhood), for each k weight sum
    const double rd4 = 1.0 / (dij * 4.0), rd5 = 1.0 / (dij * 5.0), rd6 = 1.0 / (dij * 6.0);
    // partial sums along k: A for output K-1 (terms K-2, K-1 in), B for output K (term K-1 in).
    // (Taking the planes in pairs with A and B swapping roles saves the copy of B into A, but the
    // FP64 instance then spills: 1.5x slower; the FP32 one gains 5 %.)
    S A[NQ], B[NQ];
#pragma unroll
    for (int n = 0; n < NQ; n++) { A[n] = S(0); B[n] = S(0); }
    double sabsPrev = 0.0;

    // FP32 sums: the next plane's metrics and ucat are loaded one plane ahead (the ghost planes
    // keep K + 2 <= mz + 1 inside the local array)
    Vec3<T> uN{T(0), T(0), T(0)};
    Vec3<MT> cN{0, 0, 0}, eN{0, 0, 0}, zN{0, 0, 0};
    double aN = 1.0;
    if constexpr (shifted) {
        const int c1 = c0 + (kb - 1) * SK;
        uN = ld3(ucat, c1 + SK); cN = ld3(csi, c1); eN = ld3(eta, c1); zN = ld3(zet, c1); aN = aj[c1];
    }
    for (int K = kb - 1; K <= ke; K++) {
        // loads of this plane, issued before the exchange barrier (with FP64 sums a register
        // prefetch of the next plane spills at this register budget and measured slower)
        const int cK = opaque_index(c0 + K * SK);
        Vec3<T> uKp;
        Vec3<MT> cC, eC, zC;
        double aC;
        if constexpr (shifted) {
            uKp = uN; cC = cN; eC = eN; zC = zN; aC = aN;
        } else {
            uKp = ld3(ucat, cK + SK);
            cC = ld3(csi, cK); eC = ld3(eta, cK); zC = ld3(zet, cK);
            aC = aj[cK];
        }
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
        if constexpr (shifted) {
            // issued after this plane's own loads, so that waiting for those does not wait for these
            const int cn = opaque_index(cK + SK);
            uN = ld3(ucat, cn + SK); cN = ld3(csi, cn); eN = ld3(eta, cn); zN = ld3(zet, cn); aN = aj[cn];
        }
        // ---- point quantities (les.c:466-611), zero outside the mask -------------------------------
        // base values b: u v w, U V W, Sxx Sxy Sxz Syy Syz Szz, |S|, 1/aj; the 28 filtered quantities
        // are formed from them one at a time, right before their i filter
        S b[14];
#pragma unroll
        for (int n = 0; n < 14; n++) b[n] = S(0);
        if (pt) filter_base<S>(g, i, j, K, uK, uim, uip, ujm, ujp, uKm, uKp, cC, eC, zC, aC, uref, Uref, b);
        const double sabs = b[12];
        // ---- filter along i (lanes), then j (LDS) ---------------------------------------------------
#pragma unroll
        for (int n = 0; n < NQ; n++) {
            const S v = filter_quantity(n, b);
            const S lo = shup(v), hi = shdn(v);
            sH[n][r][lane] = (n < 27) ? (lo + S(4.0) * v + hi) : (S(0.25) * lo + S(0.5) * v + S(0.25) * hi);
        }
        __syncthreads();
        const int rm = r - 1, rp = r + 1;
        // ---- filter along k, and the finished output cell K-1 --------------------------------------
        // (the halo rows 0 and DSM_NR-1, whole waves, have no output and skip this part)
        if (r >= 1 && r <= DSM_NR - 2) {
        const bool emit = outCol && K - 1 >= kb && K - 1 < ke;
        S q[NQ];
#pragma unroll
        for (int n = 0; n < NQ; n++) {
            const S f = (n < 27) ? (sH[n][rm][lane] + S(4.0) * sH[n][r][lane] + sH[n][rp][lane])
                                 : (S(0.25) * sH[n][rm][lane] + S(0.5) * sH[n][r][lane] + S(0.25) * sH[n][rp][lane]);
            const S wK = (n < 27) ? S(1.0) : S(0.25), wC = (n < 27) ? S(4.0) : S(0.5);
            q[n] = A[n] + wK * f;                    // output K-1 complete
            A[n] = B[n] + wC * f;
            B[n] = wK * f;
        }
        if (emit) {
            const int k = K - 1;
            const long c = g.L(k, j, i);
            const double ks = simpson_mask_sum(k, g.mz);
            const double rden = ks == 6.0 ? rd6 : ks == 5.0 ? rd5 : rd4;
            // The filtered values are F = q rden.  L = rden L' with L'_ab = q_ab - (rden q_a) q_b, and
            // Sh, SSh = rden (qS, qSS) make M = rden M' where M' is formed from (qS, qSS) with
            // Delta~^2 rden in place of Delta~^2; so LM = rden^2 L'M'G and MM = rden^2 M'M'G.
            double F[27];
#pragma unroll
            for (int n = 0; n < 27; n++) F[n] = double(q[n]);
            const double r3 = rden * F[3], r4 = rden * F[4], r5 = rden * F[5];
            double L[3][3], Sh[3][3], SSh[3][3];
            L[0][0] = F[6] - r3 * F[0]; L[0][1] = F[7] - r3 * F[1]; L[0][2] = F[8] - r3 * F[2];
            L[1][0] = F[9] - r4 * F[0]; L[1][1] = F[10] - r4 * F[1]; L[1][2] = F[11] - r4 * F[2];
            L[2][0] = F[12] - r5 * F[0]; L[2][1] = F[13] - r5 * F[1]; L[2][2] = F[14] - r5 * F[2];
            Sh[0][0] = F[15]; Sh[0][1] = Sh[1][0] = F[16]; Sh[0][2] = Sh[2][0] = F[17];
            Sh[1][1] = F[18]; Sh[1][2] = Sh[2][1] = F[19]; Sh[2][2] = F[20];
            SSh[0][0] = F[21]; SSh[0][1] = SSh[1][0] = F[22]; SSh[0][2] = SSh[2][0] = F[23];
            SSh[1][1] = F[24]; SSh[1][2] = SSh[2][1] = F[25]; SSh[2][2] = F[26];
            if constexpr (!shifted) { cP = ld3(csi, c); eP = ld3(eta, c); zP = ld3(zet, c); aP = aj[c]; }
            // Delta^2 = (1/aj)^(2/3) = r(aj)^2 and Delta~^2 = (8 q)^(2/3) = 4 q r(q), r(x) = x^(-1/3)
            const double rf = rcbrt_pos(aP), q27 = double(q[27]);
            double nm, dn;
            germano_sq(L, Sh, SSh, rf * rf, 4.0 * q27 * rcbrt_pos(q27) * rden, cvt3<double>(cP), cvt3<double>(eP),
                       cvt3<double>(zP), nm, dn);
            const double rd2 = rden * rden;
            LM[c] = nm * rd2;
            MM[c] = dn * rd2;
            Sabs[c] = sabsPrev;
        }
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
    // The staged values of the next plane are loaded into registers while this plane's sums are
    // formed (raw loads: aj at the point, LM / MM at its remapped point; flags: in range, remap ok,
    // ghost), and written to the other LDS buffer at the start of the next step.
    constexpr int NPT = (HP + CX * CY - 1) / (CX * CY);
    double ra[NPT], rl[NPT], rm[NPT];
    int rf[NPT];
    double ajN = 1.0, sabsN = 0.0;                // the output cell of the step fetch(K) serves, own threads
    auto fetch = [&](int K) {
#pragma unroll
        for (int u = 0; u < NPT; u++) {
            const int idx = tid + u * CX * CY;
            ra[u] = 1.0; rl[u] = 0.0; rm[u] = 0.0; rf[u] = 0;
            if (idx < HP) {
                int I = i0 - 1 + idx % HX, J = j0 - 1 + idx / HX, KK = K;
                if (I <= g.mx - 1 && J <= g.my - 1) {     // points past the owned range feed no output
                    ra[u] = aj[g.L(KK, J, I)];
                    bool ghost = false;
                    bool ok = remap_axis(I, 1, g.mx, g.per[0], false, ghost);
                    ok = remap_axis(J, 1, g.my, g.per[1], true, ghost) && ok;
                    ok = remap_axis(KK, 1, g.mz, g.per[2], false, ghost) && ok;
                    rf[u] = 1 | (ok ? 2 : 0) | (ghost ? 4 : 0);
                    if (!ghost && ok) { const int c = g.L(KK, J, I); rl[u] = LM[c]; rm[u] = MM[c]; }
                }
            }
        }
        // the cell the step that stages plane K emits: k = K-1
        if (own && K - 1 >= kb && K - 1 < ke) { const int c = g.L(K - 1, j, i); ajN = aj[c]; sabsN = Sabs[c]; }
    };
    int b = 0;
    fetch(kb - 1);
    for (int K = kb - 1; K <= ke; K++) {
#pragma unroll
        for (int u = 0; u < NPT; u++) {
            const int idx = tid + u * CX * CY;
            if (idx < HP) {
                double w = 0.0, l = 0.0, mm = 0.0;
                if (rf[u] & 1) {
                    w = 1. / ra[u];
                    if (!(rf[u] & 2)) w = 0;
                    if (!(rf[u] & 4) && w != 0) { l = rl[u]; mm = rm[u]; }
                }
                W[b][idx] = w;
                Lw[b][idx] = l;
                Mw[b][idx] = mm;
            }
        }
        const double ajC = ajN, sabsC = sabsN;     // the output cell of this step (k = K-1)
        __syncthreads();
        if (K + 1 <= ke) fetch(K + 1);
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
                const double filter = cube_root(1. / ajC);
                nut[c] = T(cs * (filter * filter) * sabsC);
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

// ---- plane-averaged dynamic model (DPASD, les.c:812-1174 and 1979-2048) ----------------------------
// The second test filter weighs the 5 x 5 x 5 neighbourhood of a cell with the products of the 1-D
// Simpson rows (1, 4, 2, 4, 1) (integrateTestfilterSimpson5x5, inline.h:3612-3633), over its
// interior cells only (I, J, K in [1, m-2] whatever the periodicity), normalised by the weights of
// those cells; its width is (64 sum bin5(R) bin5(Q) bin5(P) / aj)^(1/3) over the same cells, with
// bin5 = (1, 4, 6, 4, 1)/16 (les.c:961-1030).  Both weights and mask are products of 1-D factors,
// so the filter is applied one axis at a time:
//   k_dpasd_ij: a 1024-thread workgroup owns 60 x 12 (i, j) cells plus two halo lanes / rows a side
//     and marches along k.  Per plane each thread forms the 28 quantities of its point (those of
//     the 3-point filter), filters them along i across the lanes of the wave (two DPP shifts each
//     way) and along j through LDS, seven quantities per LDS round, and writes the (i, j)-filtered
//     sums of the tile's cells to HBM (28 arrays over the interior cells, i fastest);
//   k_dpasd_k: one thread per interior column marches over a chunk of planes: the k filter of the
//     five planes (re-read from L2), the Germano contraction (QN, NN), and the cell's LM, MM of the
//     3-point filter; the 64 cells of a wave (one j row) are summed by a cross-lane butterfly into a
//     partial per (j, i tile, k chunk);
//   k_dpasd_plane: the partials of each j plane in a fixed order -> the plane means -> Cs(j);
//   k_dpasd_cs_nut: Cs and nu_t = Cs Delta^2 |S| of every interior cell.
// Every sum has a fixed order, so Cs is bitwise reproducible from run to run.
namespace {
constexpr int DP_NR = 16;                       // rows of a k_dpasd_ij workgroup
constexpr int DP_TI = 60, DP_TJ = DP_NR - 4;    // its output cells
constexpr int DP_NC = 7;                        // quantities per LDS round of the j filter
constexpr int DK_X = 64, DK_Y = 4;              // k_dpasd_k workgroup: one wave per j row

// sum of the 1-D weights (1, 4, 2, 4, 1) over the interior points c-2 .. c+2 of an axis of size m
__device__ __forceinline__ double simpson5_mask_sum(int c, int m) {
    return ((c - 2 >= 1) ? 1.0 : 0.0) + ((c - 1 >= 1) ? 4.0 : 0.0) + 2.0 + ((c + 1 <= m - 2) ? 4.0 : 0.0) +
           ((c + 2 <= m - 2) ? 1.0 : 0.0);
}

// 5-point row of quantity n over the values at offsets -2 .. 2: Simpson weights for the 27 Germano
// quantities, the binomial (1, 4, 6, 4, 1)/16 for the volume term of the filter width
template <class S>
__device__ __forceinline__ S row5(int n, S m2, S m1, S c, S p1, S p2) {
    if (n < 27) return ((m2 + p2) + S(4.0) * (m1 + p1)) + S(2.0) * c;
    return (((m2 + p2) + S(4.0) * (m1 + p1)) + S(6.0) * c) * S(0.0625);
}
}  // namespace

template <class T>
__global__ void __launch_bounds__(64 * DP_NR) k_dpasd_ij(Geo g, const T *ucat, const double *csi, const double *eta,
                                                         const double *zet, const double *aj,
                                                         typename DsmSum<T>::type *X, long nI, int kchunk) {
    using S = typename DsmSum<T>::type;
    constexpr bool shifted = !std::is_same<S, double>::value;
    __shared__ T sU[3][DP_NR][64];            // ucat of the plane, for the j neighbours of the strain
    __shared__ S sH[DP_NC][DP_NR][64];        // i-filtered quantities of one LDS round
    const int lane = threadIdx.x, r = threadIdx.y;
    const int tile = xcd_tile();              // j fastest, then i, then the k chunk
    const int bj = tile % gridDim.y, bi = (tile / gridDim.y) % gridDim.x, bk = tile / (gridDim.x * gridDim.y);
    // lanes 2 .. 61 and rows 2 .. DP_NR-3 are the tile's cells, the others its halo
    const int i = bi * DP_TI + lane - 1, j = bj * DP_TJ + r - 1;
    const int kb = 1 + bk * kchunk, ke = min(kb + kchunk, g.mz - 1);     // planes [kb, ke)
    // with FP32 sums every point of the tile takes the velocity and contravariant velocity of one
    // interior cell out of the two factors of the Leonard term; the cell does not depend on the k
    // chunk, so the five planes that k_dpasd_k combines for a cell were filtered with one shift
    Vec3<T> uref{T(0), T(0), T(0)}, Uref{T(0), T(0), T(0)};
    if constexpr (shifted) {
        const int ic = min(max(bi * DP_TI + 30, 1), g.mx - 2), jc = min(max(bj * DP_TJ + 6, 1), g.my - 2);
        tile_shift(ucat, csi, eta, zet, g.L((g.mz - 1) / 2, jc, ic), uref, Uref);
    }
    const bool mi = i >= 1 && i <= g.mx - 2, mj = j >= 1 && j <= g.my - 2;
    const bool outCol = lane >= 2 && lane <= DP_TI + 1 && r >= 2 && r <= DP_TJ + 1 && mi && mj;
    // i >= -1 and j >= -1 here; addresses of columns past the array are clamped, those points are
    // outside the filter mask
    const int ia = min(i, g.mx + 2), ja = min(j, g.my + 2);
    const int ima = max(ia - 1, -GW), ipa = min(ia + 1, g.mx + 2), jma = max(ja - 1, -GW), jpa = min(ja + 1, g.my + 2);
    const int SK = g.gxm * g.gym;
    const int c0 = g.L(0, ja, ia);
    const int ni = g.mx - 2, nj = g.my - 2;
    const int x0 = (j - 1) * ni + (i - 1);    // interior index of the column at plane 1 (output cells)

    Vec3<T> uKm = ld3(ucat, c0 + (kb - 1) * SK), uK = ld3(ucat, c0 + kb * SK);
    for (int K = kb; K < ke; K++) {
        const int cK = opaque_index(c0 + K * SK);
        const Vec3<T> uKp = ld3(ucat, cK + SK);
        const V3 cC = ld3(csi, cK), eC = ld3(eta, cK), zC = ld3(zet, cK);
        const double aC = aj[cK];
        sU[0][r][lane] = uK.x; sU[1][r][lane] = uK.y; sU[2][r][lane] = uK.z;
        Vec3<T> uim = shup3(uK), uip = shdn3(uK);
        if (lane == 0) uim = ld3(ucat, g.L(K, ja, ima));
        if (lane == 63) uip = ld3(ucat, g.L(K, ja, ipa));
        __syncthreads();
        Vec3<T> ujm, ujp;
        if (r > 0) ujm = Vec3<T>{sU[0][r - 1][lane], sU[1][r - 1][lane], sU[2][r - 1][lane]};
        else ujm = ld3(ucat, g.L(K, jma, ia));
        if (r < DP_NR - 1) ujp = Vec3<T>{sU[0][r + 1][lane], sU[1][r + 1][lane], sU[2][r + 1][lane]};
        else ujp = ld3(ucat, g.L(K, jpa, ia));
        S b[14];
#pragma unroll
        for (int n = 0; n < 14; n++) b[n] = S(0);
        if (mi && mj) filter_base<S>(g, i, j, K, uK, uim, uip, ujm, ujp, uKm, uKp, cC, eC, zC, aC, uref, Uref, b);
        const long xo = (long)(K - 1) * ni * nj + x0;
#pragma unroll
        for (int c = 0; c < NQ; c += DP_NC) {
#pragma unroll
            for (int q = 0; q < DP_NC; q++) {
                const S v = filter_quantity(c + q, b);
                const S l1 = shup(v), h1 = shdn(v);
                const S l2 = shup(l1), h2 = shdn(h1);
                sH[q][r][lane] = row5(c + q, l2, l1, v, h1, h2);
            }
            __syncthreads();
            if (outCol) {
#pragma unroll
                for (int q = 0; q < DP_NC; q++)
                    X[(long)(c + q) * nI + xo] = row5(c + q, sH[q][r - 2][lane], sH[q][r - 1][lane], sH[q][r][lane],
                                                      sH[q][r + 1][lane], sH[q][r + 2][lane]);
            }
            // the next round overwrites sH; the last round's reads are ordered before the next
            // plane's sH writes by that plane's sU barrier
            if (c + DP_NC < NQ) __syncthreads();
        }
        uKm = uK;
        uK = uKp;
    }
}

// QN, NN of cell c from its 28 k-filtered sums F (F[0..26] not yet divided by the mask weight)
__device__ __forceinline__ void dpasd_germano(double (&F)[NQ], double rden, const double *csi, const double *eta,
                                              const double *zet, const double *aj, int c, double &qn, double &nn) {
#pragma unroll
    for (int n = 0; n < 27; n++) F[n] *= rden;
    double L[3][3], Sh[3][3], SSh[3][3];
    L[0][0] = F[6] - F[3] * F[0]; L[0][1] = F[7] - F[3] * F[1]; L[0][2] = F[8] - F[3] * F[2];
    L[1][0] = F[9] - F[4] * F[0]; L[1][1] = F[10] - F[4] * F[1]; L[1][2] = F[11] - F[4] * F[2];
    L[2][0] = F[12] - F[5] * F[0]; L[2][1] = F[13] - F[5] * F[1]; L[2][2] = F[14] - F[5] * F[2];
    Sh[0][0] = F[15]; Sh[0][1] = Sh[1][0] = F[16]; Sh[0][2] = Sh[2][0] = F[17];
    Sh[1][1] = F[18]; Sh[1][2] = Sh[2][1] = F[19]; Sh[2][2] = F[20];
    SSh[0][0] = F[21]; SSh[0][1] = SSh[1][0] = F[22]; SSh[0][2] = SSh[2][0] = F[23];
    SSh[1][1] = F[24]; SSh[1][2] = SSh[2][1] = F[25]; SSh[2][2] = F[26];
    // Delta^2 = r(aj)^2 and Delta~^2 = (64 q)^(2/3) = 16 q r(q), r(x) = x^(-1/3)
    const double rf = rcbrt_pos(aj[c]);
    germano_sq(L, Sh, SSh, rf * rf, 16.0 * F[27] * rcbrt_pos(F[27]), ld3(csi, c), ld3(eta, c), ld3(zet, c), qn, nn);
}

template <class T>
__global__ void __launch_bounds__(DK_X * DK_Y) k_dpasd_k(Geo g, const typename DsmSum<T>::type *X, long nI,
                                                         const double *csi, const double *eta, const double *zet,
                                                         const double *aj, const double *LM, const double *MM,
                                                         double *part, int kchunk) {
    using S = typename DsmSum<T>::type;
    const int lane = threadIdx.x, r = threadIdx.y;
    const int tile = xcd_tile();              // i fastest, then j, then the k chunk
    const int bi = tile % gridDim.x, bj = (tile / gridDim.x) % gridDim.y, bk = tile / (gridDim.x * gridDim.y);
    const int i = 1 + bi * DK_X + lane, j = 1 + bj * DK_Y + r;
    const int kb = 1 + bk * kchunk, ke = min(kb + kchunk, g.mz - 1);
    const int ni = g.mx - 2, nj = g.my - 2;
    double sLM = 0.0, sMM = 0.0, sQN = 0.0, sNN = 0.0;
    if (i <= g.mx - 2 && j <= g.my - 2) {
        const double dij = simpson5_mask_sum(i, g.mx) * simpson5_mask_sum(j, g.my);
        const long SP = (long)ni * nj;
        const long x0 = (long)(j - 1) * ni + (i - 1);
        // FP32 sums: four output planes per step (planes k-2 .. k+5 read once for the four, two
        // planes per output), the filtered values kept in FP32 until each output's contraction;
        // the same values and the same order of the sums as two per step
        if constexpr (std::is_same<S, float>::value) {
            for (int k = kb; k < ke; k += 4) {
                const int nout = min(4, ke - k);
                bool in[8];
#pragma unroll
                for (int o = 0; o < 8; o++) in[o] = k - 2 + o >= 1 && k - 2 + o <= g.mz - 2 && o < nout + 4;
                const S *xk = X + (long)(k - 1) * SP + x0;
                S F4[4][NQ];
#pragma unroll
                for (int n = 0; n < NQ; n++) {
                    const S *xn = xk + (long)n * nI;
                    S v[8];
#pragma unroll
                    for (int o = 0; o < 8; o++) v[o] = in[o] ? xn[(o - 2) * SP] : S(0);
#pragma unroll
                    for (int u = 0; u < 4; u++) F4[u][n] = row5(n, v[u], v[u + 1], v[u + 2], v[u + 3], v[u + 4]);
                }
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    if (u < nout) {
                        double F[NQ];
#pragma unroll
                        for (int n = 0; n < NQ; n++) F[n] = double(F4[u][n]);
                        double qn, nn;
                        const int c = g.L(k + u, j, i);
                        dpasd_germano(F, 1.0 / (dij * simpson5_mask_sum(k + u, g.mz)), csi, eta, zet, aj, c, qn, nn);
                        sLM += LM[c];
                        sMM += MM[c];
                        sQN += qn;
                        sNN += nn;
                    }
                }
            }
        } else
        // two output planes per step: planes k-2 .. k+3 are read once for both, so each output reads
        // three planes from L2 instead of five.  Planes outside [1, mz-2] are outside the mask; the
        // conditions are uniform over the workgroup.
        for (int k = kb; k < ke; k += 2) {
            const bool two = k + 1 < ke;
            bool in[6];
#pragma unroll
            for (int o = 0; o < 6; o++) in[o] = k - 2 + o >= 1 && k - 2 + o <= g.mz - 2 && (o < 5 || two);
            const S *xk = X + (long)(k - 1) * SP + x0;
            double Fa[NQ], Fb[NQ];
#pragma unroll
            for (int n = 0; n < NQ; n++) {
                const S *xn = xk + (long)n * nI;
                S v[6];
#pragma unroll
                for (int o = 0; o < 6; o++) v[o] = in[o] ? xn[(o - 2) * SP] : S(0);
                Fa[n] = double(row5(n, v[0], v[1], v[2], v[3], v[4]));
                Fb[n] = double(row5(n, v[1], v[2], v[3], v[4], v[5]));
            }
            double qn, nn;
            int c = g.L(k, j, i);
            dpasd_germano(Fa, 1.0 / (dij * simpson5_mask_sum(k, g.mz)), csi, eta, zet, aj, c, qn, nn);
            sLM += LM[c];
            sMM += MM[c];
            sQN += qn;
            sNN += nn;
            if (two) {
                c = g.L(k + 1, j, i);
                dpasd_germano(Fb, 1.0 / (dij * simpson5_mask_sum(k + 1, g.mz)), csi, eta, zet, aj, c, qn, nn);
                sLM += LM[c];
                sMM += MM[c];
                sQN += qn;
                sNN += nn;
            }
        }
    }
    // one wave per j row: the sums of its 64 cells (zero past the array)
    sLM = wave_sum(sLM);
    sMM = wave_sum(sMM);
    sQN = wave_sum(sQN);
    sNN = wave_sum(sNN);
    if (lane == 0 && j <= g.my - 2) {
        double *p = part + (((long)(j - 1) * gridDim.x + bi) * gridDim.z + bk) * 4;
        p[0] = sLM; p[1] = sMM; p[2] = sQN; p[3] = sNN;
    }
}

// plane means of LM, MM, QN, NN over the (k, i) cells of each j plane and the plane's coefficient
// (les.c:1979-2048): the partials in (i tile, k chunk) order
__global__ void k_dpasd_plane(const double *part, int nparts, double maxCs, double *csP, Geo g) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x + 1;
    if (j > g.my - 2) return;
    double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
    const double *p = part + (long)(j - 1) * nparts * 4;
    for (int q = 0; q < nparts; q++) {
        s0 += p[4 * q]; s1 += p[4 * q + 1]; s2 += p[4 * q + 2]; s3 += p[4 * q + 3];
    }
    const double cnt = (double)((long)(g.mz - 2) * (g.mx - 2));
    const double J_LM = s0 / cnt, J_MM = s1 / cnt, J_QN = s2 / cnt, J_NN = s3 / cnt;
    const double beta = (J_QN * J_MM) / (J_NN * J_LM + 1e-10);
    const double v = 0.5 * (J_LM / (J_MM + 1e-10)) / beta;
    csP[j] = pmin(pmax(v, 0.0), maxCs);
}

template <class T>
__global__ void __launch_bounds__(256) k_dpasd_cs_nut(Geo g, const double *csP, const double *Sabs, const double *aj,
                                                      double *Cs, T *nut) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1, j = blockIdx.y * blockDim.y + threadIdx.y + 1,
              k = blockIdx.z + 1;
    if (i > g.mx - 2 || j > g.my - 2 || k > g.mz - 2) return;
    const int c = g.L(k, j, i);
    const double cs = csP[j];
    Cs[c] = cs;
    if (nut) {
        const double filter = cube_root(1. / aj[c]);
        nut[c] = T(cs * (filter * filter) * Sabs[c]);
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

namespace {
// k planes per chunk of a k-marching launch over `tiles` column tiles: about `target` workgroups in
// all, at least 8 planes per chunk
int plane_chunk(int nk, int tiles, int target) {
    int chunks = (target + tiles - 1) / tiles;
    chunks = max(1, min(chunks, nk / 8 > 0 ? nk / 8 : 1));
    return (nk + chunks - 1) / chunks;
}

// LM, MM (3-point test filter + Germano contraction) and |S| of every interior cell
// m32: FP32 copies of csi, eta, zet to read instead of the mesh's (FP32 flow fields only)
template <class T>
void launch_dsm_lm(const Geo &g, const tosca_mesh *m, const T *ucat, double *LM, double *MM, double *Sabs, hipStream_t s,
                   const tosca_mesh_f32 *m32 = nullptr) {
    const int nk = g.mz - 2;
    const int ti = (g.mx - 2 + DSM_TI - 1) / DSM_TI, tj = (g.my - 2 + DSM_TJ - 1) / DSM_TJ;
    const int kchunk = plane_chunk(nk, ti * tj, 2048);
    dim3 grid(ti, tj, (nk + kchunk - 1) / kchunk);
    if constexpr (std::is_same<T, float>::value) {
        if (m32) {
            TOSCA_LAUNCH("k_dsm_lm_f32m", (k_dsm_lm<float, float>), grid, dim3(64, DSM_NR), 0, s, g, ucat, m32->csi,
                         m32->eta, m32->zet, m->aj, LM, MM, Sabs, kchunk);
            return;
        }
    }
    TOSCA_LAUNCH(sizeof(T) == 8 ? "k_dsm_lm" : "k_dsm_lm_f32", k_dsm_lm<T>, grid, dim3(64, DSM_NR), 0, s, g, ucat, m->csi,
                 m->eta, m->zet, m->aj, LM, MM, Sabs, kchunk);
}

// launch shape of k_dpasd_k, which sizes the partials buffer
struct DpasdK { int ti, tj, chunks, kchunk; };
DpasdK dpasd_k_shape(const Geo &g) {
    DpasdK d;
    const int nk = g.mz - 2;
    d.ti = (g.mx - 2 + DK_X - 1) / DK_X;
    d.tj = (g.my - 2 + DK_Y - 1) / DK_Y;
    d.kchunk = plane_chunk(nk, d.ti * d.tj, 4096);
    d.chunks = (nk + d.kchunk - 1) / d.kchunk;
    return d;
}
}  // namespace

size_t dpasd_ws_doubles(const Geo &g) {
    const size_t nI = (size_t)(g.mx - 2) * (g.my - 2) * (g.mz - 2);
    const DpasdK d = dpasd_k_shape(g);
    return 3 * (size_t)g.nloc() + NQ * nI + (size_t)(g.my - 2) * d.ti * d.chunks * 4 + (size_t)g.my + 64;
}

template <class T>
void fused_dpasd(const Geo &g, const tosca_mesh *m, const T *ucat, double maxCs, double *cs, T *nut, double *ws,
                 hipStream_t s) {
    using S = typename DsmSum<T>::type;
    const long n = g.nloc();
    const long nI = (long)(g.mx - 2) * (g.my - 2) * (g.mz - 2);
    double *LM = ws, *MM = ws + n, *Sabs = ws + 2 * n;
    S *X = reinterpret_cast<S *>(ws + 3 * n);
    const DpasdK d = dpasd_k_shape(g);
    double *part = ws + 3 * n + NQ * nI;      // an FP32 X takes half of its share
    double *csP = part + (long)(g.my - 2) * d.ti * d.chunks * 4;
    launch_dsm_lm<T>(g, m, ucat, LM, MM, Sabs, s);
    const int nk = g.mz - 2;
    const int ti = (g.mx - 2 + DP_TI - 1) / DP_TI, tj = (g.my - 2 + DP_TJ - 1) / DP_TJ;
    const int kchunk = plane_chunk(nk, ti * tj, 1536);
    TOSCA_LAUNCH(sizeof(T) == 8 ? "k_dpasd_ij" : "k_dpasd_ij_f32", k_dpasd_ij<T>, dim3(ti, tj, (nk + kchunk - 1) / kchunk),
                 dim3(64, DP_NR), 0, s, g, ucat, m->csi, m->eta, m->zet, m->aj, X, nI, kchunk);
    TOSCA_LAUNCH(sizeof(T) == 8 ? "k_dpasd_k" : "k_dpasd_k_f32", k_dpasd_k<T>, dim3(d.ti, d.tj, d.chunks), dim3(DK_X, DK_Y),
                 0, s, g, X, nI, m->csi, m->eta, m->zet, m->aj, LM, MM, part, d.kchunk);
    TOSCA_LAUNCH("k_dpasd_plane", k_dpasd_plane, dim3((g.my - 2 + 255) / 256), dim3(256), 0, s, part, d.ti * d.chunks,
                 maxCs, csP, g);
    TOSCA_LAUNCH(sizeof(T) == 8 ? "k_dpasd_cs_nut" : "k_dpasd_cs_nut_f32", k_dpasd_cs_nut<T>,
                 dim3((g.mx - 2 + 63) / 64, (g.my - 2 + 3) / 4, g.mz - 2), dim3(64, 4), 0, s, g, csP, Sabs, m->aj, cs, nut);
}

template <class T>
void fused_dsm(const Geo &g, const tosca_mesh *m, const T *ucat, double maxCs, double *cs, T *nut, double *ws,
               hipStream_t s, const tosca_mesh_f32 *m32) {
    const long n = g.nloc();
    double *LM = ws, *MM = ws + n, *Sabs = ws + 2 * n;
    const int nk = g.mz - 2;
    launch_dsm_lm<T>(g, m, ucat, LM, MM, Sabs, s, m32);
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
                               hipStream_t, const tosca_mesh_f32 *);                                         \
    template void fused_dpasd<T>(const Geo &, const tosca_mesh *, const T *, double, double *, T *, double *, \
                                 hipStream_t);                                                               \
    template void fused_nut<T>(const Geo &, const tosca_mesh *, const T *, const double *, T *, T *,         \
                               hipStream_t);                                                                 \
    template void nut_bcs<T>(T *, const tosca_params *, const Geo &, hipStream_t);
TOSCA_LES_INSTANTIATE(double)
TOSCA_LES_INSTANTIATE(float)

}  // namespace tosca
