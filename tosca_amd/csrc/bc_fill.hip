// bc_fill.hip -- boundary-condition ghost fills of the flow fields on gfx950 (src/boundary.c).
// Compiled with FP contraction off: the fills are copies, reflections and short expressions,
// evaluated as the reference writes them.
//
//   UpdateContravariantBCs  boundary.c:614-822    -> k_contra_periodic_copy, k_contra_bcs
//   UpdateCartesianBCs      boundary.c:824-2050   -> k_cart_bcs (fixedValue, slip, noSlip,
//                                                     zeroGradient, periodic patches, Cabot
//                                                     wall-function i and j patches, Shumann
//                                                     wall-function j patches)
//   Cabot wall law          wallfunctions.c:280-585 -> k_cabot_intervals, k_cabot_prefix (the
//                                                     integration table), cabot_utau (Newton)
//   UpdateTemperatureBCs    boundary.c:2052-2628  -> k_temp_bcs (zeroGradient, fixedValue,
//                                                     fixedGradient, periodic patches, surface
//                                                     temperature on jLeft)
// The reference reads the local array as it was on entry and writes a global copy; every kernel
// here writes only locations that no thread of the same launch reads (ghost planes of the cell
// fills; boundary faces of the flux fill, whose one cross-face read -- the left periodic face
// copying the right face -- runs in a launch of its own before the right face is rewritten).
// Each fill ends with the ghost scatter of the periodic-wrap axes.
#include "kernels.h"

namespace tosca {

namespace {

struct PatchSet {
    int bc[6];          // patch kinds, order iLeft iRight jLeft jRight kLeft kRight
    int wall[6];        // velocity wall model of the i and j patches (-1 Cabot, -3 Shumann), 0 elsewhere
    double rough[6];    // Cabot roughness each i and j patch uses
    double val[18];     // fixedValue vectors (velocity) or values (temperature, first 6)
};

// ---- Cabot wall law (wallfunctions.c) --------------------------------------------------------
// near_wall_eddy_viscosity (wallfunctions.c:280-284): nu_t+ = kappa y+ (1 - exp(-y+/19))^2
__device__ double near_wall_nut(double yp) { return 0.41 * yp * pow(1. - exp(-yp / 19.), 2.0); }

// integral of 1 / (1 + nu_t+) over the y+ interval [2 (i - 1), 2 i], 24 Simpson 3/8 panels, as
// pre_integrate forms it (wallfunctions.c:316-355): the panel ends from ya + dy k, the interior
// points from a running panel start; written to tab[i] and summed in order by k_cabot_prefix
__global__ void k_cabot_intervals(double *tab, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x + 1;
    if (i > n) return;
    const int N = 24;
    const double ya = (double)(i - 1) * CABOT_INTERVAL, yb = (double)i * CABOT_INTERVAL, dy = (yb - ya) / (double)N;
    double val = 0, y0 = ya;
    double E0 = 1. / (1. + near_wall_nut(ya + dy * 0.0));
    for (int k = 0; k < N; k++) {
        const double E1 = 1. / (1. + near_wall_nut(ya + dy * (double)(k + 1)));
        const double F1 = 1. / (1. + near_wall_nut(y0 + dy * 1. / 3.));
        const double F2 = 1. / (1. + near_wall_nut(y0 + dy * 2. / 3.));
        val += dy / 3. * (3 * E0 + 9 * F1 + 9 * F2 + 3 * E1) / 8.;
        y0 += dy;
        E0 = E1;
    }
    tab[i] = val;
}

// tab[0] = 0, tab[i] = tab[i-1] + interval i (integration_buffer, :313-356) as a chunked running
// sum: thread c adds the intervals of chunk c in index order (k_cabot_chunk_sums), one thread adds
// the chunk totals in order (k_cabot_chunk_offsets), and thread c walks its chunk again from that
// offset (k_cabot_chunk_apply).  The association differs from one running sum only in where the
// chunk totals are formed.
constexpr long CABOT_CHUNK = 4096;
__global__ void k_cabot_chunk_sums(const double *tab, long n, double *sums, long nc) {
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    const long i0 = 1 + c * CABOT_CHUNK, i1 = min(n, i0 + CABOT_CHUNK - 1);
    double s = 0.0;
#pragma unroll 16
    for (long i = i0; i <= i1; i++) s = s + tab[i];
    sums[c] = s;
}

__global__ void k_cabot_chunk_offsets(double *sums, long nc) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    double s = 0.0;
    for (long c = 0; c < nc; c++) {
        const double v = sums[c];
        sums[c] = s;
        s = s + v;
    }
}

__global__ void k_cabot_chunk_apply(double *tab, long n, const double *sums, long nc) {
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    if (c == 0) tab[0] = 0.0;
    const long i0 = 1 + c * CABOT_CHUNK, i1 = min(n, i0 + CABOT_CHUNK - 1);
    double s = sums[c];
#pragma unroll 16
    for (long i = i0; i <= i1; i++) {
        s = s + tab[i];
        tab[i] = s;
    }
}

// integrate_F (wallfunctions.c:364-420) divided out: returns the integral of 1 / (1 + nu_t+) from
// 0 to y+ = yb utau / nu, over utau.  Below y+ = 1e7 it interpolates the table linearly; above, the
// reference adds 4 panels from 1e7 to y+ to integration_buffer[n_yp], whose local n_yp is 0 on
// every call but the process's first (pre_integrate sets it only once), so to 0.  Where the
// reference would index outside its table (y+ < 0, or y+ at the table's last point) the index is
// clamped to the table.
__device__ double cabot_F(const double *tab, double nu, double utau, double yb) {
    const double yp = yb * utau / nu;
    if (yp <= (double)CABOT_MAX_YP) {
        const double qb = yp / (double)CABOT_INTERVAL;
        long ib = qb > 0.0 ? (long)qb : 0;              // (PetscInt) truncation; no conversion of y+ < 0
        ib = ib > CABOT_NYP - 1 ? CABOT_NYP - 1 : ib;
        const double int_b = (tab[ib + 1] - tab[ib]) / (double)CABOT_INTERVAL * (yp - (double)ib * CABOT_INTERVAL) + tab[ib];
        return (int_b - 0.0) / utau;
    }
    const int N = 4;
    const double ya = (double)CABOT_MAX_YP, dy = (yp - ya) / (double)N;
    double val = 0.0, y0 = ya;
    double E0 = 1. / (1. + near_wall_nut(ya + dy * 0.0));
    for (int k = 0; k < N; k++) {
        const double E1 = 1. / (1. + near_wall_nut(ya + dy * (double)(k + 1)));
        const double F1 = 1. / (1. + near_wall_nut(y0 + dy * 1. / 3.));
        const double F2 = 1. / (1. + near_wall_nut(y0 + dy * 2. / 3.));
        val += dy / 3. * (3 * E0 + 9 * F1 + 9 * F2 + 3 * E1) / 8.;
        y0 += dy;
        E0 = E1;
    }
    return val / utau;
}

__device__ double cabot_f(const double *tab, double nu, double u, double y, double utau) {   // f_Cabot, :467-470
    return utau * utau * cabot_F(tab, nu, utau, y) - u;
}

// uTauCabot (wallfunctions.c:519-543): Newton from 0.01 with a central-difference derivative
// (step 1e-7), at most 30 steps, stopping when a step is below 1e-10
__device__ double cabot_utau(const double *tab, double nu, double u, double y) {
    double x = 0.0, x0 = 0.01;
    const double eps = 1.e-7;
    for (int it = 0; it < 30; it++) {
        const double df = (cabot_f(tab, nu, u, y, x0 + eps) - cabot_f(tab, nu, u, y, x0 - eps)) / (2.0 * eps);
        x = x0 - cabot_f(tab, nu, u, y, x0) / df;
        if (fabs(x0 - x) < 1.e-10) break;
        x0 = x;
    }
    return x;
}

// utau_wf (wallfunctions.c:568-585): log law over 0.033 ks, or a 1/7 power law for tiny ks
__device__ double rough_utau(double nu, double ks, double sb, double ut) {
    const double k_s = ks * 0.033, kappa = 0.41;
    if (k_s > 1.e-13) return ut * kappa / log(sb / k_s);
    const double A = 8.3, B = 1.0 / 7.0;
    return pow(ut * pow(nu, B) / (A * pow(sb, B)), 1.0 / (1.0 + B));
}

__device__ __forceinline__ void across(int a, int &t1, int &t2) {
    t1 = (a == 0) ? 1 : 0;
    t2 = (a == 2) ? 1 : 2;
}

__device__ __forceinline__ int patch_type(int bc) {     // checkBCsAndSetPatchTypes, boundary.c:318-348
    return (bc == TOSCA_BCU_NOSLIP || bc == TOSCA_BCU_SLIP || bc == TOSCA_BCU_WALLFUNCTION) ? 1 : 0;
}

// left face of a periodic axis: the right face's flux from before the call (boundary.c:760-806);
// type 1 reads face m-2, type 2 the ghost face -2
__global__ void k_contra_periodic_copy(double *ucont, Geo g, PatchSet P) {
    const int a = blockIdx.z;
    if (P.bc[2 * a] != TOSCA_BCU_PERIODIC || g.per[a] == TOSCA_PERIODIC_NONE) return;
    const int m[3] = {g.mx, g.my, g.mz};
    int t1, t2;
    across(a, t1, t2);
    const int x1 = blockIdx.x * blockDim.x + threadIdx.x, x2 = blockIdx.y;
    if (x1 >= m[t1] - 1 || x2 >= m[t2] - 1) return;
    int d[3], s[3];
    d[a] = 0;
    s[a] = (g.per[a] == TOSCA_PERIODIC_COPY) ? m[a] - 2 : -2;
    d[t1] = s[t1] = x1;
    d[t2] = s[t2] = x2;
    ucont[3 * (long)g.L(d[2], d[1], d[0]) + a] = ucont[3 * (long)g.L(s[2], s[1], s[0]) + a];
}

// every other boundary face: PatchType 0 -> flux of the face-averaged ucat (boundary.c:652-688),
// PatchType 1 -> 0 (:690-713), zeroGradient -> inflow removed (:715-739)
__global__ void k_contra_bcs(double *ucont, const double *ucat, const double *icsi, const double *jeta,
                             const double *kzet, Geo g, PatchSet P) {
    const int q = blockIdx.z, a = q >> 1, side = q & 1;
    const int b = P.bc[q];
    if (side == 0 && b == TOSCA_BCU_PERIODIC && g.per[a] != TOSCA_PERIODIC_NONE) return;   // copied above
    const int m[3] = {g.mx, g.my, g.mz};
    int t1, t2;
    across(a, t1, t2);
    const int x1 = blockIdx.x * blockDim.x + threadIdx.x, x2 = blockIdx.y;
    if (x1 >= m[t1] - 1 || x2 >= m[t2] - 1) return;
    int c[3], n[3];
    c[a] = side ? m[a] - 2 : 0;
    c[t1] = x1;
    c[t2] = x2;
    n[0] = c[0]; n[1] = c[1]; n[2] = c[2];
    n[a] += 1;
    const long ci = g.L(c[2], c[1], c[0]);
    double v = ucont[3 * ci + a];
    if (patch_type(b) == 0) {
        const double *S = (a == 0) ? icsi : (a == 1) ? jeta : kzet;
        const V3 A = ld3(ucat, g.L(n[2], n[1], n[0])), B = ld3(ucat, ci), s = ld3(S, ci);
        v = 0.5 * (A.x + B.x) * s.x + 0.5 * (A.y + B.y) * s.y + 0.5 * (A.z + B.z) * s.z;
    } else {
        v = 0.0;
    }
    if (b == TOSCA_BCU_ZEROGRADIENT) {
        if (side == 0 && v > 0.0) v = 0.0;
        if (side == 1 && v < 0.0) v = 0.0;
    }
    ucont[3 * ci + a] = v;
}

// unit covariant base vector of direction a at a cell (calculateNormal, inline.h:1231-1262)
__device__ V3 unit_normal(V3 cs, V3 et, V3 ze, int a) {
    const double a11 = cs.x, a12 = cs.y, a13 = cs.z, a21 = et.x, a22 = et.y, a23 = et.z, a31 = ze.x, a32 = ze.y, a33 = ze.z;
    const double det = a11 * (a33 * a22 - a32 * a23) - a21 * (a33 * a12 - a32 * a13) + a31 * (a23 * a12 - a22 * a13);
    double x, y, z;     // column a of the inverse (Calculate_Covariant_metrics, inline.h:1210-1229)
    if (a == 0) {
        x = (a33 * a22 - a32 * a23) / det; y = -(a33 * a21 - a31 * a23) / det; z = (a32 * a21 - a31 * a22) / det;
    } else if (a == 1) {
        x = -(a33 * a12 - a32 * a13) / det; y = (a33 * a11 - a31 * a13) / det; z = -(a32 * a11 - a31 * a12) / det;
    } else {
        x = (a23 * a12 - a22 * a13) / det; y = -(a23 * a11 - a21 * a13) / det; z = (a22 * a11 - a21 * a12) / det;
    }
    const double sum = sqrt(x * x + y * y + z * z);
    return V3{x / sum, y / sum, z / sum};
}

// the kLeft inlet-function profile (type 0: none)
struct InletArgs {
    int type;
    const double *cent;
    double grnd, ymin, ly, u_tau, rough, h_inv, amp, periods, udir[3], uref_mag, fc, h_ref;
    // power law
    double uref[3], href, uprime;
    const int *noise;
    // mapped inflow planes
    const double *plane;
    int n1, n2, prds1, prds2, merge1, spline, shift2;
    double top_len, top_delta, top_z[10], top_u[10][3];
    const int *cells, *cells1, *cells2, *yids;
    const double *w, *w1, *w2, *yw;
};

namespace {
// findInterpolationWeightsWithExtrap (inline.h:5943-5966) over the 10 ascending heights pv: the
// two closest points (a point whose distance equals the runner-up's to 1e-5 relative is passed
// over, findInterpolationWeights, inline.h:5873-5938), ordered by height, and their weights
__device__ void interp_weights10(const double *pv, double pval, double w[2], int id[2]) {
    if (pval >= pv[9]) { w[0] = w[1] = 0.5; id[0] = id[1] = 9; return; }
    if (pval <= pv[0]) { w[0] = w[1] = 0.5; id[0] = id[1] = 0; return; }
    const double eqTol = 1e-5;
    int i1 = 0, i2 = 1;
    double d1 = fabs(pv[0] - pval), d2 = fabs(pv[1] - pval);
    for (int i = 0; i < 10; i++) {
        const double di = fabs(pv[i] - pval);
        const double tol1 = fabs(di - d2) / di;
        if (di < d1 && tol1 > eqTol) { i2 = i1; d2 = d1; i1 = i; d1 = di; }
        const double tol2 = fabs(di - d1) / di;
        if (di < d2 && i != i1 && tol2 > eqTol) { i2 = i; d2 = di; }
    }
    if (pv[i2] < pv[i1]) { const int t = i2; i2 = i1; i1 = t; }
    const double idx = double(i2 - i1) / (pv[i2] - pv[i1]) * (pval - pv[i1]) + double(i1);
    w[0] = (double(i2) - idx) / double(i2 - i1);
    w[1] = (idx - double(i1)) / double(i2 - i1);
    id[0] = i1; id[1] = i2;
}

// scaleHyperTangBot / scaleHyperTangTop (inline.h:5971-5999)
__device__ __forceinline__ double tanh_bot(double h, double H, double delta) {
    const double cb = H - 0.5 * delta;
    return 0.5 * (1.0 - tanh(7.0 * (h - cb) / delta));
}
__device__ __forceinline__ double tanh_top(double h, double H, double delta) {
    const double cb = H - 0.5 * delta;
    return 0.5 * (1.0 + tanh(7.0 * (h - cb) / delta));
}

// the n1Merge blend of a database value u at height h into the averaged top profile
// (boundary.c:1318-1343, 1580-1617)
__device__ V3 merge_top(const InletArgs &In, double h, V3 u) {
    double Wg[2];
    int ID[2];
    interp_weights10(In.top_z, h, Wg, ID);
    const double b = tanh_bot(h, In.top_len, In.top_delta), t = tanh_top(h, In.top_len, In.top_delta);
    const double *a0 = In.top_u[ID[0]], *a1 = In.top_u[ID[1]];
    return V3{b * u.x + t * (a0[0] * Wg[0] + a1[0] * Wg[1]), b * u.y + t * (a0[1] * Wg[0] + a1[1] * Wg[1]),
              b * u.z + t * (a0[2] * Wg[0] + a1[2] * Wg[1])};
}

// one component of the stencil sum w_0 p_0 + w_1 p_1 + ... over np points of cell (j, i) of the
// interpolation tables (stencil entries clamped into the plane)
__device__ double stencil_sum(const InletArgs &In, const int *cells, const double *w, int np, long jm, int comp) {
    const long base = jm * np;
    double s = 0.0;
    for (int q = 0; q < np; q++) {
        const int pj = min(max(cells[2 * (base + q)], 0), In.n1 + 1), pi = min(max(cells[2 * (base + q) + 1], 0), In.n2 + 1);
        const double v = In.plane[3 * ((long)pj * (In.n2 + 2) + pi) + comp];
        s = (q == 0) ? w[base + q] * v : s + w[base + q] * v;
    }
    return s;
}

// the interpolated database velocity of mesh cell (j, i), row-major index jm = j mx + i
__device__ V3 interp_cell(const InletArgs &In, long jm) {
    V3 u;
    u.x = stencil_sum(In, In.cells, In.w, 4, jm, 0);
    if (In.spline) {
        u.y = stencil_sum(In, In.cells2, In.w2, 6, jm, 1);
        u.z = stencil_sum(In, In.cells1, In.w1, 6, jm, 2);
    } else {
        u.y = stencil_sum(In, In.cells, In.w, 4, jm, 1);
        u.z = stencil_sum(In, In.cells, In.w, 4, jm, 2);
    }
    return u;
}

// the mapped-inflow ghost value of cell (j, i) at height h (typeU 3: boundary.c:1302-1371;
// typeU 4: :1372-1622)
__device__ V3 mapped_inflow(const InletArgs &In, const Geo &g, int j, int i, double h) {
    auto P = [&](int pj, int pi) { return ld3(In.plane, pj * (In.n2 + 2) + pi); };
    if (In.type == TOSCA_INLET_MAPPED) {
        int jif = j % In.n1 == 0 ? In.n1 : j % In.n1;
        int iif = i % In.n2 == 0 ? In.n2 : i % In.n2;
        if (j <= In.n1 * In.prds1 && i <= In.n2 * In.prds2) return In.merge1 ? merge_top(In, h, P(jif, iif)) : P(jif, iif);
        if (In.merge1) return V3{In.top_u[9][0], In.top_u[9][1], In.top_u[9][2]};
        if (j > In.n1 * In.prds1) jif = In.n1;
        if (i > In.n2 * In.prds2) iif = In.n2;
        return P(jif, iif);
    }
    V3 u;
    if (In.shift2) {
        const int ir = min(max(In.yids[i], 1), g.mx - 1), il = ir - 1;
        const V3 uL = interp_cell(In, (long)j * g.mx + il), uR = interp_cell(In, (long)j * g.mx + ir);
        const double yw = In.yw[i], ow = 1.0 - yw;
        u = V3{ow * uL.x + yw * uR.x, ow * uL.y + yw * uR.y, ow * uL.z + yw * uR.z};
    } else {
        u = interp_cell(In, (long)j * g.mx + i);
    }
    return In.merge1 ? merge_top(In, h, u) : u;
}
}  // namespace

// ---- the Nieuwstadt inflow profile (src/inflow.c:2732-2811) -----------------------------------------
// Complex arithmetic of its own (no std::complex in device code): products and sums as written,
// quotients by Smith's algorithm, log / exp / sqrt / sin / tan from their real and imaginary parts,
// and a real base raised to a complex power as polar(x^Re w, Im w log x), 0 for x = 0.  Each
// function follows the steps of the reference's (inline.h:372-525), with its coefficient tables.
namespace {
struct cx { double r, i; };
__device__ __forceinline__ cx operator+(cx a, cx b) { return cx{a.r + b.r, a.i + b.i}; }
__device__ __forceinline__ cx operator-(cx a, cx b) { return cx{a.r - b.r, a.i - b.i}; }
__device__ __forceinline__ cx operator+(cx a, double b) { return cx{a.r + b, a.i}; }
__device__ __forceinline__ cx operator-(cx a, double b) { return cx{a.r - b, a.i}; }
__device__ __forceinline__ cx operator-(double a, cx b) { return cx{a - b.r, -b.i}; }
__device__ __forceinline__ cx operator+(double a, cx b) { return cx{a + b.r, b.i}; }
__device__ __forceinline__ cx operator*(cx a, cx b) { return cx{a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r}; }
__device__ __forceinline__ cx operator*(double s, cx a) { return cx{s * a.r, s * a.i}; }
__device__ __forceinline__ cx operator*(cx a, double s) { return cx{a.r * s, a.i * s}; }
__device__ __forceinline__ cx operator/(cx a, cx b) {
    if (fabs(b.r) >= fabs(b.i)) {
        const double t = b.i / b.r, d = b.r + b.i * t;
        return cx{(a.r + a.i * t) / d, (a.i - a.r * t) / d};
    }
    const double t = b.r / b.i, d = b.r * t + b.i;
    return cx{(a.r * t + a.i) / d, (a.i * t - a.r) / d};
}
__device__ __forceinline__ cx operator/(double a, cx b) { return cx{a, 0.0} / b; }
__device__ __forceinline__ cx operator/(cx a, double s) { return cx{a.r / s, a.i / s}; }
__device__ __forceinline__ double cabs_(cx a) { return hypot(a.r, a.i); }
__device__ __forceinline__ cx clog_(cx a) { return cx{log(hypot(a.r, a.i)), atan2(a.i, a.r)}; }
__device__ __forceinline__ cx cexp_(cx a) {
    const double e = exp(a.r);
    return cx{e * cos(a.i), e * sin(a.i)};
}
__device__ __forceinline__ cx csqrt_(cx a) {           // principal branch
    if (a.r == 0.0 && a.i == 0.0) return cx{0.0, a.i};
    const double t = sqrt(0.5 * (hypot(a.r, a.i) + fabs(a.r)));
    if (a.r >= 0.0) return cx{t, a.i / (2.0 * t)};
    return cx{fabs(a.i) / (2.0 * t), copysign(t, a.i)};
}
__device__ __forceinline__ cx csin_(cx a) { return cx{sin(a.r) * cosh(a.i), cos(a.r) * sinh(a.i)}; }
__device__ __forceinline__ cx ctan_(cx a) {
    const double d = cos(2.0 * a.r) + cosh(2.0 * a.i);
    return cx{sin(2.0 * a.r) / d, sinh(2.0 * a.i) / d};
}
__device__ __forceinline__ cx rpow_(double x, cx w) {   // x >= 0
    if (!(x > 0.0)) return cx{0.0, 0.0};
    const double rho = pow(x, w.r), th = w.i * log(x);
    return cx{rho * cos(th), rho * sin(th)};
}
__device__ __forceinline__ bool nonpos_int(cx x) { return round(x.r) == x.r && x.i == 0.0 && x.r <= 0.0; }

__constant__ double kLanczosG9[11] = {1.000000000000000174663,     5716.400188274341379136,
                                      -14815.30426768413909044,    14291.49277657478554025,
                                      -6348.160217641458813289,    1301.608286058321874105,
                                      -108.1767053514369634679,    2.605696505611755827729,
                                      -0.7423452510201416151527e-2, 0.5384136432509564062961e-7,
                                      -0.4023533141268236372067e-8};
__constant__ double kLanczosG607[15] = {0.99999999999999709182,    57.156235665862923517,     -59.597960355475491248,
                                        14.136097974741747174,     -0.49191381609762019978,   .33994649984811888699e-4,
                                        .46523628927048575665e-4,  -.98374475304879564677e-4, .15808870322491248884e-3,
                                        -.21026444172410488319e-3, .21743961811521264320e-3,  -.16431810653676389022e-3,
                                        .84418223983852743293e-4,  -.26190838401581408670e-4, .36899182659531622704e-5};
constexpr double kPi = 3.14159265358979323846;

__device__ cx cgamma_(cx x) {                            // inline.h:372-427
    cx z = x;
    if (x.r < 0.0) z = -1.0 * z;
    const cx g{9.0, 0.0};
    cx t = z + g, s{0.0, 0.0};
    for (int k = 10; k > 0; k--) {
        s = s + cx{kLanczosG9[k], 0.0} / t;
        t = t - 1.0;
    }
    s = s + cx{kLanczosG9[0], 0.0};
    const cx ss = z + g - 0.5;
    s = clog_(s * sqrt(2.0 * kPi)) + (z - 0.5) * clog_(ss) - ss;
    cx f = cexp_(s);
    if (x.r < 0.0) f = -1.0 * kPi / (x * f * csin_(kPi * x));
    if (nonpos_int(x)) f = cx{1e20, 0.0};
    if (round(x.r) == x.r && x.i == 0.0 && x.r > 0.0) {
        cx pp{1.0, 0.0};
        for (long m = 1; m < (long)x.r; m++) pp = pp * cx{(double)m, 0.0};
        f = pp;
    }
    return f;
}

__device__ cx cdigamma_(cx x) {                          // inline.h:429-476
    cx z = x;
    if (x.r < 0.5) z = 1.0 - z;
    const cx g{607.0 / 128.0, 0.0};
    cx n{0.0, 0.0}, d = n;
    for (int k = 14; k > 0; k--) {
        const cx dz = 1.0 / (z + (double)(k + 1) - 2.0);
        const cx dd = kLanczosG607[k] * dz;
        d = d + dd;
        n = n - dd * dz;
    }
    d = d + kLanczosG607[0];
    const cx gg = z + g - 0.5;
    cx f = clog_(gg) + (n / d - g / gg);
    if (x.r < 0.5) f = f - kPi * (1.0 / ctan_(kPi * x));
    if (nonpos_int(x)) f = cx{1e20, 0.0};
    return f;
}

__device__ cx hypergeom_(cx a, cx b, cx c, double z) {   // inline.h:478-525
    if (nonpos_int(c)) return cx{1e20, 0.0};
    const double epsilon = 1e-15;
    cx term{1.0, 0.0}, result{1.0, 0.0};
    int termination = 0;
    for (int k = 0; k < 30000; k++) {
        term = term * ((a + (double)k) * (b + (double)k) * z / ((c + (double)k) * ((double)k + 1.0)));
        if (cabs_(term) < cabs_(result * epsilon)) {
            if (++termination >= 3) return result;
        } else {
            termination = 0;
        }
        if (term.r == 0.0 && term.i == 0.0) break;
        result = result + term;
        if (isnan(result.r) || isnan(result.i)) return cx{nan(""), nan("")};
    }
    return result;
}

// W(eta) = Wg - dW(eta), the capped Ekman layer's complex velocity over u_tau
__device__ cx nieuwstadt_w(const InletArgs &In, double C, double eta) {
    const double vk = 0.4;
    const cx alpha = 0.5 + 0.5 * csqrt_(cx{1.0, 4.0 * C});
    const cx Const = cdigamma_(alpha + 1.0) + cdigamma_(alpha - 1.0) - 2.0 * cdigamma_(cx{1.0, 0.0});
    const cx Wg = 1.0 / vk * (cx{log(In.h_inv / In.rough), 0.0} - Const);
    const cx ga = cgamma_(alpha);
    const cx deltaW = cx{0.0, -1.0 / vk} * alpha * alpha * ga * ga / (C * cgamma_(2.0 * alpha)) *
                      rpow_(1.0 - eta, alpha - 1.0) * hypergeom_(alpha + 1.0, alpha - 1.0, 2.0 * alpha, 1.0 - eta);
    return Wg - deltaW;
}

// NieuwstadtInflowEvaluate at height h: (U, V, 0), rotated so that the direction at hReference
// lies along udir (V_r normalised with the already normalised U_r, as in the reference)
__device__ V3 nieuwstadt(const InletArgs &In, double h) {
    const double vk = 0.4;
    const double C = In.fc * In.h_inv / (vk * In.u_tau);
    const cx Wp = nieuwstadt_w(In, C, h < In.h_inv ? h / In.h_inv : 1.0);
    const double U_p = Wp.r * In.u_tau, V_p = Wp.i * In.u_tau;
    const cx Wr = nieuwstadt_w(In, C, In.h_ref / In.h_inv);
    double U_r = Wr.r * In.u_tau, V_r = Wr.i * In.u_tau;
    U_r = U_r / sqrt(U_r * U_r + V_r * V_r);
    V_r = V_r / sqrt(U_r * U_r + V_r * V_r);
    const double theta = acos(In.udir[0] * U_r + In.udir[1] * V_r);
    return V3{cos(theta) * U_p - sin(theta) * V_p, sin(theta) * U_p + cos(theta) * V_p, 0.0};
}
}  // namespace

// ghost cell of a patch next to an interior boundary cell (transverse indices interior)
__global__ void k_cart_bcs(double *ucat, const double *csi, const double *eta, const double *zet, const double *aj,
                           double *ustar, const double *tab, double nu, Geo g, PatchSet P, InletArgs In) {
    const int q = blockIdx.z, a = q >> 1, side = q & 1;
    const int m[3] = {g.mx, g.my, g.mz};
    int t1, t2;
    across(a, t1, t2);
    const int x1 = blockIdx.x * blockDim.x + threadIdx.x + 1, x2 = blockIdx.y + 1;
    if (x1 > m[t1] - 2 || x2 > m[t2] - 2) return;
    int c[3], gh[3], c2[3];
    c[t1] = gh[t1] = c2[t1] = x1;
    c[t2] = gh[t2] = c2[t2] = x2;
    c[a] = side ? m[a] - 2 : 1;
    gh[a] = side ? m[a] - 1 : 0;
    c2[a] = side ? m[a] - 3 : 2;
    const long ci = g.L(c[2], c[1], c[0]), gi = g.L(gh[2], gh[1], gh[0]);
    const int b = P.bc[q];
    V3 out;
    if (b == TOSCA_BCU_FIXEDVALUE) {                                     // boundary.c:1631-1666
        out = V3{P.val[3 * q], P.val[3 * q + 1], P.val[3 * q + 2]};
    } else if (b == TOSCA_BCU_SLIP) {                                    // :1670-1830
        const V3 n = unit_normal(ld3(csi, ci), ld3(eta, ci), ld3(zet, ci), a), U = ld3(ucat, ci);
        const double d = U.x * n.x + U.y * n.y + U.z * n.z;
        const V3 N2{2.0 * d * n.x, 2.0 * d * n.y, 2.0 * d * n.z};
        out = V3{U.x - N2.x, U.y - N2.y, U.z - N2.z};
    } else if (b == TOSCA_BCU_NOSLIP) {                                  // :1832-1867
        const V3 U = ld3(ucat, ci);
        out = V3{-1.0 * U.x, -1.0 * U.y, -1.0 * U.z};
    } else if (b == TOSCA_BCU_WALLFUNCTION && a <= 1 && P.wall[q] == -1) {   // Cabot, :968-1150
        const V3 cs = ld3(csi, ci), et = ld3(eta, ci), ze = ld3(zet, ci), A = (a == 0) ? cs : et;
        const double area = sqrt(A.x * A.x + A.y * A.y + A.z * A.z);
        const double sb = 0.5 / aj[ci] / area;
        const V3 Ub = ld3(ucat, ci), n = unit_normal(cs, et, ze, a);
        const double d = Ub.x * n.x + Ub.y * n.y + Ub.z * n.z;
        const V3 bN{d * n.x, d * n.y, d * n.z};
        const V3 bP{Ub.x - bN.x, Ub.y - bN.y, Ub.z - bN.z};
        const double mag = sqrt(bP.x * bP.x + bP.y * bP.y + bP.z * bP.z);
        // a cell on an i and a j Cabot patch keeps the j patch's u*, which the reference computes
        // after the i patch's; the ghost of each patch uses its own
        const bool jwins = a == 0 && ((c[1] == 1 && P.bc[2] == TOSCA_BCU_WALLFUNCTION && P.wall[2] == -1) ||
                                      (c[1] == m[1] - 2 && P.bc[3] == TOSCA_BCU_WALLFUNCTION && P.wall[3] == -1));
        if (mag < 1.e-20) {                  // the reference solves for u* and then discards it
            if (!jwins) ustar[ci] = 0.0;
            out = V3{0.0, 0.0, 0.0};
        } else {
            const double us = P.rough[q] > 1.e-19 ? rough_utau(nu, P.rough[q], sb, mag) : cabot_utau(tab, nu, mag, sb);
            if (!jwins) ustar[ci] = us;
            const double tau_w = us * us;
            const double a1 = mag - 2.0 * sb * tau_w / nu, a2 = -1.0 * mag;
            const double gmag = a1 < a2 ? a2 : a1;                          // PetscMax
            const double um = sqrt(bP.x * bP.x + bP.y * bP.y + bP.z * bP.z) + 1.e-12;   // nUnit
            const V3 dir{bP.x / um, bP.y / um, bP.z / um};
            out = V3{gmag * dir.x - bN.x, gmag * dir.y - bN.y, gmag * dir.z - bN.z};
        }
    } else if (b == TOSCA_BCU_WALLFUNCTION && a == 1 && P.wall[q] == -3) {   // Shumann, :1153-1224
        // wall-parallel part extrapolated linearly from the next cell, wall-normal part reflected
        const V3 n = unit_normal(ld3(csi, ci), ld3(eta, ci), ld3(zet, ci), 1);
        const V3 Ub = ld3(ucat, ci), Uc = ld3(ucat, g.L(c2[2], c2[1], c2[0]));
        const double db = Ub.x * n.x + Ub.y * n.y + Ub.z * n.z, dc = Uc.x * n.x + Uc.y * n.y + Uc.z * n.z;
        const V3 bN{db * n.x, db * n.y, db * n.z}, cN{dc * n.x, dc * n.y, dc * n.z};
        const V3 bP{Ub.x - bN.x, Ub.y - bN.y, Ub.z - bN.z}, cP{Uc.x - cN.x, Uc.y - cN.y, Uc.z - cN.z};
        const V3 gp{2. * bP.x - cP.x, 2. * bP.y - cP.y, 2. * bP.z - cP.z};
        out = V3{gp.x - bN.x, gp.y - bN.y, gp.z - bN.z};
    } else if (b == TOSCA_BCU_GENERIC && q == 4 && In.type != 0) {      // inletFunction, :1228-1300
        const V3 cc = ld3(In.cent, ci);
        double uMag = 0.0;
        if (In.type == TOSCA_INLET_NIEUWSTADT) {                         // :1287-1292
            out = nieuwstadt(In, cc.z - In.grnd);
        } else if (In.type == TOSCA_INLET_POWERLAW) {                    // :1233-1252
            const double h = cc.z - In.grnd, alpha = 0.107027;
            const double pw = pow(h / In.href, alpha);
            const double turb = 1 + double(In.noise[(long)c[1] * g.mx + c[0]]) / 10000. * In.uprime;
            out = V3{pw * In.uref[0] * turb, pw * In.uref[1] * turb, pw * In.uref[2] * turb};
        } else if (In.type == TOSCA_INLET_MAPPED || In.type == TOSCA_INLET_INTERP) {
            out = mapped_inflow(In, g, c[1], c[0], cc.z - In.grnd);
        } else if (In.type == TOSCA_INLET_LOGLAW) {                      // :1254-1286
            const double h = cc.z - In.grnd, vk = 0.4;
            if (h <= In.h_inv) {
                const double v = (In.u_tau / vk) * log(h / In.rough);
                uMag = (v < 1e-5) ? 1e-5 : v;                            // PetscMax
            } else {
                uMag = (In.u_tau / vk) * log(In.h_inv / In.rough);
            }
        } else {                                                         // sinusoidal, :1294-1299
            const double fraction = (cc.y - In.ymin) / In.ly;
            uMag = (1.0 + In.amp * cos(In.periods * 2.0 * 3.14159265358979323846 * fraction)) * In.uref_mag;
        }
        if (In.type == TOSCA_INLET_LOGLAW || In.type == TOSCA_INLET_SINUSOIDAL)
            out = V3{uMag * In.udir[0], uMag * In.udir[1], uMag * In.udir[2]};
    } else if (b == TOSCA_BCU_ZEROGRADIENT) {                            // :1869-1928
        const V3 U = ld3(ucat, ci), U2 = ld3(ucat, g.L(c2[2], c2[1], c2[0]));
        out = V3{2 * U.x - U2.x, 2 * U.y - U2.y, 2 * U.z - U2.z};
    } else if (b == TOSCA_BCU_PERIODIC && g.per[a] != TOSCA_PERIODIC_NONE) {   // :1930-1990
        int s[3] = {c[0], c[1], c[2]};
        if (g.per[a] == TOSCA_PERIODIC_COPY) s[a] = side ? 1 : m[a] - 2;
        else s[a] = side ? m[a] + 1 : -2;
        out = ld3(ucat, g.L(s[2], s[1], s[0]));
    } else {
        return;
    }
    ucat[3 * gi] = out.x;
    ucat[3 * gi + 1] = out.y;
    ucat[3 * gi + 2] = out.z;
}

__global__ void k_temp_bcs(double *t, const double *csi, const double *eta, const double *zet, const double *aj, Geo g,
                           PatchSet P) {
    const int q = blockIdx.z, a = q >> 1, side = q & 1;
    const int m[3] = {g.mx, g.my, g.mz};
    int t1, t2;
    across(a, t1, t2);
    const int x1 = blockIdx.x * blockDim.x + threadIdx.x + 1, x2 = blockIdx.y + 1;
    if (x1 > m[t1] - 2 || x2 > m[t2] - 2) return;
    int c[3], gh[3];
    c[t1] = gh[t1] = x1;
    c[t2] = gh[t2] = x2;
    c[a] = side ? m[a] - 2 : 1;
    gh[a] = side ? m[a] - 1 : 0;
    const long ci = g.L(c[2], c[1], c[0]), gi = g.L(gh[2], gh[1], gh[0]);
    const int b = P.bc[q];
    double out;
    if (b == TOSCA_BCT_ZEROGRADIENT) {                                   // boundary.c:2379-2426
        out = t[ci];
    } else if (b == TOSCA_BCT_SURFACETEMP) {                             // :2392-2410, jLeft only
        out = 2.0 * P.val[q] - t[ci];
    } else if (b == TOSCA_BCT_FIXEDVALUE) {                              // :2429-2458
        out = P.val[q];
    } else if (b == TOSCA_BCT_FIXEDGRADIENT) {                           // :2461-2555
        const V3 s = ld3((a == 0) ? csi : (a == 1) ? eta : zet, ci);
        const double area = sqrt(s.x * s.x + s.y * s.y + s.z * s.z);
        const double d = (1.0 / aj[ci]) / area;
        out = d * P.val[q] + t[ci];
    } else if (b == TOSCA_BCT_PERIODIC && g.per[a] != TOSCA_PERIODIC_NONE) {   // :2557-2596
        int s[3] = {c[0], c[1], c[2]};
        if (g.per[a] == TOSCA_PERIODIC_COPY) s[a] = side ? 1 : m[a] - 2;
        else s[a] = side ? m[a] + 1 : -2;
        out = t[g.L(s[2], s[1], s[0])];
    } else {
        return;
    }
    t[gi] = out;
}

int maxdim(const Geo &g) { return std::max(g.mx, std::max(g.my, g.mz)); }

PatchSet patches(const int *bc, const double *vals, int nvals) {
    PatchSet P{};
    for (int q = 0; q < 6; q++) P.bc[q] = bc[q];
    if (vals)
        for (int q = 0; q < nvals; q++) P.val[q] = vals[q];
    return P;
}

}  // namespace

void launch_contravariant_bcs(double *ucont, const double *ucat, const tosca_mesh *m, const tosca_params *p,
                              const Geo &g, hipStream_t s) {
    const int n = maxdim(g);
    const PatchSet P = patches(p->bc_u, nullptr, 0);
    TOSCA_LAUNCH("k_contra_periodic_copy", k_contra_periodic_copy, dim3((n + 127) / 128, n, 3), dim3(128), 0, s, ucont, g, P);
    TOSCA_LAUNCH("k_contra_bcs", k_contra_bcs, dim3((n + 127) / 128, n, 6), dim3(128), 0, s, ucont, ucat, m->icsi,
                 m->jeta, m->kzet, g, P);
    launch_wrap(ucont, 3, g, s);
}

size_t cabot_table_doubles() { return (size_t)CABOT_NYP + 1 + (size_t)((CABOT_NYP + CABOT_CHUNK - 1) / CABOT_CHUNK); }

void launch_cabot_table(double *tab, hipStream_t s) {
    const long n = CABOT_NYP, nc = (n + CABOT_CHUNK - 1) / CABOT_CHUNK;
    double *sums = tab + n + 1;
    TOSCA_LAUNCH("k_cabot_intervals", k_cabot_intervals, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, tab, n);
    TOSCA_LAUNCH("k_cabot_chunk_sums", k_cabot_chunk_sums, dim3((unsigned)((nc + 63) / 64)), dim3(64), 0, s, tab, n, sums, nc);
    TOSCA_LAUNCH("k_cabot_chunk_offsets", k_cabot_chunk_offsets, dim3(1), dim3(64), 0, s, sums, nc);
    TOSCA_LAUNCH("k_cabot_chunk_apply", k_cabot_chunk_apply, dim3((unsigned)((nc + 63) / 64)), dim3(64), 0, s, tab, n, sums, nc);
}

void launch_cartesian_bcs(double *ucat, const double *u_fixed, double *ustar, const double *cabot_tab,
                          const tosca_inlet *inlet, const tosca_mesh *m, const tosca_params *p, const Geo &g,
                          hipStream_t s) {
    InletArgs In{};
    if (inlet) {
        In.type = inlet->type; In.cent = inlet->cent; In.grnd = inlet->grnd_level; In.ymin = inlet->ymin;
        In.ly = inlet->ly; In.u_tau = inlet->u_tau; In.rough = inlet->roughness; In.h_inv = inlet->h_inv;
        In.amp = inlet->amplitude; In.periods = inlet->periods; In.fc = inlet->fc; In.h_ref = inlet->h_ref;
        for (int d = 0; d < 3; d++) In.udir[d] = inlet->udir[d];
        const double *u = inlet->uref;
        In.uref_mag = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);   // nMag(Uref)
        for (int d = 0; d < 3; d++) In.uref[d] = u[d];
        In.href = inlet->href; In.uprime = inlet->uprime_rms; In.noise = inlet->noise;
        In.plane = inlet->plane; In.n1 = inlet->n1; In.n2 = inlet->n2; In.prds1 = inlet->prds1; In.prds2 = inlet->prds2;
        In.merge1 = inlet->merge1; In.spline = inlet->spline; In.shift2 = inlet->shift2;
        In.top_len = inlet->avg_top_length; In.top_delta = inlet->avg_top_delta;
        for (int n = 0; n < 10; n++) {
            In.top_z[n] = inlet->avg_top_coords[n];
            for (int d = 0; d < 3; d++) In.top_u[n][d] = inlet->avg_top_u[n][d];
        }
        In.cells = inlet->cells; In.cells1 = inlet->cells_1; In.cells2 = inlet->cells_2; In.yids = inlet->y_ids;
        In.w = inlet->weights; In.w1 = inlet->weights_1; In.w2 = inlet->weights_2; In.yw = inlet->y_weights;
    }
    const int n = maxdim(g);
    PatchSet P = patches(p->bc_u, u_fixed, 18);
    for (int q = 0; q < 4; q++) P.wall[q] = p->wall_u[q];
    // the roughness a Cabot patch uses: the right patch of an axis reads the left patch's model
    // when that one is Cabot too (boundary.c:978-979, 1072-1073)
    for (int a = 0; a < 2; a++) {
        const double r = p->wall_u[2 * a] == -1 ? p->wall_rough[2 * a] : p->wall_rough[2 * a + 1];
        P.rough[2 * a] = P.rough[2 * a + 1] = r;
    }
    TOSCA_LAUNCH("k_cart_bcs", k_cart_bcs, dim3((n + 127) / 128, n, 6), dim3(128), 0, s, ucat, m->csi, m->eta, m->zet,
                 m->aj, ustar, cabot_tab, p->nu, g, P, In);
    launch_wrap(ucat, 3, g, s);
}

void launch_temperature_bcs(double *t, const int *bc_t, const double *t_value, const tosca_mesh *m, const Geo &g,
                            hipStream_t s) {
    const int n = maxdim(g);
    const PatchSet P = patches(bc_t, t_value, 6);
    TOSCA_LAUNCH("k_temp_bcs", k_temp_bcs, dim3((n + 127) / 128, n, 6), dim3(128), 0, s, t, m->csi, m->eta, m->zet, m->aj, g, P);
    launch_wrap(t, 1, g, s);
}

}  // namespace tosca
