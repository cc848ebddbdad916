// krylov.hip -- vector kernels of the restarted GMRES that solves the IMEX momentum system
// (kspIMEX, src/ueqn.c:261-282) on gfx950, and the vector updates of the IMEX Runge-Kutta stages
// (src/solvers/ueqn_solvers.c:1508-1661).  Compiled with FP contraction off: every update is
// evaluated per element in the order PETSc's VecAXPY / VecMAXPY / VecScale define, so the
// element-wise results equal the CPU restatement's; only the dot products are summed in another
// (fixed, deterministic) association.
//
// Dot products: k_mdot streams w once per group of up to 8 basis vectors and writes one partial
// sum per workgroup and vector; k_mdot_finish adds the partials of each vector in workgroup order.
#include "kernels.h"

namespace tosca {

namespace {

constexpr int MDOT_BLOCKS = 1024, MDOT_THREADS = 256, MDOT_GROUP = 8;

// partial[blk * nv + j] = sum over this workgroup's elements of V_j[t] w[t], j in [j0, j0 + cnt)
template <int CNT>
__global__ void __launch_bounds__(MDOT_THREADS) k_mdot(const double *V, long stride, int j0, int nv, const double *w,
                                                       long n, double *partial) {
    __shared__ double red[CNT][MDOT_THREADS];
    double acc[CNT];
#pragma unroll
    for (int q = 0; q < CNT; q++) acc[q] = 0.0;
    for (long t = (long)blockIdx.x * MDOT_THREADS + threadIdx.x; t < n; t += (long)gridDim.x * MDOT_THREADS) {
        const double wt = w[t];
#pragma unroll
        for (int q = 0; q < CNT; q++) acc[q] = acc[q] + V[(long)(j0 + q) * stride + t] * wt;
    }
#pragma unroll
    for (int q = 0; q < CNT; q++) red[q][threadIdx.x] = acc[q];
    __syncthreads();
    for (int h = MDOT_THREADS / 2; h > 0; h >>= 1) {
        if (threadIdx.x < h) {
#pragma unroll
            for (int q = 0; q < CNT; q++) red[q][threadIdx.x] = red[q][threadIdx.x] + red[q][threadIdx.x + h];
        }
        __syncthreads();
    }
    if (threadIdx.x < CNT) partial[(long)blockIdx.x * nv + j0 + threadIdx.x] = red[threadIdx.x][0];
}

// out[j] = sum over workgroups of partial[blk * nv + j], in workgroup order (one thread per j)
__global__ void k_mdot_finish(const double *partial, int nblk, int nv, double *out) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nv) return;
    double s = 0.0;
    for (int b = 0; b < nblk; b++) s = s + partial[(long)b * nv + j];
    out[j] = s;
}

// VecMAXPY: w = w + sum_j (sign h_j) V_j, j ascending, per element
__global__ void k_maxpy(double *w, const double *V, long stride, int nv, const double *h, double sign, long n) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    double x = w[t];
    for (int j = 0; j < nv; j++) x = x + (sign * h[j]) * V[(long)j * stride + t];
    w[t] = x;
}

// y = a * x (VecScale of a copy, or in place when y == x)
__global__ void k_scale_to(double *y, const double *x, double a, long n) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    y[t] = x[t] * a;
}

// r = b + (-1) t (KSPInitialResidual: VecCopy(b, r); VecAXPY(r, -1, A x))
__global__ void k_residual(double *r, const double *b, const double *t_, long n) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    r[t] = b[t] + (-1.0) * t_[t];
}

// the owned range of a local vector field into a global vector
__global__ void k_local_to_global(double *glob, const double *loc, Geo g, long n) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const long q = t / 3;
    const int i = (int)(q % g.mx);
    const long r = q / g.mx;
    const int j = (int)(r % g.my), k = (int)(r / g.my);
    glob[t] = loc[3 * (long)g.L(k, j, i) + t % 3];
}

unsigned blocks(long n) { return (unsigned)((n + 255) / 256); }

}  // namespace

void launch_mdot(const double *V, long stride, int nv, const double *w, long n, double *partial, double *out,
                 hipStream_t s) {
    for (int j0 = 0; j0 < nv; j0 += MDOT_GROUP) {
        const int cnt = min(MDOT_GROUP, nv - j0);
        switch (cnt) {
            case 1: TOSCA_LAUNCH("k_mdot", k_mdot<1>, dim3(MDOT_BLOCKS), dim3(MDOT_THREADS), 0, s, V, stride, j0, nv, w, n, partial); break;
            case 2: TOSCA_LAUNCH("k_mdot", k_mdot<2>, dim3(MDOT_BLOCKS), dim3(MDOT_THREADS), 0, s, V, stride, j0, nv, w, n, partial); break;
            case 3: TOSCA_LAUNCH("k_mdot", k_mdot<3>, dim3(MDOT_BLOCKS), dim3(MDOT_THREADS), 0, s, V, stride, j0, nv, w, n, partial); break;
            case 4: TOSCA_LAUNCH("k_mdot", k_mdot<4>, dim3(MDOT_BLOCKS), dim3(MDOT_THREADS), 0, s, V, stride, j0, nv, w, n, partial); break;
            case 5: TOSCA_LAUNCH("k_mdot", k_mdot<5>, dim3(MDOT_BLOCKS), dim3(MDOT_THREADS), 0, s, V, stride, j0, nv, w, n, partial); break;
            case 6: TOSCA_LAUNCH("k_mdot", k_mdot<6>, dim3(MDOT_BLOCKS), dim3(MDOT_THREADS), 0, s, V, stride, j0, nv, w, n, partial); break;
            case 7: TOSCA_LAUNCH("k_mdot", k_mdot<7>, dim3(MDOT_BLOCKS), dim3(MDOT_THREADS), 0, s, V, stride, j0, nv, w, n, partial); break;
            default: TOSCA_LAUNCH("k_mdot", k_mdot<8>, dim3(MDOT_BLOCKS), dim3(MDOT_THREADS), 0, s, V, stride, j0, nv, w, n, partial); break;
        }
    }
    TOSCA_LAUNCH("k_mdot_finish", k_mdot_finish, dim3((nv + 63) / 64), dim3(64), 0, s, partial, MDOT_BLOCKS, nv, out);
}

size_t mdot_partial_doubles(int nvmax) { return (size_t)MDOT_BLOCKS * nvmax; }

void launch_maxpy(double *w, const double *V, long stride, int nv, const double *h, double sign, long n, hipStream_t s) {
    TOSCA_LAUNCH("k_maxpy", k_maxpy, dim3(blocks(n)), dim3(256), 0, s, w, V, stride, nv, h, sign, n);
}

void launch_scale_to(double *y, const double *x, double a, long n, hipStream_t s) {
    TOSCA_LAUNCH("k_scale_to", k_scale_to, dim3(blocks(n)), dim3(256), 0, s, y, x, a, n);
}

void launch_residual(double *r, const double *b, const double *t, long n, hipStream_t s) {
    TOSCA_LAUNCH("k_residual", k_residual, dim3(blocks(n)), dim3(256), 0, s, r, b, t, n);
}

void launch_local_to_global(double *glob, const double *loc, const Geo &g, hipStream_t s) {
    const long n = 3L * g.mx * g.my * g.mz;
    TOSCA_LAUNCH("k_local_to_global", k_local_to_global, dim3(blocks(n)), dim3(256), 0, s, glob, loc, g, n);
}

}  // namespace tosca
