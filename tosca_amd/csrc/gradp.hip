// gradp.hip -- GradP (src/peqn.c:3437-3773), the contravariant pressure-gradient fluxes dP that
// ExplicitRhsU subtracts, on gfx950.  Compiled with FP contraction off: each flux is evaluated in
// the reference's expression order.
//
// One thread per owned point writes all three components of dP (global vector): the fluxes of the
// faces the reference loops over (interior indices), zero on every other point, which stands for
// its VecSet(dP, 0).
#include "kernels.h"

namespace tosca {

namespace {

__device__ __forceinline__ double dot3(const V3 &a, const V3 &b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

struct GradPArgs {
    const double *p;
    const double *icsi, *ieta, *izet, *jcsi, *jeta, *jzet, *kcsi, *keta, *kzet;
    const double *iaj, *jaj, *kaj;
    double *dp;
};

__global__ void k_grad_p(GradPArgs A, Geo g) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y * blockDim.y + threadIdx.y, k = blockIdx.z;
    if (i >= g.mx || j >= g.my) return;
    const int mx = g.mx, my = g.my, mz = g.mz;
    V3 r{0.0, 0.0, 0.0};
    if (i >= 1 && i <= mx - 2 && j >= 1 && j <= my - 2 && k >= 1 && k <= mz - 2) {
        const long c = g.L(k, j, i);
        const double *P = A.p;
        auto p = [&](int kk, int jj, int ii) { return P[g.L(kk, jj, ii)]; };
        const V3 ic = ld3(A.icsi, c), ie = ld3(A.ieta, c), iz = ld3(A.izet, c);
        const V3 jc = ld3(A.jcsi, c), je = ld3(A.jeta, c), jz = ld3(A.jzet, c);
        const V3 kc = ld3(A.kcsi, c), ke = ld3(A.keta, c), kz = ld3(A.kzet, c);
        const double g11_i = dot3(ic, ic), g12_i = dot3(ie, ic), g13_i = dot3(iz, ic);
        const double g21_j = dot3(jc, je), g22_j = dot3(je, je), g23_j = dot3(jz, je);
        const double g31_k = dot3(kc, kz), g32_k = dot3(ke, kz), g33_k = dot3(kz, kz);
        double dpdc, dpde, dpdz;
        // i face: the one-sided corner rows sit at i == mx-2 whatever the periodicity (:3502-3566)
        dpdc = (i == mx - 2 && g.per[0] == TOSCA_PERIODIC_WRAP) ? p(k, j, mx + 1) - p(k, j, i) : p(k, j, i + 1) - p(k, j, i);
        if (j == my - 2 && i == mx - 2) dpde = (p(k, j, i) - p(k, j - 1, i) + p(k, j, i + 1) - p(k, j - 1, i + 1)) * 0.5;
        else if (j == 1 && i == mx - 2) dpde = (p(k, j + 1, i) - p(k, j, i) + p(k, j + 1, i + 1) - p(k, j, i + 1)) * 0.5;
        else dpde = (p(k, j + 1, i) - p(k, j - 1, i) + p(k, j + 1, i + 1) - p(k, j - 1, i + 1)) * 0.25;
        if (k == mz - 2 && i == mx - 2) dpdz = (p(k, j, i) - p(k - 1, j, i) + p(k, j, i + 1) - p(k - 1, j, i + 1)) * 0.5;
        else if (k == 1 && i == mx - 2) dpdz = (p(k + 1, j, i) - p(k, j, i) + p(k + 1, j, i + 1) - p(k, j, i + 1)) * 0.5;
        else dpdz = (p(k + 1, j, i) - p(k - 1, j, i) + p(k + 1, j, i + 1) - p(k - 1, j, i + 1)) * 0.25;
        r.x = (dpdc * g11_i + dpde * g12_i + dpdz * g13_i) * A.iaj[c];
        // j face (:3568-3651)
        if (i == mx - 2 && j == my - 2) dpdc = (p(k, j, i) - p(k, j, i - 1) + p(k, j + 1, i) - p(k, j + 1, i - 1)) * 0.5;
        else if (i == 1 && j == my - 2) dpdc = (p(k, j, i + 1) - p(k, j, i) + p(k, j + 1, i + 1) - p(k, j + 1, i)) * 0.5;
        else dpdc = (p(k, j, i + 1) - p(k, j, i - 1) + p(k, j + 1, i + 1) - p(k, j + 1, i - 1)) * 0.25;
        dpde = (j == my - 2 && g.per[1] == TOSCA_PERIODIC_WRAP) ? p(k, my + 1, i) - p(k, j, i) : p(k, j + 1, i) - p(k, j, i);
        if (k == mz - 2 && j == my - 2) dpdz = (p(k, j, i) - p(k - 1, j, i) + p(k, j + 1, i) - p(k - 1, j + 1, i)) * 0.5;
        else if (k == 1 && j == my - 2) dpdz = (p(k + 1, j, i) - p(k, j, i) + p(k + 1, j + 1, i) - p(k, j + 1, i)) * 0.5;
        else dpdz = (p(k + 1, j, i) - p(k - 1, j, i) + p(k + 1, j + 1, i) - p(k - 1, j + 1, i)) * 0.25;
        r.y = (dpdc * g21_j + dpde * g22_j + dpdz * g23_j) * A.jaj[c];
        // k face (:3653-3734)
        if (i == mx - 2 && k == mz - 2) dpdc = (p(k, j, i) - p(k, j, i - 1) + p(k + 1, j, i) - p(k + 1, j, i - 1)) * 0.5;
        else if (i == 1 && k == mz - 2) dpdc = (p(k, j, i + 1) - p(k, j, i) + p(k + 1, j, i + 1) - p(k + 1, j, i)) * 0.5;
        else dpdc = (p(k, j, i + 1) - p(k, j, i - 1) + p(k + 1, j, i + 1) - p(k + 1, j, i - 1)) * 0.25;
        if (j == my - 2 && k == mz - 2) dpde = (p(k, j, i) - p(k, j - 1, i) + p(k + 1, j, i) - p(k + 1, j - 1, i)) * 0.5;
        else if (j == 1 && k == mz - 2) dpde = (p(k, j + 1, i) - p(k, j, i) + p(k + 1, j + 1, i) - p(k + 1, j, i)) * 0.5;
        else dpde = (p(k, j + 1, i) - p(k, j - 1, i) + p(k + 1, j + 1, i) - p(k + 1, j - 1, i)) * 0.25;
        dpdz = (k == mz - 2 && g.per[2] == TOSCA_PERIODIC_WRAP) ? p(mz + 1, j, i) - p(k, j, i) : (p(k + 1, j, i) - p(k, j, i));
        r.z = (dpdc * g31_k + dpde * g32_k + dpdz * g33_k) * A.kaj[c];
        // faces whose contravariant velocity is not solved (:3736-3757)
        if (g.nonper(0) && i == mx - 2) r.x = 0.0;
        if (g.nonper(1) && j == my - 2) r.y = 0.0;
        if (g.nonper(2) && k == mz - 2) r.z = 0.0;
    }
    st3(A.dp, g.G(k, j, i), r);
}

}  // namespace

void launch_grad_p(const double *p, const tosca_mesh *m, double *dp, const Geo &g, hipStream_t s) {
    GradPArgs A;
    A.p = p;
    A.icsi = m->icsi; A.ieta = m->ieta; A.izet = m->izet;
    A.jcsi = m->jcsi; A.jeta = m->jeta; A.jzet = m->jzet;
    A.kcsi = m->kcsi; A.keta = m->keta; A.kzet = m->kzet;
    A.iaj = m->iaj; A.jaj = m->jaj; A.kaj = m->kaj;
    A.dp = dp;
    TOSCA_LAUNCH("k_grad_p", k_grad_p, dim3((g.mx + 63) / 64, (g.my + 3) / 4, g.mz), dim3(64, 4), 0, s, A, g);
}

}  // namespace tosca
