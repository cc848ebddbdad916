// farm.hip -- the wind-farm body force of uniform actuator disks on gfx950.
//
//   projectBladeForce, uniformADM branch   src/turbines/turbines_force.c:717-932
//     -> k_uadm_project: lsourceFarmCat, the Gaussian-smeared Cartesian body force per unit density
//   bodyForceCartesian2Contravariant       src/turbines/turbines_force.c:1741-1835
//     -> k_body_force_to_cont: sourceFarmCont, the face fluxes of that force
//
// The reference loops rotor t, disk point p, controlled cell c (the interior cells within
// rTip + eps prjNSigma of the rotor centre, turbines_parallel.c:57-121) and adds
// exp(-(r/eps)^2) / (eps^3 pi^1.5) B_p to the cell when r = |x_p - x_c| < eps prjNSigma.  Here one
// thread owns one cell and walks the same (t, p) order, so each cell's sum has the reference's
// association order.  A workgroup first tests its box of cells against each rotor's sphere and
// skips rotors it cannot reach; the points of a reached rotor are read with wave-uniform
// addresses (one cache line serves the wave).  Compiled with FP contraction off.
#include "kernels.h"

namespace tosca {

namespace {

constexpr int PX = 64, PY = 4;     // workgroup: 64 cells along i x 4 along j, one k plane

__global__ void __launch_bounds__(PX * PY) k_uadm_project(const double *cent, tosca_uadm F, double *scat, Geo g) {
    const int i = blockIdx.x * PX + threadIdx.x + 1, j = blockIdx.y * PY + threadIdx.y + 1, k = blockIdx.z + 1;
    const bool own = i <= g.mx - 2 && j <= g.my - 2;
    const int c = g.L(k, min(j, g.my - 2), min(i, g.mx - 2));
    const double cx = cent[3 * c], cy = cent[3 * c + 1], cz = cent[3 * c + 2];
    // bounding box of the workgroup's cell centres (LDS min / max over the workgroup)
    __shared__ double bb[6][PX * PY];
    const int tid = threadIdx.x + PX * threadIdx.y;
    bb[0][tid] = cx; bb[1][tid] = cy; bb[2][tid] = cz;
    bb[3][tid] = cx; bb[4][tid] = cy; bb[5][tid] = cz;
    __syncthreads();
    for (int w = PX * PY / 2; w > 0; w >>= 1) {
        if (tid < w) {
            for (int d = 0; d < 3; d++) {
                bb[d][tid] = fmin(bb[d][tid], bb[d][tid + w]);
                bb[3 + d][tid] = fmax(bb[3 + d][tid], bb[3 + d][tid + w]);
            }
        }
        __syncthreads();
    }
    const double lo[3] = {bb[0][0], bb[1][0], bb[2][0]}, hi[3] = {bb[3][0], bb[4][0], bb[5][0]};
    double sx = 0.0, sy = 0.0, sz = 0.0;
    for (int t = 0; t < F.n_rotors; t++) {
        const double ox = F.center[3 * t], oy = F.center[3 * t + 1], oz = F.center[3 * t + 2];
        const double rad = F.sphere[t];
        // distance from the rotor centre to the box (0 inside): no cell of the box is controlled
        // by rotor t when it exceeds the sphere radius (the box test is conservative; the cell
        // test below decides)
        const double o[3] = {ox, oy, oz};
        double d2 = 0.0;
        for (int d = 0; d < 3; d++) {
            const double e = fmax(fmax(lo[d] - o[d], o[d] - hi[d]), 0.0);
            d2 += e * e;
        }
        if (d2 > rad * rad * 1.0000001) continue;
        // the controlled-cell test of turbines_parallel.c:98-106: |cent - rotor centre| <= radius
        const double dx = cx - ox, dy = cy - oy, dz = cz - oz;
        const bool ctrl = own && sqrt(dx * dx + dy * dy + dz * dz) <= rad;
        const double eps = F.eps[t];
        const double rprj = eps * F.nsigma;
        const double den = pow(eps, 3.0) * pow(M_PI, 1.5);
        const double *P = F.points + 3L * t * F.n_points, *B = F.force + 3L * t * F.n_points;
        if (!__any(ctrl)) continue;
        for (int p = 0; p < F.n_points; p++) {
            const double rx = P[3 * p] - cx, ry = P[3 * p + 1] - cy, rz = P[3 * p + 2] - cz;
            const double r = sqrt(rx * rx + ry * ry + rz * rz);
            if (ctrl && r < rprj) {
                const double pf = exp(-(r / eps) * (r / eps)) / den;
                sx += pf * B[3 * p];
                sy += pf * B[3 * p + 1];
                sz += pf * B[3 * p + 2];
            }
        }
    }
    if (own) {
        scat[3 * c] = sx;
        scat[3 * c + 1] = sy;
        scat[3 * c + 2] = sz;
    }
}

// every local point outside the interior cells to zero (VecSet of the local vector before the
// projection; the type-2 ghost wrap then copies the owned values)
__global__ void k_zero_non_interior3(double *a, Geo g) {
    const long n = (long)g.nloc();
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        const int i = (int)(t % g.gxm) - GW;
        const long r = t / g.gxm;
        const int j = (int)(r % g.gym) - GW, k = (int)(r / g.gym) - GW;
        if (i >= 1 && i <= g.mx - 2 && j >= 1 && j <= g.my - 2 && k >= 1 && k <= g.mz - 2) continue;
        a[3 * t] = 0.0; a[3 * t + 1] = 0.0; a[3 * t + 2] = 0.0;
    }
}

// sourceFarmCont: 0 on every owned point, and on the interior cells (1 .. m-2 on every axis) the
// body force averaged to the cell's +i, +j, +k faces dotted with icsi, jeta, kzet
__global__ void __launch_bounds__(256) k_body_force_to_cont(const double *scat, const double *icsi, const double *jeta,
                                                            const double *kzet, double *scont, Geo g) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, k = blockIdx.z;
    if (i > g.mx - 1) return;
    double vx = 0.0, vy = 0.0, vz = 0.0;
    if (i >= 1 && i <= g.mx - 2 && j >= 1 && j <= g.my - 2 && k >= 1 && k <= g.mz - 2) {
        const int c = g.L(k, j, i);
        const V3 s = ld3(scat, c), si = ld3(scat, g.L(k, j, i + 1)), sj = ld3(scat, g.L(k, j + 1, i)),
                 sk = ld3(scat, g.L(k + 1, j, i));
        const V3 a = ld3(icsi, c), b = ld3(jeta, c), e = ld3(kzet, c);
        const V3 fI{0.5 * (s.x + si.x), 0.5 * (s.y + si.y), 0.5 * (s.z + si.z)};
        vx = (fI.x * a.x + fI.y * a.y + fI.z * a.z);
        const V3 fJ{0.5 * (s.x + sj.x), 0.5 * (s.y + sj.y), 0.5 * (s.z + sj.z)};
        vy = (fJ.x * b.x + fJ.y * b.y + fJ.z * b.z);
        const V3 fK{0.5 * (s.x + sk.x), 0.5 * (s.y + sk.y), 0.5 * (s.z + sk.z)};
        vz = (fK.x * e.x + fK.y * e.y + fK.z * e.z);
    }
    const long o = 3L * g.G(k, j, i);
    scont[o] = vx;
    scont[o + 1] = vy;
    scont[o + 2] = vz;
}

}  // namespace

void launch_uadm_project(const double *cent, const tosca_uadm *F, double *scat, const Geo &g, hipStream_t s) {
    TOSCA_LAUNCH("k_zero_non_interior3", k_zero_non_interior3, dim3(1024), dim3(256), 0, s, scat, g);
    TOSCA_LAUNCH("k_uadm_project", k_uadm_project, dim3((g.mx - 2 + PX - 1) / PX, (g.my - 2 + PY - 1) / PY, g.mz - 2),
                 dim3(PX, PY), 0, s, cent, *F, scat, g);
    launch_wrap(scat, 3, g, s);
}

void launch_body_force_to_cont(const double *scat, const tosca_mesh *m, double *scont, const Geo &g, hipStream_t s) {
    TOSCA_LAUNCH("k_body_force_to_cont", k_body_force_to_cont, dim3((g.mx + 255) / 256, g.my, g.mz), dim3(256), 0, s,
                 scat, m->icsi, m->jeta, m->kzet, scont, g);
}

}  // namespace tosca
