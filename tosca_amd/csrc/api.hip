// api.hip -- the C ABI (include/tosca_rhs.h): context, argument validation and dispatch.
//
// Every entry point validates its arguments on the host before any kernel is queued: enums,
// layout, and that every array is device memory large enough for the DMDA layout it is read or
// written with (hipMemGetAddressRange), so a wrong size or a host pointer is an error code, not
// a GPU memory fault.  Unsupported reference features (IBM, overset, structural LES models,
// schemes other than central/QUICK) are rejected with TOSCA_ERR_UNSUPPORTED.
#include <cstdarg>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <algorithm>
#include <array>
#include <iterator>
#include <vector>
#include <type_traits>
#include <mutex>

#include "fused.h"

using tosca::Geo;

struct tosca_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    tosca_dmda_info info{};
    Geo geo{};
    std::string err;
    double *ws = nullptr;
    size_t ws_elems = 0;
    int *chk = nullptr;
    size_t chk_elems = 0;
    double *kws = nullptr;       // Krylov basis and the IMEX predictors' work vectors
    size_t kws_elems = 0;
    double *hs = nullptr;        // device staging of the host-resident step
    size_t hs_elems = 0;
    double *cabot = nullptr;     // the Cabot wall law's integration table, built on first use
    double *adj = nullptr;       // the predictor adjoint's stage checkpoints and work vectors
    size_t adj_elems = 0;
    // meshes tosca_mesh_check accepted (their nvert, meshtag, aj, csi, eta, zet pointers): the
    // kernels do not read nvert / meshTag, so every entry point that takes a mesh requires one
    std::vector<std::array<const void *, 6>> checked;
    // the checked meshes (with their icsi, jeta, kzet pointers) whose cell-centred csi, eta, zet are
    // the averages of the face bases on their two faces, bit for bit: FormU forms them itself
    std::vector<std::array<const void *, 9>> derivable;
    tosca::KernelTimer timer;
    bool ref_path = false;   // TOSCA_PATH=ref at creation: reference-structured kernels only
    double *scratch = nullptr;   // this device's copy of `scratch` (tosca_formu_noalloc), or null
    std::string scratch_err;     // why the lookup of that copy failed
};

namespace tosca {
thread_local KernelTimer *g_timer = nullptr;
}

// Scratch of tosca_formu_noalloc: FormU's boundary-face fluxes, periodic-seam fp and j-tile seam
// halves, in a file-scope device array instead of the context's workspace.  Every GPU the library
// runs on holds its own copy (TOSCA_NOALLOC_SCRATCH_DOUBLES doubles) until the process ends.  Its
// host-side name is not a device address: host code uses only the per-device address that
// hipGetSymbolAddress returned (g_scratch_dev), never `scratch` itself.
static __device__ __align__(256) double scratch[TOSCA_NOALLOC_SCRATCH_DOUBLES];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_dev[kMaxDevices] = {};   // device address of `scratch` per device, set once

// Looks up this device's copy of `scratch` (the current device must be `device`); once per device.
hipError_t scratch_setup(int device, double **out) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    if (!g_scratch_dev[device]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) { (void)hipGetLastError(); return e; }
        g_scratch_dev[device] = static_cast<double *>(ptr);
    }
    *out = g_scratch_dev[device];
    return hipSuccess;
}

}  // namespace

namespace {

thread_local std::string g_err;   // last error of calls made without a usable context

int set_err(tosca_ctx *c, int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (c) c->err = buf;
    g_err = buf;
    return code;
}

bool device_forced_off() {
    const char *e = getenv("TOSCA_FORCE_NO_DEVICE");
    return e && e[0] && strcmp(e, "0") != 0;
}

// `need` bytes starting at p must lie inside one device allocation; p aligned to `align` bytes (the
// element size: 8 for double arrays, 4 for int and float arrays)
int check_dev(tosca_ctx *c, const void *p, size_t need, const char *name, size_t align = 8) {
    if (!p) return set_err(c, TOSCA_ERR_ARG, "%s is NULL", name);
    if (reinterpret_cast<uintptr_t>(p) % align)
        return set_err(c, TOSCA_ERR_ARG, "%s is not %zu-byte aligned", name, align);
    hipPointerAttribute_t attr;
    if (hipPointerGetAttributes(&attr, p) != hipSuccess) {
        (void)hipGetLastError();
        return set_err(c, TOSCA_ERR_ARG, "%s is not a device pointer", name);
    }
    if (attr.type != hipMemoryTypeDevice && attr.type != hipMemoryTypeManaged)
        return set_err(c, TOSCA_ERR_ARG, "%s is not device memory", name);
    hipDeviceptr_t base = nullptr;
    size_t size = 0;
    if (hipMemGetAddressRange(&base, &size, const_cast<void *>(p)) != hipSuccess || !base) {
        (void)hipGetLastError();
        return set_err(c, TOSCA_ERR_ARG, "%s: the allocation holding it is unknown to the HIP runtime, so its size "
                                         "cannot be checked", name);
    }
    const char *b = reinterpret_cast<const char *>(base), *q = reinterpret_cast<const char *>(p);
    if (q + need > b + size)
        return set_err(c, TOSCA_ERR_ARG, "%s: allocation too small (%zu bytes needed past the pointer, %zu available)",
                       name, need, (size_t)(b + size - q));
    return TOSCA_OK;
}

std::array<const void *, 6> mesh_key(const tosca_mesh *m) {
    return {m->nvert, m->meshtag, m->aj, m->csi, m->eta, m->zet};
}

// the mesh must have passed tosca_mesh_check on this context
int require_checked(tosca_ctx *c, const tosca_mesh *m) {
    if (!m) return set_err(c, TOSCA_ERR_ARG, "mesh is NULL");
    const auto key = mesh_key(m);
    for (const auto &k : c->checked)
        if (k == key) return TOSCA_OK;
    return set_err(c, TOSCA_ERR_MESH, "the mesh has not passed tosca_mesh_check on this context (its nvert / meshTag "
                                      "are validated there once, not read per call)");
}

std::array<const void *, 9> derive_key(const tosca_mesh *m) {
    return {m->nvert, m->meshtag, m->aj, m->csi, m->eta, m->zet, m->icsi, m->jeta, m->kzet};
}

// 1 when tosca_mesh_check found this mesh's cell bases to be the face averages
int cc_derivable(const tosca_ctx *c, const tosca_mesh *m) {
    const auto key = derive_key(m);
    for (const auto &k : c->derivable)
        if (k == key) return 1;
    return 0;
}

int check_ctx(tosca_ctx *c) {
    if (!c) return set_err(nullptr, TOSCA_ERR_ARG, "context is NULL");
    if (hipSetDevice(c->device) != hipSuccess) return set_err(c, TOSCA_ERR_DEVICE, "hipSetDevice(%d) failed", c->device);
    tosca::g_timer = &c->timer;
    return TOSCA_OK;
}

int check_params(tosca_ctx *c, const tosca_params *p) {
    if (!p) return set_err(c, TOSCA_ERR_ARG, "params is NULL");
    for (int a = 0; a < 3; a++)
        if (p->periodic[a] < TOSCA_PERIODIC_NONE || p->periodic[a] > TOSCA_PERIODIC_WRAP)
            return set_err(c, TOSCA_ERR_ARG, "periodic[%d] = %d is not a periodic type", a, p->periodic[a]);
    for (int q = 0; q < 6; q++) {
        if (p->bc_u[q] < TOSCA_BCU_GENERIC || p->bc_u[q] > TOSCA_BCU_FIXEDVALUE)
            return set_err(c, TOSCA_ERR_ARG, "bc_u[%d] = %d is not a TOSCA_BCU_* kind", q, p->bc_u[q]);
        if (p->bc_nut[q] < TOSCA_BCNUT_NONE || p->bc_nut[q] > TOSCA_BCNUT_PERIODIC)
            return set_err(c, TOSCA_ERR_ARG, "bc_nut[%d] = %d is not a TOSCA_BCNUT_* kind", q, p->bc_nut[q]);
    }
    if (p->ibm_active) return set_err(c, TOSCA_ERR_UNSUPPORTED, "immersed boundaries (IBM) are not supported");
    if (p->overset_active) return set_err(c, TOSCA_ERR_UNSUPPORTED, "overset meshes are not supported");
    return TOSCA_OK;
}

int check_les_model(tosca_ctx *c, const tosca_params *p) {
    switch (p->les_model) {
        case TOSCA_LES_SMAGORINSKY: case TOSCA_LES_DSM: case TOSCA_LES_DPASD: return TOSCA_OK;
        case TOSCA_LES_BV: case TOSCA_LES_BAMD:
            return set_err(c, TOSCA_ERR_UNSUPPORTED, "structural LES models (bardinaVreman/bardinaAMD) are not supported");
        default:
            if (p->les_model < 0 || p->les_model > TOSCA_LES_BAMD)
                return set_err(c, TOSCA_ERR_ARG, "les_model = %d is not an LES model", p->les_model);
            return set_err(c, TOSCA_ERR_UNSUPPORTED, "LES model %d is not supported (smagorinsky, dynamicSmagorinsky, dynamicPASD are)", p->les_model);
    }
}

int check_mesh(tosca_ctx *c, const tosca_mesh *m, bool face, bool cell, bool jac, bool need_checked = true) {
    if (!m) return set_err(c, TOSCA_ERR_ARG, "mesh is NULL");
    const size_t n = (size_t)c->geo.nloc(), v = 3 * n * sizeof(double), s = n * sizeof(double);
    int rc;
    if (need_checked && (rc = require_checked(c, m))) return rc;
    if (cell) {
        if ((rc = check_dev(c, m->csi, v, "mesh.csi")) || (rc = check_dev(c, m->eta, v, "mesh.eta")) ||
            (rc = check_dev(c, m->zet, v, "mesh.zet")) || (rc = check_dev(c, m->aj, s, "mesh.aj")))
            return rc;
    }
    if (face) {
        const double *vv[9] = {m->icsi, m->ieta, m->izet, m->jcsi, m->jeta, m->jzet, m->kcsi, m->keta, m->kzet};
        const char *nm[9] = {"mesh.icsi", "mesh.ieta", "mesh.izet", "mesh.jcsi", "mesh.jeta", "mesh.jzet", "mesh.kcsi", "mesh.keta", "mesh.kzet"};
        for (int q = 0; q < 9; q++)
            if ((rc = check_dev(c, vv[q], v, nm[q]))) return rc;
    }
    if (jac) {
        if ((rc = check_dev(c, m->iaj, s, "mesh.iaj")) || (rc = check_dev(c, m->jaj, s, "mesh.jaj")) ||
            (rc = check_dev(c, m->kaj, s, "mesh.kaj")))
            return rc;
    }
    return TOSCA_OK;
}

int ensure_ws(tosca_ctx *c, size_t elems) {
    if (c->ws_elems >= elems) return TOSCA_OK;
    if (c->ws) { hipStreamSynchronize(c->stream); hipFree(c->ws); c->ws = nullptr; c->ws_elems = 0; }
    if (hipMalloc(&c->ws, elems * sizeof(double)) != hipSuccess) {
        (void)hipGetLastError();
        return set_err(c, TOSCA_ERR_DEVICE, "workspace allocation of %zu bytes failed", elems * sizeof(double));
    }
    c->ws_elems = elems;
    return TOSCA_OK;
}

int launched(tosca_ctx *c, const char *what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_err(c, TOSCA_ERR_DEVICE, "%s: %s", what, hipGetErrorString(e));
    return TOSCA_OK;
}

void set_geo_per(tosca_ctx *c, const tosca_params *p) {
    for (int a = 0; a < 3; a++) c->geo.per[a] = p->periodic[a];
}

tosca::FormUArgs formu_args(const tosca_ctx *c, const tosca_mesh *m, const tosca_params *p, const double *ucont,
                            const double *ucat, const double *nut, double *rhs, double scale, int mode) {
    tosca::FormUArgs A{};
    A.derive_cc = cc_derivable(c, m);
    A.ucont = ucont; A.ucat = ucat; A.nut = nut;
    A.csi = m->csi; A.eta = m->eta; A.zet = m->zet;
    A.icsi = m->icsi; A.ieta = m->ieta; A.izet = m->izet;
    A.jcsi = m->jcsi; A.jeta = m->jeta; A.jzet = m->jzet;
    A.kcsi = m->kcsi; A.keta = m->keta; A.kzet = m->kzet;
    A.iaj = m->iaj; A.jaj = m->jaj; A.kaj = m->kaj;
    A.tauI = p->tau_wall_i; A.tauJ = p->tau_wall_j;
    A.rhs = rhs; A.scale = scale; A.mode = mode;
    A.quick = p->div_scheme == TOSCA_DIV_QUICK;
    A.les = nut != nullptr;
    A.nu = p->nu;
    for (int q = 0; q < 6; q++) A.bc_u[q] = p->bc_u[q];
    A.theta = nullptr; A.btheta = nullptr; A.bscale = 0.0; A.tRef = 1.0;
    return A;
}

size_t ws_need(const tosca_ctx *c, const tosca_params *p) {
    const size_t n = (size_t)c->geo.nloc();
    // reference-structured path: FormU's 6 face-flux vectors + fp (21 local fields), UpdateCs's
    // strain rows, LM/MM/QN/NN and the DPASD plane partials
    if (c->ref_path) return 22 * n + (size_t)c->geo.my * 16 * 4 + 64;
    // fused path: LM, MM, |S| then FormU's boundary-face and seam buffers; FormT's two face-flux
    // vectors; the DPASD filter sums when that model runs
    size_t need = std::max(3 * n + tosca::formu_ws_doubles(c->geo) + 64, 6 * n + 64);
    if (p && p->les_active && p->les_model == TOSCA_LES_DPASD) need = std::max(need, tosca::dpasd_ws_doubles(c->geo));
    return need;
}

// The fused UpdateCs -> UpdateNut -> Buoyancy -> FormU sequence on the context's stream, for
// double (FP64 path) or float (mixed path) flow fields; arguments are validated by the caller.
template <class T, class MT = double>
void queue_step(tosca_ctx *c, const tosca_mesh *m, const tosca_params *p, const T *ucont, const T *ucat,
                const T *theta, double *rhs, double scale, double *cs, T *nut, double *btheta, double buoy_scale,
                const tosca_mesh_f32 *m32 = nullptr, const double *farm = nullptr, double *ws = nullptr) {
    const Geo &g = c->geo;
    hipStream_t s = c->stream;
    if (!ws) ws = c->ws;
    // UpdateCs (+ the interior of UpdateNut)
    if (p->les_model == TOSCA_LES_DSM) {
        // with FormU on the FP32 metric copies, the filter kernel reads their csi, eta, zet too
        tosca::fused_dsm<T>(g, m, ucat, p->max_cs, cs, nut, ws, s, std::is_same<MT, float>::value ? m32 : nullptr);
    } else if (p->les_model == TOSCA_LES_DPASD) {
        tosca::fused_dpasd<T>(g, m, ucat, p->max_cs, cs, nut, ws, s);
    } else {
        tosca::launch_fill(cs, (long)g.nloc(), 0.0289, s);   // VecSet(lCs, std_cs), les.c:172-179
        tosca::fused_nut<T>(g, m, ucat, cs, nut, nullptr, s);
    }
    if (p->les_model == TOSCA_LES_DSM || p->les_model == TOSCA_LES_DPASD) {
        tosca::launch_wrap(cs, 1, g, s);
        tosca::launch_reset_cell_periodic(cs, 1, g, s);
    }
    // UpdateNut: zero outside the interior, boundary conditions, ghost wrap
    tosca::zero_outside_interior(nut, g, s);
    tosca::nut_bcs(nut, p, g, s);
    // Buoyancy + FormU
    tosca::FormUArgsT<T, MT> A{};
    A.ucont = ucont; A.ucat = ucat; A.nut = nut;
    if constexpr (std::is_same<MT, float>::value) {
        A.csi = m32->csi; A.eta = m32->eta; A.zet = m32->zet;
        A.icsi = m32->icsi; A.ieta = m32->ieta; A.izet = m32->izet;
        A.jcsi = m32->jcsi; A.jeta = m32->jeta; A.jzet = m32->jzet;
        A.kcsi = m32->kcsi; A.keta = m32->keta; A.kzet = m32->kzet;
        A.iaj = m32->iaj; A.jaj = m32->jaj; A.kaj = m32->kaj;
    } else {
        A.csi = m->csi; A.eta = m->eta; A.zet = m->zet;
        A.icsi = m->icsi; A.ieta = m->ieta; A.izet = m->izet;
        A.jcsi = m->jcsi; A.jeta = m->jeta; A.jzet = m->jzet;
        A.kcsi = m->kcsi; A.keta = m->keta; A.kzet = m->kzet;
        A.iaj = m->iaj; A.jaj = m->jaj; A.kaj = m->kaj;
    }
    A.tauI = p->tau_wall_i; A.tauJ = p->tau_wall_j;
    A.rhs = rhs; A.farm = farm; A.scale = scale; A.mode = 0;
    A.quick = p->div_scheme == TOSCA_DIV_QUICK;
    A.les = 1;
    A.nu = p->nu;
    for (int q = 0; q < 6; q++) A.bc_u[q] = p->bc_u[q];
    A.theta = theta;
    A.btheta = btheta;
    A.bscale = buoy_scale;
    A.tRef = p->t_ref;
    // FP32 metric copies: the cell bases are loaded (forming them there measured slower)
    A.derive_cc = std::is_same<MT, double>::value ? cc_derivable(c, m) : 0;
    tosca::fused_formu<T, MT>(g, A, ws + 3 * (size_t)g.nloc(), s);
}

// validation shared by the two mixed-precision steps (FP32 flow fields)
int check_mixed(tosca_ctx *c, const tosca_mesh *m, const tosca_params *p, const float *ucont, const float *ucat,
                const float *theta, double *rhs, double *cs, float *nut, double *btheta) {
    int rc = check_ctx(c);
    if (rc) return rc;
    if ((rc = check_params(c, p))) return rc;
    if (!p->les_active) return set_err(c, TOSCA_ERR_ARG, "tosca_rhs_les_step_mixed needs les_active = 1");
    if (p->les_model != TOSCA_LES_SMAGORINSKY && p->les_model != TOSCA_LES_DSM && p->les_model != TOSCA_LES_DPASD)
        return set_err(c, TOSCA_ERR_UNSUPPORTED, "tosca_rhs_les_step_mixed: LES model %d (supported: Smagorinsky 0, "
                                                 "dynamic Smagorinsky 2, plane-averaged dynamic 5)", p->les_model);
    if (p->div_scheme != TOSCA_DIV_CENTRAL && p->div_scheme != TOSCA_DIV_QUICK)
        return set_err(c, TOSCA_ERR_UNSUPPORTED, "div_scheme = %d: only central (0) and QUICK (1) are supported", p->div_scheme);
    if (!(p->t_ref != 0.0)) return set_err(c, TOSCA_ERR_ARG, "t_ref must be non-zero");
    set_geo_per(c, p);
    const Geo &g = c->geo;
    const size_t n = (size_t)g.nloc(), v = 3 * n * sizeof(float), ng = 3 * (size_t)g.mx * g.my * g.mz * sizeof(double);
    if ((rc = check_mesh(c, m, true, true, true))) return rc;
    if ((rc = check_dev(c, ucont, v, "ucont")) || (rc = check_dev(c, ucat, v, "ucat")) ||
        (rc = check_dev(c, theta, n * sizeof(float), "theta")) || (rc = check_dev(c, rhs, ng, "rhs")) ||
        (rc = check_dev(c, cs, n * sizeof(double), "cs")) || (rc = check_dev(c, nut, n * sizeof(float), "nut")) ||
        (rc = check_dev(c, btheta, ng, "btheta")))
        return rc;
    if (p->bc_u[0] == TOSCA_BCU_WALLFUNCTION || p->bc_u[1] == TOSCA_BCU_WALLFUNCTION)
        if ((rc = check_dev(c, p->tau_wall_i, 3 * (size_t)g.mz * g.my * sizeof(double), "tau_wall_i"))) return rc;
    if (p->bc_u[2] == TOSCA_BCU_WALLFUNCTION || p->bc_u[3] == TOSCA_BCU_WALLFUNCTION)
        if ((rc = check_dev(c, p->tau_wall_j, 3 * (size_t)g.mz * g.mx * sizeof(double), "tau_wall_j"))) return rc;
    if ((rc = ensure_ws(c, ws_need(c, p)))) return rc;
    return TOSCA_OK;
}

// the FP32 metric copies of the _m32 steps
int check_mesh32(tosca_ctx *c, const tosca_mesh_f32 *m32) {
    if (!m32) return set_err(c, TOSCA_ERR_ARG, "mesh32 is NULL");
    int rc;
    const size_t n = (size_t)c->geo.nloc(), v = 3 * n * sizeof(float), sc = n * sizeof(float);
    const float *vv[12] = {m32->csi, m32->eta, m32->zet, m32->icsi, m32->ieta, m32->izet,
                           m32->jcsi, m32->jeta, m32->jzet, m32->kcsi, m32->keta, m32->kzet};
    const char *nm[12] = {"mesh32.csi", "mesh32.eta", "mesh32.zet", "mesh32.icsi", "mesh32.ieta", "mesh32.izet",
                          "mesh32.jcsi", "mesh32.jeta", "mesh32.jzet", "mesh32.kcsi", "mesh32.keta", "mesh32.kzet"};
    for (int q = 0; q < 12; q++)
        if ((rc = check_dev(c, vv[q], v, nm[q]))) return rc;
    if ((rc = check_dev(c, m32->iaj, sc, "mesh32.iaj")) || (rc = check_dev(c, m32->jaj, sc, "mesh32.jaj")) ||
        (rc = check_dev(c, m32->kaj, sc, "mesh32.kaj")))
        return rc;
    return TOSCA_OK;
}

int check_wall_tau(tosca_ctx *c, const tosca_params *p) {
    const Geo &g = c->geo;
    int rc;
    if (p->bc_u[0] == TOSCA_BCU_WALLFUNCTION || p->bc_u[1] == TOSCA_BCU_WALLFUNCTION)
        if ((rc = check_dev(c, p->tau_wall_i, 3 * (size_t)g.mz * g.my * sizeof(double), "tau_wall_i"))) return rc;
    if (p->bc_u[2] == TOSCA_BCU_WALLFUNCTION || p->bc_u[3] == TOSCA_BCU_WALLFUNCTION)
        if ((rc = check_dev(c, p->tau_wall_j, 3 * (size_t)g.mz * g.mx * sizeof(double), "tau_wall_j"))) return rc;
    return TOSCA_OK;
}

// ExplicitRhsU's arguments (ueqn_solvers.c:1276-1407)
int check_explicit(tosca_ctx *c, const tosca_mesh *m, const tosca_params *p, double *ucont, double *ucat,
                   const double *nut, const double *dp, const double *btheta, const double *btheta_o, int ab2,
                   const double *farm, const double *ucont_o, double *rhs, int form_mode) {
    int rc = check_ctx(c);
    if (rc) return rc;
    if ((rc = check_params(c, p))) return rc;
    if (form_mode < 0 || form_mode > 2) return set_err(c, TOSCA_ERR_ARG, "form_mode = %d (0, 1 or 2)", form_mode);
    if (p->div_scheme != TOSCA_DIV_CENTRAL && p->div_scheme != TOSCA_DIV_QUICK)
        return set_err(c, TOSCA_ERR_UNSUPPORTED, "div_scheme = %d: only central (0) and QUICK (1) are supported", p->div_scheme);
    set_geo_per(c, p);
    const Geo &g = c->geo;
    const size_t n = (size_t)g.nloc(), vloc = 3 * n * sizeof(double), vglob = 3 * (size_t)g.mx * g.my * g.mz * sizeof(double);
    if ((rc = check_mesh(c, m, true, true, true))) return rc;
    if ((rc = check_dev(c, ucont, vloc, "ucont")) || (rc = check_dev(c, ucat, vloc, "ucat")) ||
        (rc = check_dev(c, dp, vglob, "dp")) || (rc = check_dev(c, ucont_o, vglob, "ucont_o")) ||
        (rc = check_dev(c, rhs, vglob, "rhs")))
        return rc;
    if (btheta) {
        if ((rc = check_dev(c, btheta, vglob, "btheta"))) return rc;
        if (ab2 && (rc = check_dev(c, btheta_o, vglob, "btheta_o"))) return rc;
    }
    if (farm && (rc = check_dev(c, farm, vglob, "farm"))) return rc;
    if (rhs == ucont_o || rhs == dp) return set_err(c, TOSCA_ERR_ARG, "rhs must not alias ucont_o or dp");
    const bool visc = form_mode != 1;
    if (p->les_active && visc) {
        if ((rc = check_les_model(c, p))) return rc;
        if ((rc = check_dev(c, nut, n * sizeof(double), "nut"))) return rc;
    }
    if (visc && (rc = check_wall_tau(c, p))) return rc;
    return ensure_ws(c, ws_need(c, p));
}

// ExplicitRhsU's kernels in the reference's order, on validated arguments
void queue_explicit(tosca_ctx *c, const tosca_mesh *m, const tosca_params *p, double *ucont, double *ucat,
                    const double *nut, const double *dp, const double *btheta, const double *btheta_o, int ab2,
                    const double *farm, const double *ucont_o, double *rhs, int form_mode) {
    const Geo &g = c->geo;
    hipStream_t s = c->stream;
    const bool visc = form_mode != 1;
    tosca::launch_no_penetration(ucont, p, g, s);                    // ueqn_solvers.c:1281
    tosca::launch_reset_face_periodic(ucont, 3, g, s);               // :1284
    tosca::launch_contra_to_cart(ucont, m, ucat, g, s);              // :1287
    tosca::launch_reset_cell_periodic(ucat, 3, g, s);                // :1301
    tosca::launch_rhs_init(rhs, dp, btheta, ab2 ? btheta_o : nullptr, btheta ? ab2 : 0, farm, g, s);   // :1304-1344
    const double *nu_t = (p->les_active && visc) ? nut : nullptr;   // :1347
    if (c->ref_path) {
        tosca::ref_formu(g, m, p, ucont, ucat, nu_t, rhs, 1.0, form_mode, c->ws, s);
    } else {
        tosca::FormUArgs A = formu_args(c, m, p, ucont, ucat, nu_t, rhs, 1.0, form_mode);
        tosca::fused_formu(g, A, c->ws + 3 * (size_t)g.nloc(), s);
    }
    tosca::launch_non_resolved(rhs, ucont_o, p, g, s);               // :1403
}

// IMEXMatVec's kernels (ueqn_solvers.c:1739-1801) on validated arguments: av = v - coef Visc(v)
void queue_imex(tosca_ctx *c, const tosca_mesh *m, const tosca_params *p, const double *v, double *ucont,
                double *ucat, const double *nut, const double *ucont_o, double coef, double *av) {
    const Geo &g = c->geo;
    hipStream_t s = c->stream;
    const long ng3 = 3L * g.mx * g.my * g.mz;
    tosca::launch_global_to_local(ucont, v, g, s);                         // :1763-1767
    tosca::launch_wrap(ucont, 3, g, s);
    tosca::launch_no_penetration(ucont, p, g, s);                          // :1770
    tosca::launch_reset_face_periodic(ucont, 3, g, s);                     // :1773
    tosca::launch_contra_to_cart(ucont, m, ucat, g, s);                    // :1777
    tosca::launch_reset_cell_periodic(ucat, 3, g, s);                      // :1790
    tosca::launch_fill(av, ng3, 0.0, s);                                   // :1793
    const double *nu_t = p->les_active ? nut : nullptr;
    if (c->ref_path) {                                                     // :1794
        tosca::ref_formu(g, m, p, ucont, ucat, nu_t, av, -1.0 * coef, 2, c->ws, s);
    } else {
        tosca::FormUArgs A = formu_args(c, m, p, ucont, ucat, nu_t, av, -1.0 * coef, 2);
        tosca::fused_formu(g, A, c->ws + 3 * (size_t)g.nloc(), s);
    }
    tosca::launch_non_resolved(av, ucont_o, p, g, s);                      // :1795
    tosca::launch_axpy(av, 1.0, v, ng3, s);                                // :1796
}

int check_imex(tosca_ctx *c, const tosca_mesh *m, const tosca_params *p, const double *ucont, const double *ucat,
               const double *nut, const double *ucont_o) {
    int rc;
    if ((rc = check_params(c, p))) return rc;
    if (p->div_scheme != TOSCA_DIV_CENTRAL && p->div_scheme != TOSCA_DIV_QUICK)
        return set_err(c, TOSCA_ERR_UNSUPPORTED, "div_scheme = %d: only central (0) and QUICK (1) are supported", p->div_scheme);
    set_geo_per(c, p);
    const Geo &g = c->geo;
    const size_t n = (size_t)g.nloc(), vloc = 3 * n * sizeof(double), vglob = 3 * (size_t)g.mx * g.my * g.mz * sizeof(double);
    if ((rc = check_mesh(c, m, true, true, true))) return rc;
    if ((rc = check_dev(c, ucont, vloc, "ucont")) || (rc = check_dev(c, ucat, vloc, "ucat")) ||
        (rc = check_dev(c, ucont_o, vglob, "ucont_o")))
        return rc;
    if (p->les_active) {
        if ((rc = check_les_model(c, p))) return rc;
        if ((rc = check_dev(c, nut, n * sizeof(double), "nut"))) return rc;
    }
    if ((rc = check_wall_tau(c, p))) return rc;
    return ensure_ws(c, ws_need(c, p));
}

int check_krylov(tosca_ctx *c, const tosca_krylov_params *kp, tosca_krylov_result *kr) {
    if (!kp || !kr) return set_err(c, TOSCA_ERR_ARG, "krylov params / result is NULL");
    if (kp->restart < 1 || kp->restart > 200) return set_err(c, TOSCA_ERR_ARG, "restart = %d (1 .. 200)", kp->restart);
    if (kp->max_it < 0) return set_err(c, TOSCA_ERR_ARG, "max_it = %d", kp->max_it);
    if (!(kp->atol >= 0.0) || !(kp->rtol >= 0.0)) return set_err(c, TOSCA_ERR_ARG, "tolerances must be >= 0");
    return TOSCA_OK;
}

// Krylov workspace: basis (restart + 1) n, tmp n, the predictors' extra vectors, mdot partials,
// device scalars
struct KrylovWs { double *V, *tmp, *extra, *partial, *dsc; };

int ensure_kws(tosca_ctx *c, int restart, int nextra, KrylovWs &K) {
    const size_t n = 3 * (size_t)c->geo.mx * c->geo.my * c->geo.mz;
    const size_t np = tosca::mdot_partial_doubles(restart + 1), nsc = 2 * (size_t)(restart + 2);
    const size_t need = (size_t)(restart + 2 + nextra) * n + np + nsc;
    if (c->kws_elems < need) {
        if (c->kws) { hipStreamSynchronize(c->stream); hipFree(c->kws); c->kws = nullptr; c->kws_elems = 0; }
        if (hipMalloc(&c->kws, need * sizeof(double)) != hipSuccess) {
            (void)hipGetLastError();
            return set_err(c, TOSCA_ERR_DEVICE, "Krylov workspace allocation of %zu bytes failed", need * sizeof(double));
        }
        c->kws_elems = need;
    }
    K.V = c->kws;
    K.tmp = K.V + (size_t)(restart + 1) * n;
    K.extra = K.tmp + n;
    K.partial = K.extra + (size_t)nextra * n;
    K.dsc = K.partial + np;
    return TOSCA_OK;
}

// device values -> host, after everything queued so far
int fetch(tosca_ctx *c, const double *d, double *h, int cnt) {
    if (hipMemcpyAsync(h, d, cnt * sizeof(double), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess)
        return set_err(c, TOSCA_ERR_DEVICE, "GMRES: device-to-host copy failed: %s", hipGetErrorString(hipGetLastError()));
    return TOSCA_OK;
}

// Restarted GMRES as PETSc 3.15 runs it for kspIMEX (src/ksp/ksp/impls/gmres/gmres.c:
// KSPSolve_GMRES, KSPGMRESCycle with classical Gram-Schmidt, KSPGMRESUpdateHessenberg,
// KSPGMRESBuildSoln; KSPConvergedDefault).  kspIMEX has a nonzero initial guess and no
// preconditioner (src/ueqn.c:264-276), so the relative and divergence tests are taken against ||b||
// (||r0|| when b = 0), dtol 1e30 (ueqn.c:277); reasons are PETSc's KSPConvergedReason values.
// Vector work on the device, the (restart + 1) x restart Hessenberg system on the host.
template <class Op>
int gmres_device(tosca_ctx *c, long n, Op A, const double *b, double *x, const tosca_krylov_params *kp,
                 tosca_krylov_result *kr, const KrylovWs &K) {
    hipStream_t s = c->stream;
    const int m = kp->restart;
    std::vector<double> HH((size_t)(m + 2) * (m + 1), 0.0), cc(m + 1), ss(m + 1), grs(m + 2), nrs(m + 1), h(m + 1);
    auto H = [&](int a, int col) -> double & { return HH[(size_t)col * (m + 2) + a]; };
    auto V = [&](int j) { return K.V + (size_t)j * n; };
    int its = 0, reason = 0, rc;
    double ttol = 0.0, res = 0.0, sq = 0.0, rnorm0 = 0.0, bsq = 0.0;
    const double divtol = 1e30;
    tosca::launch_mdot(b, n, 1, b, n, K.partial, K.dsc, s);                // ||b||
    if ((rc = fetch(c, K.dsc, &bsq, 1))) return rc;
    const double bnorm = std::sqrt(bsq);
    bool first = true;
    // KSPConvergedDefault: NaN / Inf (-9), the tolerance (3 below atol, else 2), dtol (-4)
    auto test = [&](double r) {
        if (std::isnan(r) || std::isinf(r)) return -9;
        if (r <= ttol) return r < kp->atol ? 3 : 2;
        if (r >= divtol * rnorm0) return -4;
        return 0;
    };
    while (!reason) {
        A(x, K.tmp);                                                        // KSPInitialResidual
        tosca::launch_residual(V(0), b, K.tmp, n, s);
        tosca::launch_mdot(V(0), n, 1, V(0), n, K.partial, K.dsc, s);
        if ((rc = fetch(c, K.dsc, &sq, 1))) return rc;
        res = std::sqrt(sq);
        if (res > 0.0) tosca::launch_scale_to(V(0), V(0), 1.0 / res, n, s);
        grs[0] = res;
        if (first) {
            rnorm0 = bnorm > 0.0 ? bnorm : res;
            ttol = std::max(kp->rtol * rnorm0, kp->atol);
            first = false;
        }
        reason = test(res);
        if (reason < 0) break;
        int it = 0;
        bool hapend = false;
        while (!reason && it < m && its < kp->max_it) {
            double *w = V(it + 1);
            A(V(it), w);
            tosca::launch_mdot(K.V, n, it + 1, w, n, K.partial, K.dsc, s);  // h_j = v_j . w
            if ((rc = fetch(c, K.dsc, h.data(), it + 1))) return rc;
            tosca::launch_maxpy(w, K.V, n, it + 1, K.dsc, -1.0, n, s);    // w -= sum h_j v_j
            for (int j = 0; j <= it; j++) H(j, it) = h[j];
            tosca::launch_mdot(w, n, 1, w, n, K.partial, K.dsc, s);
            if ((rc = fetch(c, K.dsc, &sq, 1))) return rc;
            const double tt = std::sqrt(sq);
            H(it + 1, it) = tt;
            double hapbnd = std::fabs(tt / grs[it]);
            if (hapbnd > 1e-30) hapbnd = 1e-30;
            hapend = tt < hapbnd;
            if (!hapend) tosca::launch_scale_to(w, w, 1.0 / tt, n, s);
            for (int j = 1; j <= it; j++) {                                 // previous rotations
                const double t0 = H(j - 1, it);
                H(j - 1, it) = cc[j - 1] * t0 + ss[j - 1] * H(j, it);
                H(j, it) = cc[j - 1] * H(j, it) - ss[j - 1] * t0;
            }
            if (!hapend) {
                const double t1 = std::sqrt(H(it, it) * H(it, it) + H(it + 1, it) * H(it + 1, it));
                if (t1 == 0.0) { reason = -2; break; }            // KSP_DIVERGED_NULL
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
        const int last = it - 1;                                            // KSPGMRESBuildSoln
        if (last >= 0) {
            nrs[last] = (H(last, last) != 0.0) ? grs[last] / H(last, last) : 0.0;
            for (int ii = 1; ii <= last; ii++) {
                const int k = last - ii;
                double tt = grs[k];
                for (int j = k + 1; j <= last; j++) tt = tt - H(k, j) * nrs[j];
                nrs[k] = tt / H(k, k);
            }
            if (hipMemcpyAsync(K.dsc, nrs.data(), (last + 1) * sizeof(double), hipMemcpyHostToDevice, s) != hipSuccess)
                return set_err(c, TOSCA_ERR_DEVICE, "GMRES: host-to-device copy failed");
            tosca::launch_fill(K.tmp, n, 0.0, s);
            tosca::launch_maxpy(K.tmp, K.V, n, last + 1, K.dsc, 1.0, n, s);
            tosca::launch_axpy(x, 1.0, K.tmp, n, s);
            if (hipStreamSynchronize(s) != hipSuccess) return set_err(c, TOSCA_ERR_DEVICE, "GMRES: stream error");
        }
        if (its >= kp->max_it && !reason) reason = -3;
    }
    kr->iterations = its;
    kr->residual = res;
    kr->reason = reason;
    return TOSCA_OK;
}

}  // namespace

extern "C" {

const char *tosca_version(void) { return "tosca-mi355x 0.1 (gfx950)"; }

const char *tosca_last_error(const tosca_ctx *ctx) { return ctx ? ctx->err.c_str() : g_err.c_str(); }

int tosca_ctx_create(tosca_ctx **out, const tosca_dmda_info *in, int device) {
    if (!out) return set_err(nullptr, TOSCA_ERR_ARG, "ctx output pointer is NULL");
    *out = nullptr;
    if (!in) return set_err(nullptr, TOSCA_ERR_ARG, "info is NULL");
    if (in->xs != 0 || in->ys != 0 || in->zs != 0 || in->xm != in->mx || in->ym != in->my || in->zm != in->mz)
        return set_err(nullptr, TOSCA_ERR_LAYOUT, "one GPU must own the whole DMDA (xs=ys=zs=0, xm=mx, ym=my, zm=mz)");
    if (in->gxs != -3 || in->gys != -3 || in->gzs != -3 || in->gxm != in->mx + 6 || in->gym != in->my + 6 ||
        in->gzm != in->mz + 6)
        return set_err(nullptr, TOSCA_ERR_LAYOUT, "local arrays must carry ghost width 3 on every side");
    if (in->mx < 4 || in->my < 4 || in->mz < 4)
        return set_err(nullptr, TOSCA_ERR_LAYOUT, "each axis needs at least 2 interior cells (m >= 4)");
    const double pts = (double)in->gxm * (double)in->gym * (double)in->gzm;
    if (pts * 10.0 >= 2147483647.0)
        return set_err(nullptr, TOSCA_ERR_LAYOUT, "local array too large for 32-bit element indexing (%.0f points)", pts);
    if (device_forced_off()) return set_err(nullptr, TOSCA_ERR_DEVICE, "device disabled by TOSCA_FORCE_NO_DEVICE");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        (void)hipGetLastError();
        return set_err(nullptr, TOSCA_ERR_DEVICE, "no HIP device available");
    }
    if (device < 0 || device >= ndev) return set_err(nullptr, TOSCA_ERR_ARG, "device %d out of range (%d devices)", device, ndev);
    if (hipSetDevice(device) != hipSuccess) return set_err(nullptr, TOSCA_ERR_DEVICE, "hipSetDevice(%d) failed", device);
    tosca_ctx *c = new tosca_ctx();
    c->device = device;
    const char *path = getenv("TOSCA_PATH");
    c->ref_path = path && strcmp(path, "ref") == 0;
    c->info = *in;
    c->geo.mx = (int)in->mx; c->geo.my = (int)in->my; c->geo.mz = (int)in->mz;
    c->geo.gxm = (int)in->gxm; c->geo.gym = (int)in->gym; c->geo.gzm = (int)in->gzm;
    // the allocation-free FormU's scratch on this device; a failed lookup disables only that entry point
    const hipError_t se = scratch_setup(device, &c->scratch);
    if (se != hipSuccess) {
        c->scratch = nullptr;
        c->scratch_err = std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(se);
    }
    *out = c;
    return TOSCA_OK;
}

int tosca_ctx_destroy(tosca_ctx *c) {
    if (!c) return TOSCA_OK;
    hipSetDevice(c->device);
    c->timer.release();
    if (tosca::g_timer == &c->timer) tosca::g_timer = nullptr;
    if (c->ws || c->chk || c->kws || c->hs || c->adj) hipStreamSynchronize(c->stream);
    if (c->kws) hipFree(c->kws);
    if (c->adj) hipFree(c->adj);
    if (c->hs) hipFree(c->hs);
    if (c->cabot) hipFree(c->cabot);
    if (c->ws) hipFree(c->ws);
    if (c->chk) hipFree(c->chk);
    delete c;
    return TOSCA_OK;
}

int tosca_ctx_set_stream(tosca_ctx *c, void *stream) {
    if (!c) return set_err(nullptr, TOSCA_ERR_ARG, "context is NULL");
    c->stream = reinterpret_cast<hipStream_t>(stream);
    return TOSCA_OK;
}

int tosca_ctx_synchronize(tosca_ctx *c) {
    int rc = check_ctx(c);
    if (rc) return rc;
    hipError_t e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return set_err(c, TOSCA_ERR_DEVICE, "stream synchronize: %s", hipGetErrorString(e));
    return TOSCA_OK;
}

int tosca_ctx_enable_timing(tosca_ctx *c, int on) {
    int rc = check_ctx(c);
    if (rc) return rc;
    c->timer.on = on != 0;
    return TOSCA_OK;
}

int tosca_ctx_timing_reset(tosca_ctx *c) {
    int rc = check_ctx(c);
    if (rc) return rc;
    c->timer.reset();
    return TOSCA_OK;
}

int tosca_ctx_timing_count(tosca_ctx *c) {
    if (check_ctx(c)) return -1;
    c->timer.collect();
    return (int)c->timer.acc.size();
}

int tosca_ctx_timing_entry(tosca_ctx *c, int idx, char *name, int name_len, double *total_ms, int64_t *launches) {
    int rc = check_ctx(c);
    if (rc) return rc;
    c->timer.collect();
    if (idx < 0 || idx >= (int)c->timer.acc.size()) return set_err(c, TOSCA_ERR_ARG, "timing index %d out of range", idx);
    auto it = c->timer.acc.begin();
    std::advance(it, idx);
    if (name && name_len > 0) { strncpy(name, it->first.c_str(), name_len - 1); name[name_len - 1] = 0; }
    if (total_ms) *total_ms = it->second.first;
    if (launches) *launches = it->second.second;
    return TOSCA_OK;
}

int tosca_mesh_check(tosca_ctx *c, const tosca_mesh *m) {
    int rc = check_ctx(c);
    if (rc) return rc;
    if ((rc = check_mesh(c, m, false, true, false, false))) return rc;
    const size_t s = (size_t)c->geo.nloc() * sizeof(double);
    if ((rc = check_dev(c, m->nvert, s, "mesh.nvert")) || (rc = check_dev(c, m->meshtag, s, "mesh.meshtag"))) return rc;
    tosca_mesh mc = *m;             // the face bases of the cell-basis identity, when all three are given
    if (m->icsi && m->jeta && m->kzet) {
        if ((rc = check_dev(c, m->icsi, 3 * s, "mesh.icsi")) || (rc = check_dev(c, m->jeta, 3 * s, "mesh.jeta")) ||
            (rc = check_dev(c, m->kzet, 3 * s, "mesh.kzet")))
            return rc;
    } else {
        mc.icsi = mc.jeta = mc.kzet = nullptr;
    }
    const auto key = mesh_key(m);
    const auto dkey = derive_key(m);
    c->checked.erase(std::remove(c->checked.begin(), c->checked.end(), key), c->checked.end());
    c->derivable.erase(std::remove(c->derivable.begin(), c->derivable.end(), dkey), c->derivable.end());
    const size_t need = 5 * (size_t)(c->geo.mz) + 8;
    if (c->chk_elems < need) {
        if (c->chk) hipFree(c->chk);
        c->chk = nullptr;
        if (hipMalloc(&c->chk, need * sizeof(int)) != hipSuccess) return set_err(c, TOSCA_ERR_DEVICE, "allocation failed");
        c->chk_elems = need;
    }
    tosca::launch_mesh_check(c->geo, &mc, c->chk, c->stream);
    if ((rc = launched(c, "mesh check"))) return rc;
    std::vector<int> h(5 * c->geo.mz);
    hipMemcpyAsync(h.data(), c->chk, h.size() * sizeof(int), hipMemcpyDeviceToHost, c->stream);
    if (hipStreamSynchronize(c->stream) != hipSuccess) return set_err(c, TOSCA_ERR_DEVICE, "mesh check failed");
    long cnt[5] = {0, 0, 0, 0, 0};
    for (int k = 0; k < c->geo.mz; k++)
        for (int q = 0; q < 5; q++) cnt[q] += h[5 * k + q];
    if (cnt[0]) return set_err(c, TOSCA_ERR_MESH, "%ld points with nvert != 0 (immersed boundaries are not supported)", cnt[0]);
    if (cnt[1]) return set_err(c, TOSCA_ERR_MESH, "%ld points with meshTag != 0 (overset meshes are not supported)", cnt[1]);
    if (cnt[2]) return set_err(c, TOSCA_ERR_MESH, "%ld interior cells with negative or close-to-zero volume (aj < 1e-20)", cnt[2]);
    if (cnt[3]) return set_err(c, TOSCA_ERR_MESH, "%ld interior cells with a close-to-zero face area (|csi|,|eta|,|zet| < 1e-20)", cnt[3]);
    if (c->checked.size() >= 64) c->checked.erase(c->checked.begin());
    c->checked.push_back(key);
    if (m->icsi && m->jeta && m->kzet && cnt[4] == 0) {
        if (c->derivable.size() >= 64) c->derivable.erase(c->derivable.begin());
        c->derivable.push_back(dkey);
    }
    return TOSCA_OK;
}

int tosca_formu(tosca_ctx *c, const tosca_mesh *m, const tosca_params *p, const double *ucont, const double *ucat,
                const double *nut, double *rhs, double scale, int form_mode) {
    int rc = check_ctx(c);
    if (rc) return rc;
    if ((rc = check_params(c, p))) return rc;
    if (form_mode < 0 || form_mode > 2) return set_err(c, TOSCA_ERR_ARG, "form_mode = %d (0, 1 or 2)", form_mode);
    if (p->div_scheme != TOSCA_DIV_CENTRAL && p->div_scheme != TOSCA_DIV_QUICK)
        return set_err(c, TOSCA_ERR_UNSUPPORTED, "div_scheme = %d: only central (0) and QUICK (1) are supported", p->div_scheme);
    set_geo_per(c, p);
    const Geo &g = c->geo;
    const size_t n = (size_t)g.nloc();
    if ((rc = check_mesh(c, m, true, true, true))) return rc;
    if ((rc = check_dev(c, ucont, 3 * n * sizeof(double), "ucont")) || (rc = check_dev(c, ucat, 3 * n * sizeof(double), "ucat")))
        return rc;
    const bool visc = form_mode != 1;
    if (p->les_active && visc) {
        if ((rc = check_les_model(c, p))) return rc;
        if ((rc = check_dev(c, nut, n * sizeof(double), "nut"))) return rc;
        if (p->bc_u[0] == TOSCA_BCU_WALLFUNCTION || p->bc_u[1] == TOSCA_BCU_WALLFUNCTION)
            if ((rc = check_dev(c, p->tau_wall_i, 3 * (size_t)g.mz * g.my * sizeof(double), "tau_wall_i"))) return rc;
        if (p->bc_u[2] == TOSCA_BCU_WALLFUNCTION || p->bc_u[3] == TOSCA_BCU_WALLFUNCTION)
            if ((rc = check_dev(c, p->tau_wall_j, 3 * (size_t)g.mz * g.mx * sizeof(double), "tau_wall_j"))) return rc;
    }
    if ((rc = check_dev(c, rhs, 3 * (size_t)g.mx * g.my * g.mz * sizeof(double), "rhs"))) return rc;
    if (c->ref_path) {
        if ((rc = ensure_ws(c, ws_need(c, p)))) return rc;
        tosca::ref_formu(g, m, p, ucont, ucat, (p->les_active && visc) ? nut : nullptr, rhs, scale, form_mode, c->ws, c->stream);
    } else {
        if ((rc = ensure_ws(c, ws_need(c, p)))) return rc;
        tosca::FormUArgs A = formu_args(c, m, p, ucont, ucat, (p->les_active && visc) ? nut : nullptr, rhs, scale, form_mode);
        tosca::fused_formu(g, A, c->ws + 3 * (size_t)g.nloc(), c->stream);
    }
    return launched(c, "FormU");
}

int tosca_formu_noalloc(tosca_ctx *c, const tosca_mesh *m, const tosca_params *p, const double *ucont,
                        const double *ucat, const double *nut, double *rhs, double scale, int form_mode) {
    int rc = check_ctx(c);
    if (rc) return rc;
    tosca::g_timer = nullptr;      // the context's kernel timer creates events: not used here
    if (!c->scratch) return set_err(c, TOSCA_ERR_DEVICE, "tosca_formu_noalloc: no scratch on device %d (%s)", c->device,
                                    c->scratch_err.c_str());
    if ((rc = check_params(c, p))) return rc;
    if (form_mode < 0 || form_mode > 2) return set_err(c, TOSCA_ERR_ARG, "form_mode = %d (0, 1 or 2)", form_mode);
    if (p->div_scheme != TOSCA_DIV_CENTRAL && p->div_scheme != TOSCA_DIV_QUICK)
        return set_err(c, TOSCA_ERR_UNSUPPORTED, "div_scheme = %d: only central (0) and QUICK (1) are supported", p->div_scheme);
    set_geo_per(c, p);
    const Geo &g = c->geo;
    const size_t need = tosca::formu_ws_doubles(g);
    if (need > TOSCA_NOALLOC_SCRATCH_DOUBLES)
        return set_err(c, TOSCA_ERR_LAYOUT, "tosca_formu_noalloc: this layout needs %zu doubles of scratch, more than the "
                                            "%zu the library holds per device (use tosca_formu)", need,
                       (size_t)TOSCA_NOALLOC_SCRATCH_DOUBLES);
    const size_t n = (size_t)g.nloc();
    if ((rc = check_mesh(c, m, true, true, true))) return rc;
    if ((rc = check_dev(c, ucont, 3 * n * sizeof(double), "ucont")) || (rc = check_dev(c, ucat, 3 * n * sizeof(double), "ucat")))
        return rc;
    const bool visc = form_mode != 1;
    if (p->les_active && visc) {
        if ((rc = check_les_model(c, p))) return rc;
        if ((rc = check_dev(c, nut, n * sizeof(double), "nut"))) return rc;
        if ((rc = check_wall_tau(c, p))) return rc;
    }
    if ((rc = check_dev(c, rhs, 3 * (size_t)g.mx * g.my * g.mz * sizeof(double), "rhs"))) return rc;
    // the fused kernels write every scratch element they read before reading it: nothing to clear
    tosca::FormUArgs A = formu_args(c, m, p, ucont, ucat, (p->les_active && visc) ? nut : nullptr, rhs, scale, form_mode);
    tosca::fused_formu(g, A, c->scratch, c->stream);
    return launched(c, "FormU (allocation-free)");
}

int tosca_explicit_rhs_u(tosca_ctx *c, const tosca_mesh *m, const tosca_params *p, double *ucont, double *ucat,
                         const double *nut, const double *dp, const double *btheta, const double *btheta_o, int ab2,
                         const double *farm, const double *ucont_o, double *rhs, int form_mode) {
    int rc = check_explicit(c, m, p, ucont, ucat, nut, dp, btheta, btheta_o, ab2, farm, ucont_o, rhs, form_mode);
    if (rc) return rc;
    queue_explicit(c, m, p, ucont, ucat, nut, dp, btheta, btheta_o, ab2, farm, ucont_o, rhs, form_mode);
    return launched(c, "ExplicitRhsU");
}

int tosca_ueqn_fe(tosca_ctx *c, const tosca_mesh *m, const tosca_params *p, double *ucont, double *ucat,
                  const double *nut, const double *dp, const double *btheta, const double *btheta_o, int ab2,
                  const double *farm, const double *ucont_o, double dt, double *rhs, int form_mode) {
    int rc = check_explicit(c, m, p, ucont, ucat, nut, dp, btheta, btheta_o, ab2, farm, ucont_o, rhs, form_mode);
    if (rc) return rc;
    queue_explicit(c, m, p, ucont, ucat, nut, dp, btheta, btheta_o, ab2, farm, ucont_o, rhs, form_mode);  // :1424
    tosca::launch_fe_update(rhs, ucont_o, dt, ucont, c->geo, c->stream);    // :1427-1433
    tosca::launch_wrap(ucont, 3, c->geo, c->stream);                        // :1436-1437
    return launched(c, "UeqnFE");
}

int tosca_ueqn_rk4(tosca_ctx *c, const tosca_mesh *m, const tosca_params *p, double *ucont, double *ucat,
                   const double *nut, const double *dp, const double *btheta, const double *btheta_o, int ab2,
                   const double *farm, const double *ucont_o, double dt, double *rhs, double *utmp, int form_mode) {
    int rc = check_explicit(c, m, p, ucont, ucat, nut, dp, btheta, btheta_o, ab2, farm, ucont_o, rhs, form_mode);
    if (rc) return rc;
    const Geo &g = c->geo;
    const size_t vglob = 3 * (size_t)g.mx * g.my * g.mz * sizeof(double);
    if ((rc = check_dev(c, utmp, vglob, "utmp"))) return rc;
    if (utmp == rhs || utmp == ucont_o) return set_err(c, TOSCA_ERR_ARG, "utmp must not alias rhs or ucont_o");
    const double b[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};   // :1458-1466
    const double a[4] = {0.0, 0.5, 0.5, 1.0};
    hipStream_t s = c->stream;
    if (hipMemcpyAsync(utmp, ucont_o, vglob, hipMemcpyDeviceToDevice, s) != hipSuccess)   // :1471
        return set_err(c, TOSCA_ERR_DEVICE, "UeqnRK4: copy of ucont_o failed");
    for (int i = 0; i < 4; i++) {
        queue_explicit(c, m, p, ucont, ucat, nut, dp, btheta, btheta_o, ab2, farm, ucont_o, rhs, form_mode);   // :1489
        // utmp += dt b_i Rhs (:1492) and, in the same pass, the next stage's ucont = Ucont_o +
        // a_{i+1} dt Rhs (:1480-1486) or, after the last stage, ucont = utmp (:1495-1497)
        tosca::launch_rk_stage(utmp, dt * b[i], rhs, ucont, i < 3 ? a[i + 1] * dt : 0.0, ucont_o, i == 3, g, s);
        tosca::launch_wrap(ucont, 3, g, s);
    }
    return launched(c, "UeqnRK4");
}

}  // extern "C"

namespace {

constexpr double kWrayBeta[3] = {8.0 / 15.0, 5.0 / 12.0, 3.0 / 4.0};   // ueqn_solvers.c:1518
constexpr double kWrayDelta[3] = {0.0, 17.0 / 60.0, 5.0 / 12.0};       // :1520

// the explicit Wray Runge-Kutta stages on validated arguments; chk (nullable): the global state at
// the start of each stage, before ExplicitRhsU's resets, is stored at chk + s * ng3, and the state
// after the last stage at chk + n_stages * ng3
void queue_rk3w(tosca_ctx *c, const tosca_mesh *m, const tosca_params *p, double *ucont, double *ucat,
                const double *nut, const double *dp, const double *btheta, const double *btheta_o, int ab2,
                const double *farm, double dt, int n_stages, double *ucont_o, double *rhs, double *rhs_o,
                int form_mode, double *chk) {
    const Geo &g = c->geo;
    hipStream_t s = c->stream;
    const long ng3 = 3L * g.mx * g.my * g.mz;
    for (int st = 0; st < n_stages; st++) {
        const int i = st % 3;
        if (i == 0) {                                       // a time step begins
            tosca::launch_local_to_global(ucont_o, ucont, g, s);       // Ucont_o = Ucont (time loop)
            tosca::launch_fill(rhs_o, ng3, 0.0, s);                    // VecSet(Rhs_o, 0), :1522
        }
        if (chk) tosca::launch_local_to_global(chk + st * ng3, ucont, g, s);
        queue_explicit(c, m, p, ucont, ucat, nut, dp, btheta, btheta_o, ab2, farm, ucont_o, rhs, form_mode);   // :1527
        // Ucont = (Ucont + dt beta_i Rhs) + (-1.0 dt delta_i) Rhs_o; Rhs_o = Rhs (:1533-1541, 1563-1569)
        tosca::launch_rk3w_update(ucont, rhs, rhs_o, dt * kWrayBeta[i], -1.0 * dt * kWrayDelta[i], g, s);
        tosca::launch_wrap(ucont, 3, g, s);
    }
    if (chk) tosca::launch_local_to_global(chk + (long)n_stages * ng3, ucont, g, s);
}

int check_rk3w(tosca_ctx *c, const tosca_mesh *m, const tosca_params *p, double *ucont, double *ucat,
               const double *nut, const double *dp, const double *btheta, const double *btheta_o, int ab2,
               const double *farm, int n_stages, double *ucont_o, double *rhs, double *rhs_o, int form_mode) {
    int rc = check_explicit(c, m, p, ucont, ucat, nut, dp, btheta, btheta_o, ab2, farm, ucont_o, rhs, form_mode);
    if (rc) return rc;
    if (n_stages < 0) return set_err(c, TOSCA_ERR_ARG, "n_stages = %d", n_stages);
    const Geo &g = c->geo;
    const size_t vglob = 3 * (size_t)g.mx * g.my * g.mz * sizeof(double);
    if ((rc = check_dev(c, rhs_o, vglob, "rhs_o"))) return rc;
    const void *v[6] = {ucont_o, rhs, rhs_o, dp, btheta, farm};
    for (int a = 0; a < 3; a++)
        for (int b = a + 1; b < 6; b++)
            if (v[a] == v[b]) return set_err(c, TOSCA_ERR_ARG, "ucont_o, rhs and rhs_o must be distinct from each other "
                                                               "and from the sources");
    return TOSCA_OK;
}

}  // namespace

extern "C" {

int tosca_ueqn_rk3w(tosca_ctx *c, const tosca_mesh *m, const tosca_params *p, double *ucont, double *ucat,
                    const double *nut, const double *dp, const double *btheta, const double *btheta_o, int ab2,
                    const double *farm, double dt, int n_stages, double *ucont_o, double *rhs, double *rhs_o,
                    int form_mode) {
    int rc = check_rk3w(c, m, p, ucont, ucat, nut, dp, btheta, btheta_o, ab2, farm, n_stages, ucont_o, rhs, rhs_o,
                        form_mode);
    if (rc) return rc;
    queue_rk3w(c, m, p, ucont, ucat, nut, dp, btheta, btheta_o, ab2, farm, dt, n_stages, ucont_o, rhs, rhs_o, form_mode,
               nullptr);
    return launched(c, "UeqnRK3W");
}

int tosca_ueqn_rk3w_adjoint(tosca_ctx *c, const tosca_mesh *m, const tosca_params *p, double *ucont, double *ucat,
                            const double *nut, const double *dp, const double *btheta, const double *btheta_o,
                            int ab2, const double *farm, double dt, int n_stages, double *ucont_o, double *rhs,
                            double *rhs_o, double *ucont_bar, double *nut_bar) {
    int rc = check_rk3w(c, m, p, ucont, ucat, nut, dp, btheta, btheta_o, ab2, farm, n_stages, ucont_o, rhs, rhs_o, 0);
    if (rc) return rc;
    const Geo &g = c->geo;
    const size_t n = (size_t)g.nloc(), ng3 = 3 * (size_t)g.mx * g.my * g.mz;
    if ((rc = check_dev(c, ucont_bar, ng3 * sizeof(double), "ucont_bar"))) return rc;
    const bool les = p->les_active != 0;
    if (les && (rc = check_dev(c, nut_bar, n * sizeof(double), "nut_bar"))) return rc;
    if (ucont_bar == rhs || ucont_bar == rhs_o || ucont_bar == ucont_o)
        return set_err(c, TOSCA_ERR_ARG, "ucont_bar must not alias the work vectors");
    // the FormU adjoint's workspace, then the checkpoints (n_stages + 1 states) and the work vectors
    if ((rc = ensure_ws(c, std::max(ws_need(c, p), 52 * n + 64)))) return rc;
    const size_t need = ((size_t)n_stages + 1) * ng3 + 2 * ng3 + 9 * n;
    if (c->adj_elems < need) {
        if (c->adj) { hipStreamSynchronize(c->stream); hipFree(c->adj); c->adj = nullptr; c->adj_elems = 0; }
        if (hipMalloc(&c->adj, need * sizeof(double)) != hipSuccess) {
            (void)hipGetLastError();
            return set_err(c, TOSCA_ERR_DEVICE, "checkpoint store of %zu bytes failed", need * sizeof(double));
        }
        c->adj_elems = need;
    }
    hipStream_t s = c->stream;
    double *chk = c->adj, *carry = chk + ((size_t)n_stages + 1) * ng3, *rbar = carry + ng3, *vbar = rbar + ng3,
           *ubar = vbar + 3 * n, *qbar = ubar + 3 * n;
    // forward sweep, storing each stage's starting state
    queue_rk3w(c, m, p, ucont, ucat, nut, dp, btheta, btheta_o, ab2, farm, dt, n_stages, ucont_o, rhs, rhs_o, 0, chk);
    // backward sweep: each stage's steps transposed, in reverse order
    tosca::launch_fill(carry, (long)ng3, 0.0, s);
    if (les) tosca::launch_fill(nut_bar, (long)n, 0.0, s);
    for (int st = n_stages - 1; st >= 0; st--) {
        const int i = st % 3;
        const double *x = chk + (size_t)st * ng3, *uo = chk + (size_t)(st - i) * ng3;
        // the stage's reset fluxes and Cartesian velocity, recomputed from its checkpoint
        tosca::launch_global_to_local(ucont, x, g, s);
        tosca::launch_wrap(ucont, 3, g, s);
        tosca::launch_no_penetration(ucont, p, g, s);
        tosca::launch_reset_face_periodic(ucont, 3, g, s);
        tosca::launch_contra_to_cart(ucont, m, ucat, g, s);
        tosca::launch_reset_cell_periodic(ucat, 3, g, s);
        // cotangent of the stage's Rhs; resetNonResolvedCellFaces zeroes the same entries of it
        tosca::launch_rk3w_adj_combine(ucont_bar, carry, rbar, dt * kWrayBeta[i], -1.0 * dt * kWrayDelta[i], g, s);
        tosca::launch_non_resolved(rbar, uo, p, g, s);
        // FormU (the sources are constant): cotangents of the local ucont, ucat and nu_t
        tosca::vjp_formu_nut(g, m, p, ucont, ucat, nut, rbar, 1.0, vbar, ubar, les ? nut_bar : nullptr, c->ws, s);
        // contravariantToCartesian and the periodic resets of ucat, transposed into the ucont cotangent
        tosca::launch_reset_cell_periodic_adj(ubar, 3, g, s);
        tosca::launch_c2c_adj(ubar, m, qbar, vbar, g, s);
        // the stage update reads the reset state: its cotangent is the next state's plus ExplicitRhsU's
        tosca::launch_add_owned(ucont_bar, vbar, g, s);
        tosca::launch_face_resets_adj(ucont_bar, x, p, g, s);
    }
    // leave ucont and ucat as the forward sweep left them
    if (n_stages > 0) {
        tosca::launch_global_to_local(ucont, chk + (size_t)(n_stages - 1) * ng3, g, s);
        tosca::launch_wrap(ucont, 3, g, s);
        tosca::launch_no_penetration(ucont, p, g, s);
        tosca::launch_reset_face_periodic(ucont, 3, g, s);
        tosca::launch_contra_to_cart(ucont, m, ucat, g, s);
        tosca::launch_reset_cell_periodic(ucat, 3, g, s);
        tosca::launch_global_to_local(ucont, chk + (size_t)n_stages * ng3, g, s);
        tosca::launch_wrap(ucont, 3, g, s);
    }
    return launched(c, "UeqnRK3W adjoint");
}

int tosca_imex_matvec(tosca_ctx *c, const tosca_mesh *m, const tosca_params *p, const double *v, double *ucont,
                      double *ucat, const double *nut, const double *ucont_o, double coef, double *av) {
    int rc = check_ctx(c);
    if (rc) return rc;
    if ((rc = check_imex(c, m, p, ucont, ucat, nut, ucont_o))) return rc;
    const Geo &g = c->geo;
    const size_t vglob = 3 * (size_t)g.mx * g.my * g.mz * sizeof(double);
    if ((rc = check_dev(c, v, vglob, "v")) || (rc = check_dev(c, av, vglob, "av"))) return rc;
    if (av == v) return set_err(c, TOSCA_ERR_ARG, "av must not alias v");
    queue_imex(c, m, p, v, ucont, ucat, nut, ucont_o, coef, av);
    return launched(c, "IMEXMatVec");
}

int tosca_imex_solve(tosca_ctx *c, const tosca_mesh *m, const tosca_params *p, const double *b, double *x,
                     double *ucont, double *ucat, const double *nut, const double *ucont_o, double coef,
                     const tosca_krylov_params *kp, tosca_krylov_result *kr) {
    int rc = check_ctx(c);
    if (rc) return rc;
    if ((rc = check_krylov(c, kp, kr))) return rc;
    if ((rc = check_imex(c, m, p, ucont, ucat, nut, ucont_o))) return rc;
    const Geo &g = c->geo;
    const long n = 3L * g.mx * g.my * g.mz;
    if ((rc = check_dev(c, b, n * sizeof(double), "b")) || (rc = check_dev(c, x, n * sizeof(double), "x"))) return rc;
    if (b == x) return set_err(c, TOSCA_ERR_ARG, "x must not alias b");
    KrylovWs K;
    if ((rc = ensure_kws(c, kp->restart, 0, K))) return rc;
    auto A = [&](const double *v, double *out) { queue_imex(c, m, p, v, ucont, ucat, nut, ucont_o, coef, out); };
    if ((rc = gmres_device(c, n, A, b, x, kp, kr, K))) return rc;
    return launched(c, "KSPSolve(IMEX)");
}

int tosca_ueqn_rk3cn(tosca_ctx *c, const tosca_mesh *m, const tosca_params *p, int scheme, double *ucont,
                     double *ucat, const double *nut, const double *dp, const double *btheta, const double *btheta_o,
                     int ab2, const double *farm, const double *ucont_o, double dt, double *rhs,
                     const tosca_krylov_params *kp, tosca_krylov_result *kr) {
    int rc = check_explicit(c, m, p, ucont, ucat, nut, dp, btheta, btheta_o, ab2, farm, ucont_o, rhs, 1);
    if (rc) return rc;
    if (scheme != 0 && scheme != 1) return set_err(c, TOSCA_ERR_ARG, "scheme = %d (0 RK3WCN, 1 RK3SOCN)", scheme);
    if ((rc = check_krylov(c, kp, kr))) return rc;
    if ((rc = check_imex(c, m, p, ucont, ucat, nut, ucont_o))) return rc;
    const Geo &g = c->geo;
    hipStream_t s = c->stream;
    const long n = 3L * g.mx * g.my * g.mz;
    KrylovWs K;
    if ((rc = ensure_kws(c, kp->restart, 5, K))) return rc;
    double *rhs_o = K.extra, *visc = K.extra + n, *bU = K.extra + 2 * n, *utmp = K.extra + 3 * n, *ucg = K.extra + 4 * n;
    const double wb[3] = {8.0 / 15.0, 5.0 / 12.0, 3.0 / 4.0}, wg[3] = {4.0 / 15.0, 1.0 / 15.0, 1.0 / 6.0},
                 wd[3] = {0.0, 17.0 / 60.0, 5.0 / 12.0};                   // ueqn_solvers.c:1518-1520
    const double sa[3] = {1.0, 3.0 / 4.0, 1.0 / 3.0}, sb[3] = {1.0, 1.0 / 4.0, 2.0 / 3.0},
                 sg[3] = {1.0 / 2.0, 1.0 / 8.0, 1.0 / 3.0};                 // :1596-1598
    const size_t vb = n * sizeof(double);
    const double *nu_t = p->les_active ? nut : nullptr;
    tosca::launch_fill(rhs_o, n, 0.0, s);                                    // :1526
    for (int i = 0; i < 3; i++) {
        queue_explicit(c, m, p, ucont, ucat, nut, dp, btheta, btheta_o, ab2, farm, ucont_o, rhs, 1);
        tosca::launch_fill(visc, n, 0.0, s);                                 // FormU(RhsVisc, 1, 2)
        if (c->ref_path) {
            tosca::ref_formu(g, m, p, ucont, ucat, nu_t, visc, 1.0, 2, c->ws, s);
        } else {
            tosca::FormUArgs A = formu_args(c, m, p, ucont, ucat, nu_t, visc, 1.0, 2);
            tosca::fused_formu(g, A, c->ws + 3 * (size_t)g.nloc(), s);
        }
        tosca::launch_non_resolved(visc, ucont_o, p, g, s);
        tosca::launch_local_to_global(ucg, ucont, g, s);                     // Ucont
        double gam;
        if (scheme == 0) {
            if (hipMemcpyAsync(bU, ucg, vb, hipMemcpyDeviceToDevice, s) != hipSuccess)
                return set_err(c, TOSCA_ERR_DEVICE, "RK3WCN: copy failed");
            tosca::launch_axpy(bU, dt * wb[i], rhs, n, s);
            tosca::launch_axpy(bU, -1.0 * dt * wd[i], rhs_o, n, s);
            tosca::launch_axpy(bU, dt * wg[i], visc, n, s);
            gam = wg[i];
        } else {
            tosca::launch_scale_to(bU, ucont_o, sa[i], n, s);
            if (i > 0) tosca::launch_axpy(bU, 1.0 - sa[i], ucg, n, s);
            tosca::launch_axpy(bU, dt * sb[i], rhs, n, s);
            tosca::launch_axpy(bU, dt * sg[i], visc, n, s);
            gam = sg[i];
        }
        if (hipMemcpyAsync(utmp, bU, vb, hipMemcpyDeviceToDevice, s) != hipSuccess)   // explicit-Euler guess
            return set_err(c, TOSCA_ERR_DEVICE, "RK3CN: copy failed");
        tosca::launch_axpy(utmp, dt * gam, visc, n, s);
        const double coef = dt * gam;
        auto A = [&](const double *v, double *out) { queue_imex(c, m, p, v, ucont, ucat, nut, ucont_o, coef, out); };
        if ((rc = gmres_device(c, n, A, bU, utmp, kp, kr, K))) return rc;
        tosca::launch_global_to_local(ucont, utmp, g, s);
        tosca::launch_wrap(ucont, 3, g, s);
        if (scheme == 0 && hipMemcpyAsync(rhs_o, rhs, vb, hipMemcpyDeviceToDevice, s) != hipSuccess)
            return set_err(c, TOSCA_ERR_DEVICE, "RK3WCN: copy failed");
    }
    return launched(c, scheme == 0 ? "UeqnRK3CN_W" : "UeqnRK3CN_SO");
}

int tosca_ueqn_abcn(tosca_ctx *c, const tosca_mesh *m, const tosca_params *p, double *ucont, double *ucat,
                    const double *nut, const double *dp, const double *btheta, const double *btheta_o, int ab2,
                    const double *farm, const double *ucont_o, double dt, double *rhs, double *conv_o,
                    const tosca_krylov_params *kp, tosca_krylov_result *kr) {
    int rc = check_explicit(c, m, p, ucont, ucat, nut, dp, btheta, btheta_o, ab2, farm, ucont_o, rhs, 1);
    if (rc) return rc;
    if ((rc = check_krylov(c, kp, kr))) return rc;
    if ((rc = check_imex(c, m, p, ucont, ucat, nut, ucont_o))) return rc;
    const Geo &g = c->geo;
    hipStream_t s = c->stream;
    const long n = 3L * g.mx * g.my * g.mz;
    if ((rc = check_dev(c, conv_o, n * sizeof(double), "conv_o"))) return rc;
    if (conv_o == rhs || conv_o == ucont_o) return set_err(c, TOSCA_ERR_ARG, "conv_o must not alias rhs or ucont_o");
    KrylovWs K;
    if ((rc = ensure_kws(c, kp->restart, 4, K))) return rc;
    double *visc = K.extra, *cnv = K.extra + n, *bU = K.extra + 2 * n, *utmp = K.extra + 3 * n;
    const double *nu_t = p->les_active ? nut : nullptr;
    queue_explicit(c, m, p, ucont, ucat, nut, dp, btheta, btheta_o, ab2, farm, ucont_o, rhs, 1);   // :1675
    for (int mode = 2; mode >= 1; mode--) {                                  // :1678-1681
        double *out = mode == 2 ? visc : cnv;
        tosca::launch_fill(out, n, 0.0, s);
        if (c->ref_path) {
            tosca::ref_formu(g, m, p, ucont, ucat, nu_t, out, 1.0, mode, c->ws, s);
        } else {
            tosca::FormUArgs A = formu_args(c, m, p, ucont, ucat, nu_t, out, 1.0, mode);
            tosca::fused_formu(g, A, c->ws + 3 * (size_t)g.nloc(), s);
        }
        tosca::launch_non_resolved(out, ucont_o, p, g, s);
    }
    const size_t vb = n * sizeof(double);
    if (hipMemcpyAsync(bU, rhs, vb, hipMemcpyDeviceToDevice, s) != hipSuccess)   // :1684
        return set_err(c, TOSCA_ERR_DEVICE, "ABCN: copy failed");
    if (ab2) {
        tosca::launch_axpy(bU, 0.5, cnv, n, s);                               // :1690-1691
        tosca::launch_axpy(bU, -0.5, conv_o, n, s);
        tosca::launch_axpy(bU, 0.5, visc, n, s);                              // :1699-1700
    } else {
        tosca::launch_axpy(bU, 1.0, rhs, n, s);                               // :1695
    }
    tosca::launch_scale_to(bU, bU, dt, n, s);                                 // :1703
    tosca::launch_axpy(bU, 1.0, ucont_o, n, s);                               // :1706
    tosca::launch_fill(utmp, n, 0.0, s);                                      // :1709-1713
    tosca::launch_axpy(utmp, ab2 ? 0.5 * dt : dt, visc, n, s);
    tosca::launch_axpy(utmp, 1.0, bU, n, s);
    const double coef = dt * (ab2 ? 0.5 : 1.0);
    auto A = [&](const double *v, double *out) { queue_imex(c, m, p, v, ucont, ucat, nut, ucont_o, coef, out); };
    if ((rc = gmres_device(c, n, A, bU, utmp, kp, kr, K))) return rc;
    tosca::launch_global_to_local(ucont, utmp, g, s);                         // :1723-1725
    tosca::launch_wrap(ucont, 3, g, s);
    if (hipMemcpyAsync(conv_o, cnv, vb, hipMemcpyDeviceToDevice, s) != hipSuccess)   // :1730
        return set_err(c, TOSCA_ERR_DEVICE, "ABCN: copy failed");
    return launched(c, "UeqnABCN");
}

int tosca_cn_rhs(tosca_ctx *c, const tosca_params *p, const double *dp, const double *btheta,
                 const double *btheta_o, int ab2, const double *rhs_o, const double *farm, double *bU) {
    int rc = check_ctx(c);
    if (rc) return rc;
    if ((rc = check_params(c, p))) return rc;
    set_geo_per(c, p);
    const Geo &g = c->geo;
    const long n = 3L * g.mx * g.my * g.mz;
    const size_t vb = n * sizeof(double);
    if ((rc = check_dev(c, dp, vb, "dp")) || (rc = check_dev(c, bU, vb, "bU"))) return rc;
    if (btheta) {
        if ((rc = check_dev(c, btheta, vb, "btheta"))) return rc;
        if (ab2 && (rc = check_dev(c, btheta_o, vb, "btheta_o"))) return rc;
    }
    if (ab2 && (rc = check_dev(c, rhs_o, vb, "rhs_o"))) return rc;
    if (farm && (rc = check_dev(c, farm, vb, "farm"))) return rc;
    if (bU == dp || bU == btheta || bU == btheta_o || bU == rhs_o || bU == farm)
        return set_err(c, TOSCA_ERR_ARG, "bU must not alias an input");
    // -dP (ueqn_solvers.c:1102-1103), buoyancy AB2 or first step (:1125-1135), 1/2 Rhs_o after the
    // first step (:1151-1152), farm (:1155-1156), in one pass
    tosca::launch_cn_rhs(bU, dp, btheta, (btheta && ab2) ? btheta_o : nullptr, ab2 ? rhs_o : nullptr, farm,
                         ab2 ? 3.0 / 2.0 : 1.0, n, c->stream);
    return launched(c, "UeqnSNES constant RHS");
}

int tosca_snes_func_eval(tosca_ctx *c, const tosca_mesh *m, const tosca_params *p, const double *u, double *ucont,
                         double *ucat, const double *nut, const double *bU, const double *ucont_o, double dt, int ab2,
                         double *F) {
    int rc = check_ctx(c);
    if (rc) return rc;
    if ((rc = check_imex(c, m, p, ucont, ucat, nut, ucont_o))) return rc;
    const Geo &g = c->geo;
    const long n = 3L * g.mx * g.my * g.mz;
    const size_t vb = n * sizeof(double);
    if ((rc = check_dev(c, u, vb, "u")) || (rc = check_dev(c, bU, vb, "bU")) || (rc = check_dev(c, F, vb, "F"))) return rc;
    if (F == u || F == bU || F == ucont_o) return set_err(c, TOSCA_ERR_ARG, "F must not alias an input");
    hipStream_t s = c->stream;
    tosca::launch_global_to_local(ucont, u, g, s);                           // :1195-1199
    tosca::launch_wrap(ucont, 3, g, s);
    tosca::launch_no_penetration(ucont, p, g, s);                            // :1202
    tosca::launch_reset_face_periodic(ucont, 3, g, s);                       // :1205
    tosca::launch_contra_to_cart(ucont, m, ucat, g, s);                      // :1208
    tosca::launch_reset_cell_periodic(ucat, 3, g, s);                        // :1221
    if (hipMemcpyAsync(F, bU, vb, hipMemcpyDeviceToDevice, s) != hipSuccess)  // :1224
        return set_err(c, TOSCA_ERR_DEVICE, "SNESFuncEval: copy failed");
    const double scale = ab2 ? 0.5 : 1.0;                                     // :1227
    const double *nu_t = p->les_active ? nut : nullptr;
    if (c->ref_path) {                                                        // :1230
        tosca::ref_formu(g, m, p, ucont, ucat, nu_t, F, scale, 0, c->ws, s);
    } else {
        tosca::FormUArgs A = formu_args(c, m, p, ucont, ucat, nu_t, F, scale, 0);
        tosca::fused_formu(g, A, c->ws + 3 * (size_t)g.nloc(), s);
    }
    tosca::launch_snes_finish(F, u, ucont_o, dt, p, g, s);                    // :1264-1271
    return launched(c, "SNESFuncEval");
}

int tosca_grad_p(tosca_ctx *c, const tosca_mesh *m, const tosca_params *p, const double *pres, double *dp) {
    int rc = check_ctx(c);
    if (rc) return rc;
    if ((rc = check_params(c, p))) return rc;
    set_geo_per(c, p);
    const Geo &g = c->geo;
    const size_t n = (size_t)g.nloc();
    if ((rc = check_mesh(c, m, true, false, true))) return rc;
    if ((rc = check_dev(c, pres, n * sizeof(double), "p")) ||
        (rc = check_dev(c, dp, 3 * (size_t)g.mx * g.my * g.mz * sizeof(double), "dp")))
        return rc;
    tosca::launch_grad_p(pres, m, dp, g, c->stream);
    return launched(c, "GradP");
}

int tosca_formt(tosca_ctx *c, const tosca_mesh *m, const tosca_params *p, const tosca_teqn_params *tp,
                const double *ucont, const double *theta, const double *nut, const double *kt, double *rhs,
                double scale, int form_mode) {
    int rc = check_ctx(c);
    if (rc) return rc;
    if ((rc = check_params(c, p))) return rc;
    if (!tp) return set_err(c, TOSCA_ERR_ARG, "teqn params is NULL");
    if (form_mode < 0 || form_mode > 2) return set_err(c, TOSCA_ERR_ARG, "form_mode = %d is not 0, 1 or 2", form_mode);
    if (!(tp->pr > 0.0)) return set_err(c, TOSCA_ERR_ARG, "Pr = %g must be positive", tp->pr);
    const bool les = p->les_active != 0;
    if (les) {
        if (p->les_model < 0 || p->les_model > TOSCA_LES_BAMD)
            return set_err(c, TOSCA_ERR_ARG, "les_model = %d is not an LES model", p->les_model);
        if (p->les_model == TOSCA_LES_STABILITY)
            return set_err(c, TOSCA_ERR_UNSUPPORTED, "the stability-based LES model is not supported");
        if (p->les_model == TOSCA_LES_BV || p->les_model == TOSCA_LES_BAMD)
            return set_err(c, TOSCA_ERR_UNSUPPORTED, "structural LES models (bardinaVreman/bardinaAMD) are not supported");
    }
    set_geo_per(c, p);
    const Geo &g = c->geo;
    const size_t n = (size_t)g.nloc(), ng = (size_t)g.mx * g.my * g.mz * sizeof(double);
    if ((rc = check_mesh(c, m, true, true, true))) return rc;
    if ((rc = check_dev(c, ucont, 3 * n * sizeof(double), "ucont")) ||
        (rc = check_dev(c, theta, n * sizeof(double), "theta")) || (rc = check_dev(c, rhs, ng, "rhs")))
        return rc;
    if (les) {
        const bool ktm = p->les_model == TOSCA_LES_AMD || p->les_model == TOSCA_LES_DSM ||
                         p->les_model == TOSCA_LES_DLASI || p->les_model == TOSCA_LES_DLASD ||
                         p->les_model == TOSCA_LES_DPASD;
        if ((rc = check_dev(c, nut, n * sizeof(double), "nut"))) return rc;
        if (ktm && (rc = check_dev(c, kt, n * sizeof(double), "kt"))) return rc;
        if ((tp->wall_t[0] || tp->wall_t[1]) &&
            (rc = check_dev(c, tp->q_wall_i, 3 * (size_t)g.mz * g.my * sizeof(double), "q_wall_i")))
            return rc;
        if ((tp->wall_t[2] || tp->wall_t[3]) &&
            (rc = check_dev(c, tp->q_wall_j, 3 * (size_t)g.mz * g.mx * sizeof(double), "q_wall_j")))
            return rc;
    }
    if ((rc = ensure_ws(c, ws_need(c, p)))) return rc;
    tosca::ref_formt(g, m, p, tp, ucont, theta, nut, kt, rhs, scale, form_mode, c->ws, c->stream);
    return launched(c, "FormT");
}

int tosca_update_contravariant_bcs(tosca_ctx *c, const tosca_mesh *m, const tosca_params *p, double *ucont,
                                   const double *ucat) {
    int rc = check_ctx(c);
    if (rc) return rc;
    if ((rc = check_params(c, p))) return rc;
    set_geo_per(c, p);
    const size_t v = 3 * (size_t)c->geo.nloc() * sizeof(double);
    if ((rc = require_checked(c, m))) return rc;
    if ((rc = check_dev(c, m->icsi, v, "mesh.icsi")) || (rc = check_dev(c, m->jeta, v, "mesh.jeta")) ||
        (rc = check_dev(c, m->kzet, v, "mesh.kzet")) || (rc = check_dev(c, ucont, v, "ucont")) ||
        (rc = check_dev(c, ucat, v, "ucat")))
        return rc;
    tosca::launch_contravariant_bcs(ucont, ucat, m, p, c->geo, c->stream);
    return launched(c, "UpdateContravariantBCs");
}

int tosca_update_cartesian_bcs(tosca_ctx *c, const tosca_mesh *m, const tosca_params *p, double *ucat,
                               const double *u_fixed, double *ustar, const tosca_inlet *inlet) {
    int rc = check_ctx(c);
    if (rc) return rc;
    if ((rc = check_params(c, p))) return rc;
    bool cabot = false, smooth = false;
    for (int q = 0; q < 6; q++) {
        if (p->bc_u[q] == TOSCA_BCU_GENERIC && !(q == 4 && inlet))
            return set_err(c, TOSCA_ERR_UNSUPPORTED, "bc_u[%d]: inlet-function velocity patches are covered on kLeft only, "
                                                     "with their profile in inlet", q);
        if (p->bc_u[q] == TOSCA_BCU_WALLFUNCTION) {
            if (q >= 4)
                return set_err(c, TOSCA_ERR_UNSUPPORTED, "bc_u[%d]: k patches have no velocity wall model", q);
            if (p->wall_u[q] != -1 && p->wall_u[q] != -3)
                return set_err(c, TOSCA_ERR_ARG, "bc_u[%d] is a wall function but wall_u[%d] = %d is neither -1 (Cabot) "
                                                 "nor -3 (Shumann)", q, q, p->wall_u[q]);
            if (p->wall_u[q] == -1) {
                cabot = true;
                const int lead = (p->wall_u[q & ~1] == -1) ? (q & ~1) : q;     // the roughness this patch uses
                if (!(p->wall_rough[lead] > 1.e-19)) smooth = true;
            }
        }
        if (p->bc_u[q] == TOSCA_BCU_FIXEDVALUE && !u_fixed)
            return set_err(c, TOSCA_ERR_ARG, "bc_u[%d] is fixedValue but u_fixed is NULL", q);
    }
    set_geo_per(c, p);
    if ((rc = check_mesh(c, m, false, true, false))) return rc;
    const size_t n = (size_t)c->geo.nloc();
    if ((rc = check_dev(c, ucat, 3 * n * sizeof(double), "ucat"))) return rc;
    if (cabot) {
        if (!(p->nu > 0.0)) return set_err(c, TOSCA_ERR_ARG, "Cabot wall function with nu = %g", p->nu);
        if (!ustar) return set_err(c, TOSCA_ERR_ARG, "a Cabot wall-function patch writes ustar, which is NULL");
        if ((rc = check_dev(c, m->aj, n * sizeof(double), "mesh.aj")) || (rc = check_dev(c, ustar, n * sizeof(double), "ustar")))
            return rc;
    }
    if (inlet && p->bc_u[4] == TOSCA_BCU_GENERIC) {
        const int t = inlet->type;
        if (t < TOSCA_INLET_POWERLAW || t > TOSCA_INLET_SINUSOIDAL)
            return set_err(c, TOSCA_ERR_UNSUPPORTED, "inlet function type %d: the power law (1), log law (2), mapped (3) "
                                                     "and interpolated (4) inflow, Nieuwstadt (5) and sinusoidal (6) "
                                                     "profiles are covered", t);
        if ((rc = check_dev(c, inlet->cent, 3 * n * sizeof(double), "inlet.cent"))) return rc;
        const size_t plane_pts = (size_t)c->geo.mx * c->geo.my;        // [my][mx] tables
        if (t == TOSCA_INLET_POWERLAW && (rc = check_dev(c, inlet->noise, plane_pts * sizeof(int), "inlet.noise", 4)))
            return rc;
        if (t == TOSCA_INLET_MAPPED || t == TOSCA_INLET_INTERP) {
            if (inlet->n1 < 1 || inlet->n2 < 1)
                return set_err(c, TOSCA_ERR_ARG, "inflow plane size n1 = %d, n2 = %d", inlet->n1, inlet->n2);
            if (t == TOSCA_INLET_MAPPED && (inlet->prds1 < 1 || inlet->prds2 < 1))
                return set_err(c, TOSCA_ERR_ARG, "inflow periods n1Periods = %d, n2Periods = %d", inlet->prds1, inlet->prds2);
            const size_t np = (size_t)(inlet->n1 + 2) * (inlet->n2 + 2);
            if ((rc = check_dev(c, inlet->plane, 3 * np * sizeof(double), "inlet.plane"))) return rc;
        }
        if (t == TOSCA_INLET_INTERP) {
            if ((rc = check_dev(c, inlet->cells, 8 * plane_pts * sizeof(int), "inlet.cells", 4)) ||
                (rc = check_dev(c, inlet->weights, 4 * plane_pts * sizeof(double), "inlet.weights")))
                return rc;
            if (inlet->spline &&
                ((rc = check_dev(c, inlet->cells_1, 12 * plane_pts * sizeof(int), "inlet.cells_1", 4)) ||
                 (rc = check_dev(c, inlet->cells_2, 12 * plane_pts * sizeof(int), "inlet.cells_2", 4)) ||
                 (rc = check_dev(c, inlet->weights_1, 6 * plane_pts * sizeof(double), "inlet.weights_1")) ||
                 (rc = check_dev(c, inlet->weights_2, 6 * plane_pts * sizeof(double), "inlet.weights_2"))))
                return rc;
            if (inlet->shift2 &&
                ((rc = check_dev(c, inlet->y_ids, (size_t)c->geo.mx * sizeof(int), "inlet.y_ids", 4)) ||
                 (rc = check_dev(c, inlet->y_weights, (size_t)c->geo.mx * sizeof(double), "inlet.y_weights"))))
                return rc;
        }
    }
    if (smooth && !c->cabot) {
        if (hipMalloc(&c->cabot, tosca::cabot_table_doubles() * sizeof(double)) != hipSuccess) {
            (void)hipGetLastError();
            c->cabot = nullptr;
            return set_err(c, TOSCA_ERR_DEVICE, "Cabot table allocation of %zu bytes failed",
                           tosca::cabot_table_doubles() * sizeof(double));
        }
        tosca::launch_cabot_table(c->cabot, c->stream);
    }
    tosca::launch_cartesian_bcs(ucat, u_fixed, ustar, c->cabot, (inlet && p->bc_u[4] == TOSCA_BCU_GENERIC) ? inlet : nullptr,
                                m, p, c->geo, c->stream);
    return launched(c, "UpdateCartesianBCs");
}

int tosca_update_temperature_bcs(tosca_ctx *c, const tosca_mesh *m, const tosca_params *p, const int *bc_t,
                                 const double *t_value, double *theta) {
    int rc = check_ctx(c);
    if (rc) return rc;
    if ((rc = check_params(c, p))) return rc;
    if (!bc_t || !t_value) return set_err(c, TOSCA_ERR_ARG, "bc_t / t_value is NULL");
    for (int q = 0; q < 6; q++) {
        if (bc_t[q] < TOSCA_BCT_ZEROGRADIENT || bc_t[q] > TOSCA_BCT_SURFACETEMP)
            return set_err(c, TOSCA_ERR_ARG, "bc_t[%d] = %d is not a TOSCA_BCT_* kind", q, bc_t[q]);
        if (bc_t[q] == TOSCA_BCT_SURFACETEMP && q != 2)
            return set_err(c, TOSCA_ERR_ARG, "bc_t[%d]: the surface-temperature ghost applies to the jLeft patch only", q);
    }
    set_geo_per(c, p);
    if ((rc = check_mesh(c, m, false, true, false))) return rc;
    if ((rc = check_dev(c, theta, (size_t)c->geo.nloc() * sizeof(double), "theta"))) return rc;
    tosca::launch_temperature_bcs(theta, bc_t, t_value, m, c->geo, c->stream);
    return launched(c, "UpdateTemperatureBCs");
}

int tosca_update_cs(tosca_ctx *c, const tosca_mesh *m, const tosca_params *p, const double *ucat, double *cs) {
    int rc = check_ctx(c);
    if (rc) return rc;
    if ((rc = check_params(c, p)) || (rc = check_les_model(c, p))) return rc;
    set_geo_per(c, p);
    const size_t n = (size_t)c->geo.nloc();
    if ((rc = check_dev(c, cs, n * sizeof(double), "cs"))) return rc;
    if (p->les_model != TOSCA_LES_SMAGORINSKY) {
        if ((rc = check_mesh(c, m, false, true, false))) return rc;
        if ((rc = check_dev(c, ucat, 3 * n * sizeof(double), "ucat"))) return rc;
    }
    if ((rc = ensure_ws(c, ws_need(c, p)))) return rc;
    if (!c->ref_path && (p->les_model == TOSCA_LES_DSM || p->les_model == TOSCA_LES_DPASD)) {
        if (p->les_model == TOSCA_LES_DSM) tosca::fused_dsm<double>(c->geo, m, ucat, p->max_cs, cs, nullptr, c->ws, c->stream);
        else tosca::fused_dpasd<double>(c->geo, m, ucat, p->max_cs, cs, nullptr, c->ws, c->stream);
        tosca::launch_wrap(cs, 1, c->geo, c->stream);
        tosca::launch_reset_cell_periodic(cs, 1, c->geo, c->stream);
    } else {
        tosca::ref_update_cs(c->geo, m, p, ucat, cs, c->ws, c->stream);
    }
    return launched(c, "UpdateCs");
}

int tosca_update_nut(tosca_ctx *c, const tosca_mesh *m, const tosca_params *p, const double *ucat, const double *cs,
                     double *nut, double *ksgs) {
    int rc = check_ctx(c);
    if (rc) return rc;
    if ((rc = check_params(c, p)) || (rc = check_les_model(c, p))) return rc;
    set_geo_per(c, p);
    const size_t n = (size_t)c->geo.nloc();
    if ((rc = check_mesh(c, m, false, true, false))) return rc;
    if ((rc = check_dev(c, ucat, 3 * n * sizeof(double), "ucat")) || (rc = check_dev(c, cs, n * sizeof(double), "cs")) ||
        (rc = check_dev(c, nut, n * sizeof(double), "nut")))
        return rc;
    if (ksgs && (rc = check_dev(c, ksgs, n * sizeof(double), "ksgs"))) return rc;
    if (c->ref_path) {
        tosca::ref_update_nut(c->geo, m, p, ucat, cs, nut, ksgs, c->stream);
    } else {
        tosca::fused_nut(c->geo, m, ucat, cs, nut, ksgs, c->stream);
        tosca::zero_outside_interior(nut, c->geo, c->stream);
        if (ksgs) tosca::zero_outside_interior(ksgs, c->geo, c->stream);
        tosca::nut_bcs(nut, p, c->geo, c->stream);
        if (ksgs) tosca::launch_wrap(ksgs, 1, c->geo, c->stream);
    }
    return launched(c, "UpdateNut");
}

int tosca_buoyancy(tosca_ctx *c, const tosca_mesh *m, const tosca_params *p, const double *theta, double scale,
                   double *btheta, double *gcont) {
    int rc = check_ctx(c);
    if (rc) return rc;
    if ((rc = check_params(c, p))) return rc;
    set_geo_per(c, p);
    const Geo &g = c->geo;
    const size_t n = (size_t)g.nloc(), ng = 3 * (size_t)g.mx * g.my * g.mz * sizeof(double);
    if ((rc = require_checked(c, m))) return rc;
    if ((rc = check_dev(c, m->icsi, 3 * n * sizeof(double), "mesh.icsi")) ||
        (rc = check_dev(c, m->jeta, 3 * n * sizeof(double), "mesh.jeta")) ||
        (rc = check_dev(c, m->kzet, 3 * n * sizeof(double), "mesh.kzet")) ||
        (rc = check_dev(c, theta, n * sizeof(double), "theta")) || (rc = check_dev(c, btheta, ng, "btheta")))
        return rc;
    if (gcont && (rc = check_dev(c, gcont, ng, "gcont"))) return rc;
    if (!(p->t_ref != 0.0)) return set_err(c, TOSCA_ERR_ARG, "t_ref must be non-zero");
    tosca::ref_buoyancy(g, m, p, theta, scale, btheta, gcont, c->stream);
    return launched(c, "Buoyancy");
}

}  // extern "C"

namespace {
// tosca_rhs_les_step and tosca_rhs_les_step_farm (farm: the VecAXPY of sourceFarmCont, or null)
int rhs_les_step_impl(tosca_ctx *c, const tosca_mesh *m, const tosca_params *p, const double *ucont,
                      const double *ucat, const double *theta, const double *farm, double *rhs, double scale,
                      double *cs, double *nut, double *btheta, double buoy_scale) {
    int rc = check_ctx(c);
    if (rc) return rc;
    if ((rc = check_params(c, p))) return rc;
    if (!p->les_active) return set_err(c, TOSCA_ERR_ARG, "tosca_rhs_les_step needs les_active = 1");
    const size_t ngv = 3 * (size_t)c->geo.mx * c->geo.my * c->geo.mz;
    if (farm && (rc = check_dev(c, farm, ngv * sizeof(double), "farm"))) return rc;
    if (c->ref_path) {
        if (farm) {
            if ((rc = check_dev(c, rhs, ngv * sizeof(double), "rhs"))) return rc;
            tosca::launch_axpy(rhs, 1.0, farm, (long)ngv, c->stream);
        }
        if ((rc = tosca_update_cs(c, m, p, ucat, cs))) return rc;
        if ((rc = tosca_update_nut(c, m, p, ucat, cs, nut, nullptr))) return rc;
        if ((rc = tosca_buoyancy(c, m, p, theta, buoy_scale, btheta, nullptr))) return rc;
        return tosca_formu(c, m, p, ucont, ucat, nut, rhs, scale, 0);
    }
    // fused path: validate everything the four calls would, then queue the fused kernels
    if ((rc = check_les_model(c, p))) return rc;
    if (p->div_scheme != TOSCA_DIV_CENTRAL && p->div_scheme != TOSCA_DIV_QUICK)
        return set_err(c, TOSCA_ERR_UNSUPPORTED, "div_scheme = %d: only central (0) and QUICK (1) are supported", p->div_scheme);
    if (!(p->t_ref != 0.0)) return set_err(c, TOSCA_ERR_ARG, "t_ref must be non-zero");
    set_geo_per(c, p);
    const Geo &g = c->geo;
    const size_t n = (size_t)g.nloc(), v = 3 * n * sizeof(double), ng = 3 * (size_t)g.mx * g.my * g.mz * sizeof(double);
    if ((rc = check_mesh(c, m, true, true, true))) return rc;
    if ((rc = check_dev(c, ucont, v, "ucont")) || (rc = check_dev(c, ucat, v, "ucat")) ||
        (rc = check_dev(c, theta, n * sizeof(double), "theta")) || (rc = check_dev(c, rhs, ng, "rhs")) ||
        (rc = check_dev(c, cs, n * sizeof(double), "cs")) || (rc = check_dev(c, nut, n * sizeof(double), "nut")) ||
        (rc = check_dev(c, btheta, ng, "btheta")))
        return rc;
    if (p->bc_u[0] == TOSCA_BCU_WALLFUNCTION || p->bc_u[1] == TOSCA_BCU_WALLFUNCTION)
        if ((rc = check_dev(c, p->tau_wall_i, 3 * (size_t)g.mz * g.my * sizeof(double), "tau_wall_i"))) return rc;
    if (p->bc_u[2] == TOSCA_BCU_WALLFUNCTION || p->bc_u[3] == TOSCA_BCU_WALLFUNCTION)
        if ((rc = check_dev(c, p->tau_wall_j, 3 * (size_t)g.mz * g.mx * sizeof(double), "tau_wall_j"))) return rc;
    if ((rc = ensure_ws(c, ws_need(c, p)))) return rc;
    queue_step<double>(c, m, p, ucont, ucat, theta, rhs, scale, cs, nut, btheta, buoy_scale, nullptr, farm);
    return launched(c, "rhs_les_step");
}
}  // namespace

extern "C" {

int tosca_rhs_les_step(tosca_ctx *c, const tosca_mesh *m, const tosca_params *p, const double *ucont,
                       const double *ucat, const double *theta, double *rhs, double scale, double *cs, double *nut,
                       double *btheta, double buoy_scale) {
    return rhs_les_step_impl(c, m, p, ucont, ucat, theta, nullptr, rhs, scale, cs, nut, btheta, buoy_scale);
}

int tosca_rhs_les_step_farm(tosca_ctx *c, const tosca_mesh *m, const tosca_params *p, const double *ucont,
                            const double *ucat, const double *theta, const double *farm, double *rhs, double scale,
                            double *cs, double *nut, double *btheta, double buoy_scale) {
    if (c && !farm) return set_err(c, TOSCA_ERR_ARG, "farm is NULL");
    return rhs_les_step_impl(c, m, p, ucont, ucat, theta, farm, rhs, scale, cs, nut, btheta, buoy_scale);
}

int tosca_rhs_les_step_host(tosca_ctx *c, const tosca_mesh *m, const tosca_params *p, const double *ucont,
                            const double *ucat, const double *theta, double *rhs, double scale, double *cs,
                            double *nut, double *btheta, double buoy_scale, double *times_ms) {
    int rc = check_ctx(c);
    if (rc) return rc;
    if (!ucont || !ucat || !theta || !rhs || !cs || !nut || !btheta)
        return set_err(c, TOSCA_ERR_ARG, "host arrays must not be NULL");
    if ((rc = check_params(c, p))) return rc;
    set_geo_per(c, p);
    const Geo &g = c->geo;
    const size_t n = (size_t)g.nloc(), ng = 3 * (size_t)g.mx * g.my * g.mz;
    const size_t need = 3 * n + 3 * n + n + ng + n + n + ng;
    if (c->hs_elems < need) {
        if (c->hs) { hipStreamSynchronize(c->stream); hipFree(c->hs); c->hs = nullptr; c->hs_elems = 0; }
        if (hipMalloc(&c->hs, need * sizeof(double)) != hipSuccess) {
            (void)hipGetLastError();
            return set_err(c, TOSCA_ERR_DEVICE, "staging allocation of %zu bytes failed", need * sizeof(double));
        }
        c->hs_elems = need;
    }
    double *dV = c->hs, *dU = dV + 3 * n, *dT = dU + 3 * n, *dR = dT + n, *dC = dR + ng, *dN = dC + n, *dB = dN + n;
    hipStream_t s = c->stream;
    hipEvent_t e[4];
    for (int q = 0; q < 4; q++) hipEventCreate(&e[q]);
    hipEventRecord(e[0], s);
    hipError_t he = hipSuccess;
    auto cp = [&](void *dst, const void *src, size_t bytes, hipMemcpyKind kind) {
        if (he == hipSuccess) he = hipMemcpyAsync(dst, src, bytes, kind, s);
    };
    cp(dV, ucont, 3 * n * sizeof(double), hipMemcpyHostToDevice);
    cp(dU, ucat, 3 * n * sizeof(double), hipMemcpyHostToDevice);
    cp(dT, theta, n * sizeof(double), hipMemcpyHostToDevice);
    cp(dR, rhs, ng * sizeof(double), hipMemcpyHostToDevice);             // rhs accumulates
    cp(dN, nut, n * sizeof(double), hipMemcpyHostToDevice);              // ghosts the step leaves alone
    cp(dC, cs, n * sizeof(double), hipMemcpyHostToDevice);
    hipEventRecord(e[1], s);
    if (he == hipSuccess) rc = tosca_rhs_les_step(c, m, p, dV, dU, dT, dR, scale, dC, dN, dB, buoy_scale);
    hipEventRecord(e[2], s);
    if (he == hipSuccess && rc == TOSCA_OK) {
        cp(rhs, dR, ng * sizeof(double), hipMemcpyDeviceToHost);
        cp(cs, dC, n * sizeof(double), hipMemcpyDeviceToHost);
        cp(nut, dN, n * sizeof(double), hipMemcpyDeviceToHost);
        cp(btheta, dB, ng * sizeof(double), hipMemcpyDeviceToHost);
    }
    hipEventRecord(e[3], s);
    const hipError_t se = hipStreamSynchronize(s);
    if (times_ms && rc == TOSCA_OK && he == hipSuccess && se == hipSuccess) {
        float a = 0, b = 0, d = 0;
        hipEventElapsedTime(&a, e[0], e[1]);
        hipEventElapsedTime(&b, e[1], e[2]);
        hipEventElapsedTime(&d, e[2], e[3]);
        times_ms[0] = a + d;
        times_ms[1] = b;
    }
    for (int q = 0; q < 4; q++) hipEventDestroy(e[q]);
    if (rc) return rc;
    if (he != hipSuccess || se != hipSuccess)
        return set_err(c, TOSCA_ERR_DEVICE, "host staging failed: %s", hipGetErrorString(he != hipSuccess ? he : se));
    return TOSCA_OK;
}

int tosca_rhs_les_step_mixed(tosca_ctx *c, const tosca_mesh *m, const tosca_params *p, const float *ucont,
                             const float *ucat, const float *theta, double *rhs, double scale, double *cs, float *nut,
                             double *btheta, double buoy_scale) {
    int rc = check_mixed(c, m, p, ucont, ucat, theta, rhs, cs, nut, btheta);
    if (rc) return rc;
    queue_step<float>(c, m, p, ucont, ucat, theta, rhs, scale, cs, nut, btheta, buoy_scale);
    return launched(c, "rhs_les_step_mixed");
}

int tosca_rhs_les_step_mixed_farm(tosca_ctx *c, const tosca_mesh *m, const tosca_mesh_f32 *m32, const tosca_params *p,
                                  const float *ucont, const float *ucat, const float *theta, const double *farm,
                                  double *rhs, double scale, double *cs, float *nut, double *btheta, double buoy_scale) {
    int rc = check_mixed(c, m, p, ucont, ucat, theta, rhs, cs, nut, btheta);
    if (rc) return rc;
    if ((rc = check_dev(c, farm, 3 * (size_t)c->geo.mx * c->geo.my * c->geo.mz * sizeof(double), "farm"))) return rc;
    if (m32) {
        if ((rc = check_mesh32(c, m32))) return rc;
        queue_step<float, float>(c, m, p, ucont, ucat, theta, rhs, scale, cs, nut, btheta, buoy_scale, m32, farm);
    } else {
        queue_step<float>(c, m, p, ucont, ucat, theta, rhs, scale, cs, nut, btheta, buoy_scale, nullptr, farm);
    }
    return launched(c, "rhs_les_step_mixed_farm");
}

int tosca_project_uadm(tosca_ctx *c, const tosca_params *p, const double *cent, const tosca_uadm *F, double *scat) {
    int rc = check_ctx(c);
    if (rc) return rc;
    if ((rc = check_params(c, p))) return rc;
    if (!F) return set_err(c, TOSCA_ERR_ARG, "farm is NULL");
    if (F->n_rotors < 0 || (F->n_rotors > 0 && F->n_points < 1))
        return set_err(c, TOSCA_ERR_ARG, "farm: n_rotors = %d, n_points = %d", F->n_rotors, F->n_points);
    if (!(F->nsigma > 0.0)) return set_err(c, TOSCA_ERR_ARG, "farm: nsigma must be positive");
    set_geo_per(c, p);
    const Geo &g = c->geo;
    const size_t n = (size_t)g.nloc();
    if ((rc = check_dev(c, cent, 3 * n * sizeof(double), "cent")) || (rc = check_dev(c, scat, 3 * n * sizeof(double), "scat")))
        return rc;
    if (F->n_rotors > 0) {
        const size_t nr = (size_t)F->n_rotors, np = nr * (size_t)F->n_points;
        if ((rc = check_dev(c, F->center, 3 * nr * sizeof(double), "farm.center")) ||
            (rc = check_dev(c, F->sphere, nr * sizeof(double), "farm.sphere")) ||
            (rc = check_dev(c, F->eps, nr * sizeof(double), "farm.eps")) ||
            (rc = check_dev(c, F->points, 3 * np * sizeof(double), "farm.points")) ||
            (rc = check_dev(c, F->force, 3 * np * sizeof(double), "farm.force")))
            return rc;
    }
    tosca::launch_uadm_project(cent, F, scat, g, c->stream);
    return launched(c, "project_uadm");
}

int tosca_body_force_to_cont(tosca_ctx *c, const tosca_mesh *m, const tosca_params *p, const double *scat,
                             double *scont) {
    int rc = check_ctx(c);
    if (rc) return rc;
    if ((rc = check_params(c, p))) return rc;
    set_geo_per(c, p);
    const Geo &g = c->geo;
    if ((rc = check_mesh(c, m, true, false, false))) return rc;
    if ((rc = check_dev(c, scat, 3 * (size_t)g.nloc() * sizeof(double), "scat")) ||
        (rc = check_dev(c, scont, 3 * (size_t)g.mx * g.my * g.mz * sizeof(double), "scont")))
        return rc;
    tosca::launch_body_force_to_cont(scat, m, scont, g, c->stream);
    return launched(c, "body_force_to_cont");
}

int tosca_rhs_les_step_mixed_m32(tosca_ctx *c, const tosca_mesh *m, const tosca_mesh_f32 *m32, const tosca_params *p,
                                 const float *ucont, const float *ucat, const float *theta, double *rhs, double scale,
                                 double *cs, float *nut, double *btheta, double buoy_scale) {
    int rc = check_mixed(c, m, p, ucont, ucat, theta, rhs, cs, nut, btheta);
    if (rc) return rc;
    if ((rc = check_mesh32(c, m32))) return rc;
    queue_step<float, float>(c, m, p, ucont, ucat, theta, rhs, scale, cs, nut, btheta, buoy_scale, m32);
    return launched(c, "rhs_les_step_mixed_m32");
}

int tosca_rhs_vjp(tosca_ctx *c, const tosca_mesh *m, const tosca_params *p, const double *ucont, const double *ucat,
                  const double *theta, const double *cs, const double *rbar, double scale, double buoy_scale,
                  double *ucont_bar, double *ucat_bar, double *theta_bar, double *cs_bar) {
    int rc = check_ctx(c);
    if (rc) return rc;
    if ((rc = check_params(c, p))) return rc;
    if (p->les_active && p->les_model != TOSCA_LES_SMAGORINSKY && p->les_model != TOSCA_LES_DSM &&
        p->les_model != TOSCA_LES_DPASD)
        return set_err(c, TOSCA_ERR_UNSUPPORTED, "LES model %d: nu_t = Cs Delta^2 |S| models only", p->les_model);
    if (p->div_scheme != TOSCA_DIV_CENTRAL && p->div_scheme != TOSCA_DIV_QUICK)
        return set_err(c, TOSCA_ERR_UNSUPPORTED, "div_scheme = %d: only central (0) and QUICK (1) are supported", p->div_scheme);
    if (!(p->t_ref != 0.0)) return set_err(c, TOSCA_ERR_ARG, "t_ref must be non-zero");
    set_geo_per(c, p);
    const Geo &g = c->geo;
    const size_t n = (size_t)g.nloc(), v = 3 * n * sizeof(double), ng = 3 * (size_t)g.mx * g.my * g.mz * sizeof(double);
    if ((rc = check_mesh(c, m, true, true, true))) return rc;
    if ((rc = check_dev(c, ucont, v, "ucont")) || (rc = check_dev(c, ucat, v, "ucat")) ||
        (rc = check_dev(c, theta, n * sizeof(double), "theta")) || (rc = check_dev(c, cs, n * sizeof(double), "cs")) ||
        (rc = check_dev(c, rbar, ng, "rbar")) || (rc = check_dev(c, ucont_bar, v, "ucont_bar")) ||
        (rc = check_dev(c, ucat_bar, v, "ucat_bar")) || (rc = check_dev(c, theta_bar, n * sizeof(double), "theta_bar")) ||
        (rc = check_dev(c, cs_bar, n * sizeof(double), "cs_bar")))
        return rc;
    // nu_t, fpbar, the nu_eff cotangents and the ucat-cotangent packets of the faces and cells
    if ((rc = ensure_ws(c, std::max(ws_need(c, p), 52 * n + 64)))) return rc;
    tosca::vjp_rhs(g, m, p, ucont, ucat, theta, cs, rbar, scale, buoy_scale, ucont_bar, ucat_bar, theta_bar, cs_bar,
                   c->ws, c->stream);
    return launched(c, "rhs_vjp");
}

int tosca_rhs_vjp_mixed(tosca_ctx *c, const tosca_mesh *m, const tosca_params *p, const float *ucont,
                        const float *ucat, const float *theta, const double *cs, const double *rbar, double scale,
                        double buoy_scale, float *ucont_bar, float *ucat_bar, float *theta_bar, double *cs_bar) {
    int rc = check_ctx(c);
    if (rc) return rc;
    if ((rc = check_params(c, p))) return rc;
    if (p->les_active && p->les_model != TOSCA_LES_SMAGORINSKY && p->les_model != TOSCA_LES_DSM &&
        p->les_model != TOSCA_LES_DPASD)
        return set_err(c, TOSCA_ERR_UNSUPPORTED, "LES model %d: nu_t = Cs Delta^2 |S| models only", p->les_model);
    if (p->div_scheme != TOSCA_DIV_CENTRAL && p->div_scheme != TOSCA_DIV_QUICK)
        return set_err(c, TOSCA_ERR_UNSUPPORTED, "div_scheme = %d: only central (0) and QUICK (1) are supported", p->div_scheme);
    if (!(p->t_ref != 0.0)) return set_err(c, TOSCA_ERR_ARG, "t_ref must be non-zero");
    set_geo_per(c, p);
    const Geo &g = c->geo;
    const size_t n = (size_t)g.nloc(), v = 3 * n * sizeof(float), ng = 3 * (size_t)g.mx * g.my * g.mz * sizeof(double);
    if ((rc = check_mesh(c, m, true, true, true))) return rc;
    if ((rc = check_dev(c, ucont, v, "ucont")) || (rc = check_dev(c, ucat, v, "ucat")) ||
        (rc = check_dev(c, theta, n * sizeof(float), "theta")) || (rc = check_dev(c, cs, n * sizeof(double), "cs")) ||
        (rc = check_dev(c, rbar, ng, "rbar")) || (rc = check_dev(c, ucont_bar, v, "ucont_bar")) ||
        (rc = check_dev(c, ucat_bar, v, "ucat_bar")) || (rc = check_dev(c, theta_bar, n * sizeof(float), "theta_bar")) ||
        (rc = check_dev(c, cs_bar, n * sizeof(double), "cs_bar")))
        return rc;
    if ((rc = ensure_ws(c, std::max(ws_need(c, p), 52 * n + 64)))) return rc;
    tosca::vjp_rhs_mixed(g, m, p, ucont, ucat, cs, rbar, scale, buoy_scale, ucont_bar, ucat_bar, theta_bar, cs_bar,
                         c->ws, c->stream);
    return launched(c, "rhs_vjp_mixed");
}

}  // extern "C"

// ---------------------------------------------------------------------------------------------
// mesh check kernel: one block per k plane of the owned range, fixed-size per-plane counters
// ---------------------------------------------------------------------------------------------
namespace tosca {

// counters per plane: nvert != 0, meshTag != 0, aj < 1e-20 and a close-to-zero face area at interior
// cells, and interior cells whose csi, eta, zet differ from 0.5 a + 0.5 b of the face bases on
// their two faces (or every interior cell, when icsi / jeta / kzet are not given)
__global__ void k_mesh_check(Geo g, const double *nvert, const double *meshtag, const double *aj, const double *csi,
                             const double *eta, const double *zet, const double *icsi, const double *jeta,
                             const double *kzet, int *out) {
    const int k = blockIdx.x;
    int cnt[5] = {0, 0, 0, 0, 0};
    const int n = g.my * g.mx;
    for (int t = threadIdx.x; t < n; t += blockDim.x) {
        int j = t / g.mx, i = t % g.mx;
        int c = g.L(k, j, i);
        if (nvert[c] != 0.0) cnt[0]++;
        if (meshtag[c] != 0.0) cnt[1]++;
        bool in = i > 0 && i < g.mx - 1 && j > 0 && j < g.my - 1 && k > 0 && k < g.mz - 1;
        if (in) {
            if (!(aj[c] >= 1e-20)) cnt[2]++;
            V3 a = ld3(csi, c), b = ld3(eta, c), d = ld3(zet, c);
            double na = sqrt(a.x * a.x + a.y * a.y + a.z * a.z), nb = sqrt(b.x * b.x + b.y * b.y + b.z * b.z),
                   nd = sqrt(d.x * d.x + d.y * d.y + d.z * d.z);
            if (!(na >= 1e-20) || !(nb >= 1e-20) || !(nd >= 1e-20)) cnt[3]++;
            if (!icsi || !jeta || !kzet) {
                cnt[4]++;
            } else {
                auto avg_is = [](V3 cc, V3 f, V3 fl) {
                    return cc.x == 0.5 * f.x + 0.5 * fl.x && cc.y == 0.5 * f.y + 0.5 * fl.y && cc.z == 0.5 * f.z + 0.5 * fl.z;
                };
                if (!avg_is(a, ld3(icsi, c), ld3(icsi, g.L(k, j, i - 1))) || !avg_is(b, ld3(jeta, c), ld3(jeta, g.L(k, j - 1, i))) ||
                    !avg_is(d, ld3(kzet, c), ld3(kzet, g.L(k - 1, j, i))))
                    cnt[4]++;
            }
        }
    }
    __shared__ int red[5][256];
    for (int q = 0; q < 5; q++) red[q][threadIdx.x] = cnt[q];
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w)
            for (int q = 0; q < 5; q++) red[q][threadIdx.x] += red[q][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int q = 0; q < 5; q++) out[5 * k + q] = red[q][0];
}

void launch_mesh_check(const Geo &g, const tosca_mesh *m, int *out, hipStream_t s) {
    TOSCA_LAUNCH("k_mesh_check", k_mesh_check, dim3(g.mz), dim3(256), 0, s, g, m->nvert, m->meshtag, m->aj, m->csi, m->eta,
                 m->zet, m->icsi, m->jeta, m->kzet, out);
}

}  // namespace tosca
