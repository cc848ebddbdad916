/*
 * tosca_rhs.h -- C ABI of the MI355X (gfx950) momentum right-hand-side + LES library.
 *
 * The entry points replace, one for one, the reference solver routines that evaluate the
 * momentum-predictor right-hand side of TOSCA (sebastipa/TOSCA):
 *
 *   tosca_formu       <-> PetscErrorCode FormU(ueqn_*, Vec &Rhs, PetscReal scale, PetscInt formMode)
 *                         src/solvers/ueqn_solvers.h:10, body src/solvers/ueqn_solvers.c:6-1083
 *   tosca_update_cs   <-> PetscErrorCode UpdateCs(les_*)        src/include/les.h:72, src/les.c:165-2115
 *   tosca_update_nut  <-> PetscErrorCode UpdateNut(les_*)       src/include/les.h:75, src/les.c:2119-2578
 *                         (includes UpdateNutBCs, src/boundary.c:2630-2929)
 *   tosca_buoyancy    <-> PetscErrorCode Buoyancy(ueqn_*, PetscReal scale)
 *                         src/include/ueqn_sources.h:22, src/sources/ueqn_sources.c:2726-2880
 *
 * Storage contract (the reference's PETSc DMDA layout, src/mesh.c:136-203):
 *   - every array is [k][j][i] with i fastest; k <-> x (streamwise), j <-> z (vertical), i <-> y;
 *   - "local" arrays are the ghosted DMDA local vectors (ghost width 3): gxm*gym*gzm points,
 *     the pointer addresses element [gzs][gys][gxs], i.e. what VecGetArray returns;
 *   - "global" arrays (Rhs, bTheta, gCont) hold the owned range only: xm*ym*zm points,
 *     pointer at element [zs][ys][xs];
 *   - vectors are AoS {x,y,z} doubles (Cmpnts, src/include/base.h:81-84).
 * One GPU owns the whole domain: xs=ys=zs=0, xm=mx, ym=my, zm=mz, gxs=gys=gzs=-3,
 * gxm=mx+6, gym=my+6, gzm=mz+6.  Anything else returns TOSCA_ERR_LAYOUT.
 *
 * All pointers are DEVICE pointers (hipMalloc'd or torch CUDA tensors); the caller owns them.
 * Calls are asynchronous on the context's HIP stream (default: the null stream); results are
 * ready after tosca_ctx_synchronize() or any later work ordered on that stream.
 * Errors are returned as non-zero codes (never exit), with a message in tosca_last_error().
 */
#ifndef TOSCA_RHS_H
#define TOSCA_RHS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes ------------------------------------------------------------------------ */
#define TOSCA_OK               0
#define TOSCA_ERR_ARG          1   /* null / misaligned pointer (device arrays: 8-byte aligned,
                                        int tables 4-byte), bad enum value, bad form_mode        */
#define TOSCA_ERR_UNSUPPORTED  2   /* IBM, overset, structural LES, schemes other than central/QUICK */
#define TOSCA_ERR_LAYOUT       3   /* ghost width != 3 or not a whole-domain single-GPU layout   */
#define TOSCA_ERR_DEVICE       4   /* HIP runtime error or device unavailable                   */
#define TOSCA_ERR_MESH         5   /* tosca_mesh_check: non-zero nvert/meshTag, bad volume/area */

/* ---- layout: mirrors PETSc's DMDALocalInfo (mesh->info, src/include/mesh.h:40) ----------- */
typedef struct {
    int64_t xs, ys, zs, xm, ym, zm;        /* owned range                                 */
    int64_t gxs, gys, gzs, gxm, gym, gzm;  /* ghosted range                               */
    int64_t mx, my, mz;                    /* global DMDA sizes (cells + 2 boundary ghosts) */
} tosca_dmda_info;

/* ---- static mesh arrays (mesh_ fields, src/include/mesh.h:44-60), all local/ghosted ----- */
typedef struct {
    const double *csi, *eta, *zet;         /* cell-centred contravariant basis, AoS        */
    const double *icsi, *ieta, *izet;      /* i-face basis                                 */
    const double *jcsi, *jeta, *jzet;      /* j-face basis                                 */
    const double *kcsi, *keta, *kzet;      /* k-face basis                                 */
    const double *aj, *iaj, *jaj, *kaj;    /* inverse cell / face Jacobians (scalars)      */
    const double *nvert, *meshtag;         /* IBM / overset markers; only read by tosca_mesh_check */
} tosca_mesh;

/* periodic connectivity per axis (src/boundary.c:356-612): type 1 = i_periodic, type 2 = ii_periodic */
enum { TOSCA_PERIODIC_NONE = 0, TOSCA_PERIODIC_COPY = 1, TOSCA_PERIODIC_WRAP = 2 };

/* convective schemes accepted by tosca_formu (ueqn->centralDiv / ueqn->quickDiv, src/ueqn.c:81-110) */
enum { TOSCA_DIV_CENTRAL = 0, TOSCA_DIV_QUICK = 1 };

/* boundaryU patch kinds (mesh->boundaryU.iLeft ... kRight).  SLIP and WALLFUNCTION change FormU's
 * viscous flux (ueqn_solvers.c:336-356,548-568,758-768); NOSLIP, SLIP, WALLFUNCTION and
 * ZEROGRADIENT select the face resets of tosca_explicit_rhs_u (resetNoPenetrationFluxes,
 * inline.h:5262-5355; resetNonResolvedCellFaces, inline.h:1674-1838).  Every other velocity BC
 * inletFunction and oversetInterpolate patches are GENERIC.  FIXEDVALUE patches take the values
 * of tosca_update_cartesian_bcs; everywhere else FIXEDVALUE behaves as GENERIC.                   */
enum {
    TOSCA_BCU_GENERIC = 0, TOSCA_BCU_SLIP = 1, TOSCA_BCU_WALLFUNCTION = 2, TOSCA_BCU_PERIODIC = 3,
    TOSCA_BCU_NOSLIP = 4, TOSCA_BCU_ZEROGRADIENT = 5, TOSCA_BCU_FIXEDVALUE = 6
};

/* boundaryT patch kinds for tosca_update_temperature_bcs (src/boundary.c:2378-2555).  A
 * thetaWallFunction i or j patch fills its ghosts as zeroGradient (boundary.c:2380-2415) and is
 * passed as TOSCA_BCT_ZEROGRADIENT, except the jLeft patch with the surface-temperature model
 * (boundaryT.jLWF == -4) after the first iteration, whose ghost is 2 Ts - t(1) with Ts the
 * surface temperature interpolated to the current time (boundary.c:2392-2410): that patch is
 * passed as TOSCA_BCT_SURFACETEMP with Ts as its t_value. */
enum {
    TOSCA_BCT_ZEROGRADIENT = 0, TOSCA_BCT_FIXEDVALUE = 1, TOSCA_BCT_FIXEDGRADIENT = 2, TOSCA_BCT_PERIODIC = 3,
    TOSCA_BCT_SURFACETEMP = 4
};

/* boundaryNut patch kinds applied by UpdateNutBCs (src/boundary.c:2812-2917) */
enum { TOSCA_BCNUT_NONE = 0, TOSCA_BCNUT_ZEROGRADIENT = 1, TOSCA_BCNUT_FIXEDVALUE = 2, TOSCA_BCNUT_PERIODIC = 3 };

/* LES model numbering of the reference's LesModelNames (src/les.c:16-28) */
enum {
    TOSCA_LES_SMAGORINSKY = 0, TOSCA_LES_STABILITY = 1, TOSCA_LES_DSM = 2, TOSCA_LES_DLASI = 3,
    TOSCA_LES_DLASD = 4, TOSCA_LES_DPASD = 5, TOSCA_LES_AMD = 6, TOSCA_LES_VREMAN = 7,
    TOSCA_LES_BV = 8, TOSCA_LES_BAMD = 9
};

/* patch order used by bc_u / bc_nut / nut_value: iLeft, iRight, jLeft, jRight, kLeft, kRight */
typedef struct {
    int    periodic[3];      /* i, j, k: TOSCA_PERIODIC_*                                       */
    int    bc_u[6];          /* TOSCA_BCU_*                                                     */
    int    bc_nut[6];        /* TOSCA_BCNUT_*                                                   */
    double nut_value[6];     /* fixedValue nut per patch (boundaryNut.iLval ...)                */
    int    div_scheme;       /* TOSCA_DIV_*                                                     */
    int    les_active;       /* flags->isLesActive                                              */
    int    les_model;        /* TOSCA_LES_*; supported: SMAGORINSKY, DSM, DPASD                 */
    int    ibm_active;       /* flags->isIBMActive: must be 0                                   */
    int    overset_active;   /* flags->isOversetActive: must be 0                               */
    double nu;               /* constants->nu                                                   */
    double max_cs;           /* les->maxCs (default 0.5, src/les.c:49)                          */
    double t_ref;            /* abl->tRef or constants->tRef (ueqn_sources.c:2749-2751)         */
    /* wall-function shear stress (ueqn->iLWM / jLWM ->tauWall), used only where a patch is
     * TOSCA_BCU_WALLFUNCTION: three planes x|y|z, each [zm][ym] (i patches) or [zm][xm]
     * (j patches), device pointers; the i pair and the j pair each share one array, as in
     * the reference (ueqn_solvers.c:336-345, 548-556).                                      */
    const double *tau_wall_i;
    const double *tau_wall_j;
    /* velocity wall model of the iLeft, iRight, jLeft, jRight patches where they are
     * TOSCA_BCU_WALLFUNCTION (boundaryU.iLWF ... jRWF, boundary.c:4700-4980): -1 Cabot,
     * -3 Shumann, 0 unset.  Read only by tosca_update_cartesian_bcs, which fills the ghost
     * cells of Cabot i and j patches (boundary.c:968-1150) and of Shumann j patches
     * (:1153-1224), and leaves those of Shumann i patches as they are, as the reference does. */
    int wall_u[4];
    /* kRough of the Cabot wall models of those patches (iLWM->wmCabot->roughness ...); a value
     * above 1e-19 selects the rough-wall law utau_wf, otherwise the smooth-wall Cabot law    */
    double wall_rough[4];
} tosca_params;

/* ---- context ------------------------------------------------------------------------------ */
typedef struct tosca_ctx tosca_ctx;

int         tosca_ctx_create(tosca_ctx **ctx, const tosca_dmda_info *info, int device);
int         tosca_ctx_destroy(tosca_ctx *ctx);
int         tosca_ctx_set_stream(tosca_ctx *ctx, void *hip_stream);   /* hipStream_t, NULL = default */
int         tosca_ctx_synchronize(tosca_ctx *ctx);
const char *tosca_last_error(const tosca_ctx *ctx);
const char *tosca_version(void);

/* Optional per-kernel timing (instrumentation; no counterpart in the reference, whose solvers
 * print PetscTime wall timers).  When enabled, every kernel the context queues is bracketed by
 * two hipEvents on its stream; the readers synchronise and return, per kernel name, the total
 * event time in milliseconds and the number of launches since the last reset. */
int tosca_ctx_enable_timing(tosca_ctx *ctx, int on);
int tosca_ctx_timing_reset(tosca_ctx *ctx);
int tosca_ctx_timing_count(tosca_ctx *ctx);   /* number of kernel names, or -1 on error */
int tosca_ctx_timing_entry(tosca_ctx *ctx, int idx, char *name, int name_len, double *total_ms,
                           int64_t *launches);

/* Validates the static mesh once (replaces the checks of SetMeshMetrics, src/mesh.c:1306-1346,
 * and the IBM/overset flags): nvert and meshtag all zero, aj > 1e-20 and |csi|,|eta|,|zet| >= 1e-20
 * on the interior cells (NaN fails both).  Synchronous.  Returns TOSCA_ERR_MESH with a message on
 * failure.  REQUIRED once per mesh and context: the kernels never read nvert / meshTag, so every
 * entry point below that takes a mesh returns TOSCA_ERR_MESH unless a mesh with the same nvert,
 * meshtag, aj, csi, eta and zet pointers passed this check on the same context (the context keeps
 * the last 64 such meshes; a failed check removes the mesh).  The check does not see later writes
 * into those arrays: the mesh is static in the reference (SetMeshMetrics runs once).
 * When icsi, jeta and kzet are given it also records whether every interior cell's csi, eta, zet
 * equal 0.5 f + 0.5 f' of the face bases icsi, jeta, kzet on its two faces of that direction, bit
 * for bit, as SetMeshMetrics forms them (mesh.c:925-960, the face bases and the cells' dummy bases
 * being the same cross products).  For such a mesh FormU's face dot on FP64 metrics forms the cell
 * bases from the face bases it reads anyway instead of loading them (the same values); for any
 * other mesh, and on FP32 metric copies, it loads them. */
int tosca_mesh_check(tosca_ctx *ctx, const tosca_mesh *mesh);

/* FormU (src/solvers/ueqn_solvers.c:6-1083).
 *   ucont: lUcont (local AoS), ucat: lUcat (local AoS), nut: les->lNu_t (local scalar; read only
 *   when params->les_active), rhs: global AoS, ACCUMULATED (rhs += scale * ...).
 *   form_mode: 0 = convective + viscous, 1 = convective only, 2 = viscous only. */
int tosca_formu(tosca_ctx *ctx, const tosca_mesh *mesh, const tosca_params *params,
                const double *ucont, const double *ucat, const double *nut,
                double *rhs, double scale, int form_mode);

/* No counterpart in the reference.  tosca_formu for callers that cannot allocate device memory
 * (code under HIP graph capture, tight solver loops such as a Krylov or Newton iteration, hosts
 * that forbid hipMalloc): same arguments, same kernels (always the fused ones), same results,
 * but its scratch (boundary-face fluxes and tile seams) lives in a fixed device array of
 * TOSCA_NOALLOC_SCRATCH_DOUBLES doubles (256 MiB) that every GPU the library runs on holds until
 * the process ends, instead of the context's workspace.  It calls no allocating routine and does
 * not use the context's kernel timer.  A layout whose FormU needs more scratch than that is
 * rejected with TOSCA_ERR_LAYOUT (a 512x512x192 shard needs about 22 M doubles).  All calls on one
 * GPU share the array: calls that overlap on one GPU (several streams or host threads) race. */
#define TOSCA_NOALLOC_SCRATCH_DOUBLES ((size_t)32 * 1024 * 1024)
int tosca_formu_noalloc(tosca_ctx *ctx, const tosca_mesh *mesh, const tosca_params *params,
                        const double *ucont, const double *ucat, const double *nut,
                        double *rhs, double scale, int form_mode);

/* UpdateCs (src/les.c:165-2115): writes les->lCs (local scalar) for SMAGORINSKY (Cs = 0.0289
 * everywhere), DSM (test filter + Germano + 3x3x3 box average + clip) and DPASD (3- and 5-point
 * test filters + j-plane averages + clip).  Uses context workspace. */
int tosca_update_cs(tosca_ctx *ctx, const tosca_mesh *mesh, const tosca_params *params,
                    const double *ucat, double *cs);

/* UpdateNut (src/les.c:2119-2578, Smagorinsky/dynamic branch :2510-2535) followed by
 * UpdateNutBCs (src/boundary.c:2630-2929): writes lNu_t (local scalar, all 6 boundary patches)
 * and, when ksgs != NULL, lKsgs. */
int tosca_update_nut(tosca_ctx *ctx, const tosca_mesh *mesh, const tosca_params *params,
                     const double *ucat, const double *cs, double *nut, double *ksgs);

/* Buoyancy (src/sources/ueqn_sources.c:2726-2880): theta = teqn->lTmprt (local scalar);
 * btheta = ueqn->bTheta (global AoS) is zeroed and filled; gcont (global AoS, may be NULL)
 * receives gravity . face area vector on the faces the reference computes it on. */
int tosca_buoyancy(tosca_ctx *ctx, const tosca_mesh *mesh, const tosca_params *params,
                   const double *theta, double scale, double *btheta, double *gcont);

/* One full right-hand-side + LES evaluation, the unit of work of the benchmark:
 * UpdateCs -> UpdateNut -> Buoyancy(buoy_scale) -> FormU(rhs, scale, form_mode 0).
 * Outputs are exactly those of the four calls above (cs, nut, btheta written, rhs accumulated);
 * the kernels are fused where the data flow allows. */
int tosca_rhs_les_step(tosca_ctx *ctx, const tosca_mesh *mesh, const tosca_params *params,
                       const double *ucont, const double *ucat, const double *theta,
                       double *rhs, double scale, double *cs, double *nut,
                       double *btheta, double buoy_scale);

/* The same step with host-resident flow fields (the staged residency mode): ucont, ucat, theta
 * (local), rhs (global, accumulated), cs, nut (local) and btheta (global) are host arrays of the
 * same layouts; the mesh and params arrays stay on the device.  The context copies the inputs
 * into device buffers it owns (allocated on first use), runs tosca_rhs_les_step and copies the
 * outputs back before returning (synchronous).  Pinned host memory (hipHostMalloc) moves at the
 * link's full rate; pageable memory works but is slower.  Host sizes cannot be checked: they must
 * match the layout.  times_ms (may be NULL) receives {copy ms (both directions), step ms}. */
int tosca_rhs_les_step_host(tosca_ctx *ctx, const tosca_mesh *mesh, const tosca_params *params,
                            const double *ucont, const double *ucat, const double *theta, double *rhs,
                            double scale, double *cs, double *nut, double *btheta, double buoy_scale,
                            double *times_ms);

/* Mixed-precision form of tosca_rhs_les_step (no reference counterpart: a new precision mode of
 * the same unit of work).  The flow fields are FP32 (ucont, ucat, theta in; nut out); the metrics
 * and Jacobians stay FP64 (tosca_mesh), as do rhs (accumulated), cs and btheta.  Arithmetic:
 * velocity, strain and face-flux arithmetic in FP32; the test-filter sums in FP32 after a
 * per-tile reference velocity is taken out of both factors of the Leonard term (which leaves it
 * unchanged); the cell-face flux differences (fp), the face dot into rhs, the buoyancy source, the
 * Germano contraction, the box average, the plane means and the coefficient in FP64.  LES models:
 * SMAGORINSKY, DSM and DPASD (others: TOSCA_ERR_UNSUPPORTED). */
int tosca_rhs_les_step_mixed(tosca_ctx *ctx, const tosca_mesh *mesh, const tosca_params *params,
                             const float *ucont, const float *ucat, const float *theta,
                             double *rhs, double scale, double *cs, float *nut,
                             double *btheta, double buoy_scale);

/* FP32 copies of the metrics FormU reads, for tosca_rhs_les_step_mixed_m32 (no reference
 * counterpart).  Same layout as the FP64 fields of tosca_mesh (local AoS vectors, local scalars).
 * The mesh is static in the reference (SetMeshMetrics runs once), so the copies are made once. */
typedef struct {
    const float *csi, *eta, *zet;
    const float *icsi, *ieta, *izet;
    const float *jcsi, *jeta, *jzet;
    const float *kcsi, *keta, *kzet;
    const float *iaj, *jaj, *kaj;
} tosca_mesh_f32;

/* tosca_rhs_les_step_mixed with FormU reading the FP32 metric copies of mesh32 instead of the FP64
 * metrics (no reference counterpart).  FormU's flux arithmetic already runs in FP32 on FP32-rounded
 * face metrics; here the face dot into rhs and the buoyancy source also see FP32-rounded metrics
 * and Jacobians (arithmetic still FP64).  The dynamic Smagorinsky filter reads the FP32 csi, eta,
 * zet too (its strain already used them rounded to FP32; its Germano contraction now sees them
 * rounded); aj, the box average, DPASD and UpdateNut read the FP64 mesh.  FormU's metric bytes
 * halve (312 -> 156 B/cell). */
int tosca_rhs_les_step_mixed_m32(tosca_ctx *ctx, const tosca_mesh *mesh, const tosca_mesh_f32 *mesh32,
                                 const tosca_params *params, const float *ucont, const float *ucat,
                                 const float *theta, double *rhs, double scale, double *cs, float *nut,
                                 double *btheta, double buoy_scale);

/* Adjoint (vector-Jacobian product) of the right-hand side with respect to velocity, temperature
 * and Cs (no reference counterpart: the reference has no adjoint).  The function is
 *   R(ucont, ucat, theta, cs) = FormU(ucont, ucat, nu_t, scale, form_mode 0) + Buoyancy(theta, buoy_scale),
 *   nu_t = UpdateNut(ucat, cs)  (Cs Delta^2 |S| at interior cells, UpdateNutBCs ghosts),
 * i.e. what tosca_update_nut + tosca_formu (rhs from zero) + tosca_buoyancy produce, summed.
 * Every input is a full local array (ghost cells included); R and rbar are global AoS arrays.
 * Given rbar, writes (overwrites) ucont_bar, ucat_bar (local AoS), theta_bar, cs_bar (local
 * scalars) = J^T rbar.  Forward intermediates are recomputed, not stored.  The sums use atomic
 * adds, so repeated calls agree to round-off, not bit for bit.  Where |S| = 0 the subgradient 0
 * is used; QUICK's upwind switch is differentiated with the branch held. */
int tosca_rhs_vjp(tosca_ctx *ctx, const tosca_mesh *mesh, const tosca_params *params,
                  const double *ucont, const double *ucat, const double *theta, const double *cs,
                  const double *rbar, double scale, double buoy_scale,
                  double *ucont_bar, double *ucat_bar, double *theta_bar, double *cs_bar);

/* Mixed-precision form of tosca_rhs_vjp (no reference counterpart): the adjoint of the function
 * tosca_rhs_les_step_mixed evaluates, for its FP32 flow fields.  ucont, ucat, theta (local, FP32)
 * in; cs (local), rbar (global) FP64; ucont_bar, ucat_bar, theta_bar (local, FP32) and cs_bar
 * (local, FP64) overwritten.  The forward intermediates (nu_t, |S|, the face gradients and
 * interpolated velocities) are recomputed in FP32 as the mixed step computes them; the cotangent
 * arithmetic is FP64; the cotangent packets between the kernels are stored in FP32, which halves
 * their HBM traffic.  Same function, error codes and conventions as tosca_rhs_vjp. */
int tosca_rhs_vjp_mixed(tosca_ctx *ctx, const tosca_mesh *mesh, const tosca_params *params,
                        const float *ucont, const float *ucat, const float *theta, const double *cs,
                        const double *rbar, double scale, double buoy_scale,
                        float *ucont_bar, float *ucat_bar, float *theta_bar, double *cs_bar);

/* ExplicitRhsU (src/solvers/ueqn_solvers.h:24; body src/solvers/ueqn_solvers.c:1276-1407), the
 * immediate caller of FormU, with the
 * terms this library covers, in the reference's order:
 *   resetNoPenetrationFluxes (src/include/inline.h:5262-5355) on ucont: the normal flux at noSlip,
 *     velocityWallFunction and slip patches set to 0, at zeroGradient patches clipped to outflow;
 *   resetFacePeriodicFluxesVector (inline.h:5357-5451) on ucont;
 *   contravariantToCartesian (src/ueqn.c:528-635): ucat at the interior cells from ucont;
 *   resetCellPeriodicFluxes (inline.h:5453-5606) on ucat;
 *   rhs = 0; rhs -= dp; rhs += 3/2 btheta - 1/2 btheta_o (ab2 != 0, clock->it > itStart) or
 *     rhs += btheta (ab2 == 0); rhs += farm; FormU(rhs, 1.0, form_mode) -- VecAXPY order;
 *   resetNonResolvedCellFaces (inline.h:1674-1838) on rhs, with ucont_o.
 * ucont: lUcont (local AoS, in/out); ucat: lUcat (local AoS, out: interior cells and periodic
 * boundary cells/ghosts; other ghost cells keep their values); nut: lNu_t (read when les_active);
 * dp, btheta, btheta_o, farm, ucont_o, rhs: global AoS (dP, bTheta, bTheta_o, sourceFarmCont,
 * Ucont_o, Rhs).  btheta == NULL: temperature equation off; farm == NULL: no wind farm.
 * The reference's other source terms (Coriolis, canopy, driving controller, damping layers,
 * pressure-gradient forcings, hyperviscosity, the pTilde buoyancy form) are not part of this
 * path; a caller that uses them adds them to rhs before resetNonResolvedCellFaces would. */
int tosca_explicit_rhs_u(tosca_ctx *ctx, const tosca_mesh *mesh, const tosca_params *params,
                         double *ucont, double *ucat, const double *nut, const double *dp,
                         const double *btheta, const double *btheta_o, int ab2, const double *farm,
                         const double *ucont_o, double *rhs, int form_mode);

/* Settings of the IMEX momentum solve, kspIMEX (src/ueqn.c:217-282): GMRES with restart
 * `restart` (ueqn->gmresRestart, 30), no preconditioner, nonzero initial guess, converged when the
 * GMRES residual estimate is <= max(rtol * ||r0||, atol) (relTolU 1e-30, absTolU 1e-5), at most
 * max_it iterations (1000). */
typedef struct {
    int    restart;
    int    max_it;
    double rtol;
    double atol;
} tosca_krylov_params;

/* Outcome of a solve: total GMRES iterations, the last residual estimate, and reason, a PETSc
 * KSPConvergedReason value: 2 KSP_CONVERGED_RTOL, 3 KSP_CONVERGED_ATOL, -2 KSP_DIVERGED_NULL (zero
 * Givens rotation), -3 KSP_DIVERGED_ITS, -4 KSP_DIVERGED_DTOL (residual >= 1e30 ||b||),
 * -5 KSP_DIVERGED_BREAKDOWN, -9 KSP_DIVERGED_NANORINF.  The tolerance is max(rtol ||b||, atol), with
 * ||r0|| in place of ||b|| when b = 0 (KSPConvergedDefault with a nonzero initial guess). */
typedef struct {
    int    iterations;
    double residual;
    int    reason;
} tosca_krylov_result;

/* Replaces KSPSolve(ueqn->kspIMEX, b, x) of the IMEX schemes (ueqn_solvers.c:1564, 1641, 1716)
 * with IMEXMatVec as the operator: restarted GMRES on the device (PETSc 3.15's KSPGMRES algorithm:
 * classical Gram-Schmidt, Givens rotations, no preconditioner).  b, x: global AoS, x the initial
 * guess on entry and the solution on return; ucont, ucat: the operator's local work arrays;
 * coef = dt scale as for tosca_imex_matvec.  The Krylov basis, (restart + 2) global vectors, lives
 * in the context and is allocated on first use. */
int tosca_imex_solve(tosca_ctx *ctx, const tosca_mesh *mesh, const tosca_params *params,
                     const double *b, double *x, double *ucont, double *ucat, const double *nut,
                     const double *ucont_o, double coef, const tosca_krylov_params *kp,
                     tosca_krylov_result *kr);

/* Replaces UeqnRK3CN_W (scheme 0, "RK3WCN", src/solvers/ueqn_solvers.c:1508-1583) and
 * UeqnRK3CN_SO (scheme 1, "RK3SOCN", :1586-1661): three IMEX Runge-Kutta stages, each with
 * ExplicitRhsU(formMode 1) into rhs, the viscous FormU(formMode 2) with its non-resolved faces
 * reset, the stage's right-hand side bU and explicit-Euler initial guess assembled with the
 * reference's VecAXPY sequence, and tosca_imex_solve with coef = dt gamma_i.  Arguments as
 * tosca_ueqn_rk4; kr reports the last stage's solve.  Work vectors live in the context. */
int tosca_ueqn_rk3cn(tosca_ctx *ctx, const tosca_mesh *mesh, const tosca_params *params, int scheme,
                     double *ucont, double *ucat, const double *nut, const double *dp,
                     const double *btheta, const double *btheta_o, int ab2, const double *farm,
                     const double *ucont_o, double dt, double *rhs, const tosca_krylov_params *kp,
                     tosca_krylov_result *kr);

/* Replaces UeqnABCN (src/solvers/ueqn_solvers.c:1665-1737): Adams-Bashforth 2 convection and
 * Crank-Nicolson diffusion in one IMEX solve.  ab2 is clock->it > clock->itStart: it selects the
 * AB2 bTheta terms of ExplicitRhsU, the AB2 convection with conv_o and the Crank-Nicolson half
 * (coef = dt / 2); on the first step (ab2 == 0) the viscous term is fully implicit (coef = dt) and
 * the reference adds Rhs to bU = Rhs a second time (:1694-1695), which is reproduced.
 * conv_o: RhsConv_o (global AoS), read and replaced by this step's convective term.  Other
 * arguments as tosca_ueqn_rk3cn. */
int tosca_ueqn_abcn(tosca_ctx *ctx, const tosca_mesh *mesh, const tosca_params *params,
                    double *ucont, double *ucat, const double *nut, const double *dp,
                    const double *btheta, const double *btheta_o, int ab2, const double *farm,
                    const double *ucont_o, double dt, double *rhs, double *conv_o,
                    const tosca_krylov_params *kp, tosca_krylov_result *kr);

/* The constant right-hand side of the Crank-Nicolson momentum solve, built by UeqnSNES before
 * SNESSolve (src/solvers/ueqn_solvers.c:1099-1156, the terms of this path): bU = -dp, + bTheta
 * terms (3/2 btheta - 1/2 btheta_o when ab2, btheta otherwise; btheta NULL: off), + 0.5 rhs_o
 * when ab2 (the explicit Crank-Nicolson half), + farm (NULL: off).  All global AoS; ab2 is
 * clock->it > clock->itStart. */
int tosca_cn_rhs(tosca_ctx *ctx, const tosca_params *params, const double *dp, const double *btheta,
                 const double *btheta_o, int ab2, const double *rhs_o, const double *farm, double *bU);

/* Replaces SNESFuncEval (src/solvers/ueqn_solvers.c:1185-1272), the residual of the Crank-Nicolson
 * momentum solve that UeqnSNES hands to SNESSolve, and with it every function evaluation of the
 * matrix-free Jacobian (MatMFFD):
 *   F(u) = dt (bU + s FormU(u, formMode 0)), non-resolved faces reset, - u + ucont_o,
 * with s = 1 on the first step (ab2 == 0, backward Euler) and 1/2 after (Crank-Nicolson), FormU
 * evaluated after the ExplicitRhsU resets and contravariantToCartesian of u.  u, bU, ucont_o, F:
 * global AoS; ucont, ucat: the local work arrays lUcont, lUcat (overwritten, as in the
 * reference); nut: lNu_t.  A PETSc caller registers it with SNESSetFunction on device Vecs. */
int tosca_snes_func_eval(tosca_ctx *ctx, const tosca_mesh *mesh, const tosca_params *params,
                         const double *u, double *ucont, double *ucat, const double *nut,
                         const double *bU, const double *ucont_o, double dt, int ab2, double *F);

/* Replaces GradP (src/peqn.c:3437-3773): the contravariant pressure-gradient fluxes
 * dP^q = (g^q1 dp/dxi + g^q2 dp/deta + g^q3 dp/dzeta) aj_q at the faces 1..m-2 of the interior
 * cells, with the reference's one-sided transverse differences on the corner face rows and its
 * type-2 periodic neighbour (ghost m+1) on face m-2; the right boundary face of a non-periodic
 * axis and every other point of dp are 0.  p: local pressure lP (ghosts as the caller's
 * boundary conditions left them); dp: global AoS (ueqn->dP), overwritten. */
int tosca_grad_p(tosca_ctx *ctx, const tosca_mesh *mesh, const tosca_params *params, const double *p, double *dp);

/* Replaces UeqnFE (src/solvers/ueqn_solvers.c:1411-1441), the forward-Euler momentum predictor:
 * rhs = ExplicitRhsU(form_mode) as tosca_explicit_rhs_u; rhs = dt rhs; rhs += ucont_o;
 * ucont (owned range) = rhs, then the periodic ghost wrap of DMGlobalToLocal.  On return rhs
 * holds the new Ucont, as the reference's Rhs does.  Arguments as tosca_explicit_rhs_u. */
int tosca_ueqn_fe(tosca_ctx *ctx, const tosca_mesh *mesh, const tosca_params *params,
                  double *ucont, double *ucat, const double *nut, const double *dp,
                  const double *btheta, const double *btheta_o, int ab2, const double *farm,
                  const double *ucont_o, double dt, double *rhs, int form_mode);

/* Replaces UeqnRK4 (src/solvers/ueqn_solvers.c:1446-1504), the classical four-stage Runge-Kutta
 * predictor with b = (1/6, 1/3, 1/3, 1/6) and a = (0, 1/2, 1/2, 1): utmp = ucont_o; per stage i,
 * for i > 0 ucont = a_i dt rhs + ucont_o (VecWAXPY, owned range + ghost wrap), then
 * rhs = ExplicitRhsU and utmp += b_i dt rhs; finally ucont = utmp.  The sources (dp, btheta,
 * farm) and nut are held fixed over the stages, as in the reference.  utmp: global AoS work
 * vector (Ueqn->Utmp), not aliasing rhs or ucont_o. */
int tosca_ueqn_rk4(tosca_ctx *ctx, const tosca_mesh *mesh, const tosca_params *params,
                   double *ucont, double *ucat, const double *nut, const double *dp,
                   const double *btheta, const double *btheta_o, int ab2, const double *farm,
                   const double *ucont_o, double dt, double *rhs, double *utmp, int form_mode);

/* The explicit Wray Runge-Kutta momentum predictor: the three-stage low-storage scheme of
 * UeqnRK3CN_W (src/solvers/ueqn_solvers.c:1508-1583) with its coefficients beta = (8/15, 5/12, 3/4),
 * delta = (0, 17/60, 5/12) (:1518-1520) and the whole right-hand side explicit (the viscous term
 * too: beta_i - delta_i = 2 gamma_i).  Stage i: rhs = ExplicitRhsU(form_mode) as
 * tosca_explicit_rhs_u; Ucont = (Ucont + dt beta_i rhs) + (-1.0 dt delta_i) rhs_o, the reference's
 * VecCopy / VecAXPY order (:1533-1541); Ucont into the owned range of ucont with the periodic ghost
 * wrap; rhs_o = rhs (:1563-1569).  n_stages stages run back to back and every third one begins a
 * time step, where ucont_o = Ucont (as the time loop sets Ucont_o before the predictor) and rhs_o = 0
 * (:1522).  The sources (dp, btheta, farm) and nut are held fixed.  ucont, ucat: lUcont, lUcat (in
 * place); ucont_o, rhs, rhs_o: global work vectors (on return: the last step's Ucont_o, the last
 * stage's Rhs twice).  Other arguments as tosca_explicit_rhs_u. */
int tosca_ueqn_rk3w(tosca_ctx *ctx, const tosca_mesh *mesh, const tosca_params *params,
                    double *ucont, double *ucat, const double *nut, const double *dp,
                    const double *btheta, const double *btheta_o, int ab2, const double *farm,
                    double dt, int n_stages, double *ucont_o, double *rhs, double *rhs_o, int form_mode);

/* Checkpointed adjoint of tosca_ueqn_rk3w (form_mode 0) over n_stages stages (no reference
 * counterpart: the reference has no adjoint).  The function is the final Ucont (global) of the
 * predictor as a function of the initial Ucont (the owned range of ucont) and of the local nu_t
 * array; the sources, ucat's non-periodic boundary cells and the ghosts outside periodic axes are
 * constants.  Runs the forward sweep (leaving ucont, ucat, ucont_o, rhs, rhs_o as tosca_ueqn_rk3w
 * leaves them), storing every stage's starting state in the context (n_stages + 1 global vectors),
 * then the backward sweep: per stage, in reverse order, the transposes of the stage update,
 * resetNonResolvedCellFaces, FormU (tosca_rhs_vjp's kernels with nu_t an input), the periodic
 * resets of ucat, contravariantToCartesian and the face resets.
 *   ucont_bar: global, in: the cotangent of the final Ucont; out: that of the initial Ucont.
 *   nut_bar: local, out: the cotangent of nu_t (read only when params->les_active).
 * The zeroGradient inflow clip and QUICK's upwind switch are differentiated with the branch held. */
int tosca_ueqn_rk3w_adjoint(tosca_ctx *ctx, const tosca_mesh *mesh, const tosca_params *params,
                            double *ucont, double *ucat, const double *nut, const double *dp,
                            const double *btheta, const double *btheta_o, int ab2, const double *farm,
                            double dt, int n_stages, double *ucont_o, double *rhs, double *rhs_o,
                            double *ucont_bar, double *nut_bar);

/* Replaces IMEXMatVec (src/solvers/ueqn_solvers.c:1739-1801), the matrix-free operator of the
 * implicit (Crank-Nicolson / IMEX Runge-Kutta) momentum solve: av = v - coef Visc(v), where
 * coef = dt scale (scale 1/2 for the CN steps of ABCN, 1 for RK3WCN / RK3SOCN, chosen by the
 * caller) and Visc is FormU(formMode 2) of v after the ExplicitRhsU resets and
 * contravariantToCartesian, followed by resetNonResolvedCellFaces with ucont_o.
 * v, ucont_o, av: global AoS; ucont, ucat: local work arrays (lUcont, lUcat) left as the
 * reference leaves them. */
int tosca_imex_matvec(tosca_ctx *ctx, const tosca_mesh *mesh, const tosca_params *params,
                      const double *v, double *ucont, double *ucat, const double *nut,
                      const double *ucont_o, double coef, double *av);

/* ---- temperature equation (the sibling of FormU) ----------------------------------------------- */

/* Inputs of FormT beyond tosca_params (teqn_ / constants_ members). */
typedef struct {
    double pr;              /* constants->Pr                                                         */
    int    wall_t[4];       /* thetaWallFunction on iLeft, iRight, jLeft, jRight (boundaryT)         */
    /* wall heat flux of the thermal wall models, read where wall_t is set and LES is active:
     * three planes x|y|z, each [zm][ym] (i patches: teqn->iRWM->qWall, for both i patches as the
     * reference reads it, teqn_solvers.c:212-229) or [zm][xm] (j patches: jLWM->qWall, :337-354);
     * device pointers, may be NULL where no patch needs them                                     */
    const double *q_wall_i;
    const double *q_wall_j;
} tosca_teqn_params;

/* FormT (src/solvers/teqn_solvers.c:6-592): rhs (global scalar) is OVERWRITTEN at the interior
 * cells with scale * (divergence budget [form_mode != 2] + diffusion budget [form_mode != 1]) * aj.
 * Face fluxes: convection -ucont * weno3 of theta on the four-cell stencil (inline.h:5147-5179,
 * getFace2Cell4Stencil*); diffusion (g . grad theta) * face Jacobian * kappa_eff with
 *   LES off:  nu / Pr;
 *   LES on (the molecular part is dropped, as in the reference):
 *     AMD, dynamicSmagorinsky, dynamicLASI, dynamicLASD, dynamicPASD: 0.5 (kt_L + kt_R) (les->lk_t);
 *     smagorinsky, vreman: nu_t / Prt with the stability-corrected Prt (teqn_solvers.c:190-200);
 *     thetaWallFunction faces: the wall heat flux, kappa 0.
 * Then the periodic reset of the face fluxes (resetFacePeriodicFluxesVector).  kt is read for the
 * dynamic models only (may be NULL otherwise); nut whenever LES is active.  Stability-based and
 * structural models (bardinaVreman / bardinaAMD) return TOSCA_ERR_UNSUPPORTED. */
int tosca_formt(tosca_ctx *ctx, const tosca_mesh *mesh, const tosca_params *params, const tosca_teqn_params *tparams,
                const double *ucont, const double *theta, const double *nut, const double *kt, double *rhs,
                double scale, int form_mode);

/* ---- boundary-condition ghost fills (the step before the path) ------------------------------ */

/* UpdateContravariantBCs (src/boundary.c:614-822) on lUcont (local AoS, in place).  On every
 * boundary patch's faces (loop range k < mz-1, j < my-1, i < mx-1): patches other than noSlip /
 * slip / velocityWallFunction (PatchType 0, boundary.c:318-348) take the flux of the face-averaged
 * ucat, no-penetration patches (PatchType 1) get zero flux, zeroGradient patches then lose an
 * inflowing flux, and the left face of a periodic axis takes the right face's flux from before the
 * call (the right face itself is re-interpolated, as in the reference).  Then the ghost scatter. */
int tosca_update_contravariant_bcs(tosca_ctx *ctx, const tosca_mesh *mesh, const tosca_params *params,
                                   double *ucont, const double *ucat);

/* The kLeft inletFunction profile of tosca_update_cartesian_bcs (mesh->inletF.kLeft, read in
 * src/inflow.c:40-120 and 422-870; applied at boundary.c:1228-1600).  h = cent.z - grnd_level.
 *   TOSCA_INLET_POWERLAW (typeU 1): (h / href)^0.107027 uref (1 + noise / 10000 uprime_rms).  The
 *     reference draws noise = rand() % 20000 - 10000 per ghost after srand(time(NULL)), so its
 *     values cannot be reproduced; the caller passes the draws it wants, one per (j, i).
 *   TOSCA_INLET_LOGLAW (2), TOSCA_INLET_SINUSOIDAL (6): the analytic profiles.
 *   TOSCA_INLET_NIEUWSTADT (5): NieuwstadtInflowEvaluate (src/inflow.c:2732-2811), the Ekman-layer
 *     profile with veer: (U, V, 0) from the complex velocity W(h) = Wg - dW(h) of a boundary layer
 *     capped at h_inv, rotated so that the profile at h_ref points along udir.
 *   TOSCA_INLET_MAPPED (3): the inflow-database plane repeated prds1 x prds2 times: ghost (j, i)
 *     takes plane[jif][iif], jif = j mod n1 in 1..n1 and iif = i mod n2 in 1..n2, while
 *     j <= n1 prds1 and i <= n2 prds2; beyond that the last plane row / column (or, with merge1,
 *     avg_top_u[9]).  With merge1 the value below the database top is blended into the height
 *     profile avg_top_u at avg_top_coords (two-point interpolation, findInterpolationWeightsWith-
 *     Extrap, inline.h:5873-5966) by the tanh weights of scaleHyperTang{Bot,Top} (inline.h:
 *     5971-5999) with H = avg_top_length and delta = avg_top_delta.
 *   TOSCA_INLET_INTERP (4): the database plane interpolated onto this mesh with the per-cell
 *     stencils and weights of SetInflowWeights (src/inflow.c:941-1870): 4 points for every
 *     component, or with spline the y (z) component from the 6-point stencils cells_2/weights_2
 *     (cells_1/weights_1); with shift2 the values at i_left = y_ids[i] - 1 and i_right = y_ids[i]
 *     mixed as (1 - y_weights[i]) left + y_weights[i] right (boundary.c:885-935 computes y_ids and
 *     y_weights on the host each call); then the merge1 blend of type 3.
 * plane is ifPtr->ucat_plane after readInflowU (the caller reads and time-interpolates the
 * database files and copies the plane to the device). */
enum {
    TOSCA_INLET_POWERLAW = 1, TOSCA_INLET_LOGLAW = 2, TOSCA_INLET_MAPPED = 3, TOSCA_INLET_INTERP = 4,
    TOSCA_INLET_NIEUWSTADT = 5, TOSCA_INLET_SINUSOIDAL = 6
};
typedef struct {
    int type;                        /* typeU: one of TOSCA_INLET_*                                 */
    const double *cent;              /* mesh->lCent: cell centres, local AoS, device                */
    double grnd_level;               /* mesh->grndLevel (log law: h = cent.z - grnd_level)          */
    double ymin, ly;                 /* mesh->bounds.ymin, mesh->bounds.Ly (sinusoidal)             */
    double u_tau, roughness, h_inv;  /* uTau, roughness, hInv (log law)                             */
    double amplitude, periods;       /* sinusoidal amplitude and number of periods across Ly        */
    double udir[3];                  /* Udir, the unit flow direction                               */
    double uref[3];                  /* Uref (sinusoidal: its magnitude scales the profile)         */
    double fc;                       /* Nieuwstadt: Coriolis parameter fc = 2 Omega sin(latitude)   */
    double h_ref;                    /* Nieuwstadt: hReference, the height whose direction is udir  */
    /* power law (typeU 1); its velocity at href is uref                                          */
    double href, uprime_rms;         /* Href, uPrimeRMS                                             */
    const int *noise;                /* [my][mx] the draws in [-10000, 9999] at the kLeft ghosts     */
    /* mapped inflow databases (typeU 3 and 4)                                                   */
    const double *plane;             /* ucat_plane [n1 + 2][n2 + 2][3], device                      */
    int n1, n2;                      /* database plane size without its side ghosts                 */
    int prds1, prds2;                /* n1Periods, n2Periods (typeU 3)                              */
    int merge1;                      /* n1Merge: blend into avg_top_u near the database top         */
    double avg_top_length;           /* avgTopLength                                                */
    double avg_top_delta;            /* avgTopDelta                                                 */
    double avg_top_coords[10];       /* avgTopPointCoords (ascending heights)                       */
    double avg_top_u[10][3];         /* uBarAvgTopX                                                 */
    /* typeU 4: stencils into plane as (j, i) pairs, [my][mx][4][2] and [my][mx][6][2], device   */
    int spline;                      /* interpMethod == "spline"                                    */
    int shift2;                      /* n2Shift                                                     */
    const int *cells;                /* closestCells                                                */
    const double *weights;           /* inflowWeights [my][mx][4]                                   */
    const int *cells_1, *cells_2;    /* closestCells_1, closestCells_2 (spline only)                */
    const double *weights_1;         /* inflowWeights_1 [my][mx][6] (spline only)                   */
    const double *weights_2;         /* inflowWeights_2 [my][mx][6] (spline only)                   */
    const int *y_ids;                /* yIDs [mx] (shift2 only)                                     */
    const double *y_weights;         /* yWeights [mx] (shift2 only)                                 */
} tosca_inlet;

/* UpdateCartesianBCs (src/boundary.c:824-2050) on lUcat (local AoS, in place): the ghost cells of
 * every patch next to the interior cells (edges and corners untouched) from
 *   fixedValue   the patch value (u_fixed: host array of 6 x 3 doubles, iLval ... kRval; may be
 *                NULL when no patch is TOSCA_BCU_FIXEDVALUE),
 *   noSlip       -u of the boundary cell,
 *   slip         the boundary cell's velocity mirrored in the patch (calculateNormal, inline.h:1231),
 *   zeroGradient 2 u(1) - u(2) (linear extrapolation),
 *   periodic     the opposite interior cell (type 1) or ghost (type 2),
 *   velocityWallFunction with the Shumann model (params->wall_u = -3) on a j patch
 *                (boundary.c:1153-1224): 2 Ub_par - Uc_par - Ub_nor, the parts parallel and
 *                normal to the patch (calculateNormal's j vector) of the boundary cell b and the
 *                next interior cell c; on an i patch its ghosts are left as they are,
 *   velocityWallFunction with the Cabot model (wall_u = -1) on an i or j patch (boundary.c:
 *                968-1150): u* of the boundary cell from the parallel speed |Ub_par| at the
 *                wall distance sb = 0.5 / aj / |area| (uTauCabot, the Newton solve of the
 *                integrated near-wall eddy-viscosity law, wallfunctions.c:280-543, or utau_wf
 *                for rough walls, :568-585), written to ustar at the boundary cell; the ghost
 *                gets the parallel speed max(|Ub_par| - 2 sb u*^2 / nu, -|Ub_par|) along Ub_par
 *                minus Ub_nor (0 and u* = 0 where |Ub_par| < 1e-20).  The right patch of an
 *                axis whose left patch is Cabot too takes the left patch's roughness, as in the
 *                reference (boundary.c:978-979, 1072-1073),
 *   inletFunction on kLeft (bc_u[4] = TOSCA_BCU_GENERIC with inlet != NULL, boundary.c:1228-1600):
 *                the profile of tosca_inlet: the power law, the log law max(uTau/0.4 ln(h/roughness),
 *                1e-5) Udir below hInv and its value at hInv above, h = cent.z - grnd_level, the
 *                mapped and interpolated inflow-database planes, the Nieuwstadt profile, or the
 *                sinusoidal profile (1 + amplitude cos(2 pi periods (cent.y - ymin)/Ly)) |Uref| Udir,
 * then the ghost scatter.  ustar: local scalar array (lUstar), device; may be NULL when no patch
 * is Cabot.  inlet: may be NULL when kLeft is not an inlet function.  Inlet functions on other
 * patches and wall-function k patches return TOSCA_ERR_UNSUPPORTED.  The first call with a
 * smooth Cabot patch builds the context's copy of the reference's integration table (5e6 + 1
 * doubles). */
int tosca_update_cartesian_bcs(tosca_ctx *ctx, const tosca_mesh *mesh, const tosca_params *params, double *ucat,
                               const double *u_fixed, double *ustar, const tosca_inlet *inlet);

/* UpdateTemperatureBCs (src/boundary.c:2052-2628) on lTmprt (local scalar, in place): ghosts of
 * zeroGradient (t(1)), fixedValue (value), fixedGradient ((1/aj)/|area| value + t(1), the area of
 * the boundary cell's own basis vector) and periodic patches, then the ghost scatter.  bc_t and
 * t_value: host arrays of 6 (TOSCA_BCT_* kinds, values) in patch order.  The jLeft patch may be
 * TOSCA_BCT_SURFACETEMP (2 Ts - t(1), Ts = t_value[2]; boundary.c:2392-2410); thetaWallFunction
 * i and j patches are otherwise zeroGradient.  Inlet functions are not covered. */
int tosca_update_temperature_bcs(tosca_ctx *ctx, const tosca_mesh *mesh, const tosca_params *params,
                                 const int *bc_t, const double *t_value, double *theta);

/* ---- wind-farm body force (the sourceFarmCont term of ExplicitRhsU) -------------------------- */

/* Uniform actuator-disk rotors (the uniformADM turbine model; windTurbine members, src/include/
 * turbines.h).  Device arrays of float64:
 *   center [n_rotors][3]            rotor centres (wt->rotCenter)
 *   sphere [n_rotors]               radius of the controlled-cell sphere, rTip + eps * nsigma
 *                                   (turbines_parallel.c:57-80)
 *   eps    [n_rotors]               Gaussian projection width (wt->eps, "epsilon")
 *   points [n_rotors][n_points][3]  disk points (wt->uadm.points, meshUADM)
 *   force  [n_rotors][n_points][3]  body force per unit density at each point (wt->uadm.B)
 * nsigma is the projection cut-off in widths, sqrt(log(1/0.001)) in the reference
 * (wt->prjNSigma, turbines_io.c:3469). */
typedef struct {
    int n_rotors;
    int n_points;
    double nsigma;
    const double *center;
    const double *sphere;
    const double *eps;
    const double *points;
    const double *force;
} tosca_uadm;

/* projectBladeForce, uniformADM branch (src/turbines/turbines_force.c:717-932, 1179-1185): scat
 * (farm->lsourceFarmCat, local AoS) is set to zero and, at every interior cell within a rotor's
 * sphere (|cent - center| <= sphere), receives exp(-(r/eps)^2) / (eps^3 pi^1.5) * force for each
 * of that rotor's points at a distance r < eps * nsigma, rotors and points in order; then the
 * ghost scatter.  cent: mesh->lCent (local AoS).  The other turbine models (ADM, ALM, AFM) and the
 * tower and nacelle forces are not covered. */
int tosca_project_uadm(tosca_ctx *ctx, const tosca_params *params, const double *cent, const tosca_uadm *farm,
                       double *scat);

/* bodyForceCartesian2Contravariant (src/turbines/turbines_force.c:1741-1835): scont
 * (farm->sourceFarmCont, global AoS) = 0, then at every interior cell the body force of scat
 * averaged to the cell's +i, +j, +k faces and dotted with icsi, jeta, kzet. */
int tosca_body_force_to_cont(tosca_ctx *ctx, const tosca_mesh *mesh, const tosca_params *params, const double *scat,
                             double *scont);

/* tosca_rhs_les_step with the wind-farm source: rhs += farm (VecAXPY(Rhs, 1, sourceFarmCont),
 * src/solvers/ueqn_solvers.c:1340-1344, over every owned point) before the FormU accumulation,
 * in one pass with it.  farm: global AoS (tosca_body_force_to_cont's scont). */
int tosca_rhs_les_step_farm(tosca_ctx *ctx, const tosca_mesh *mesh, const tosca_params *params,
                            const double *ucont, const double *ucat, const double *theta, const double *farm,
                            double *rhs, double scale, double *cs, double *nut, double *btheta, double buoy_scale);

/* tosca_rhs_les_step_mixed (mesh32 == NULL) or tosca_rhs_les_step_mixed_m32 (mesh32 != NULL) with
 * the wind-farm source (farm: float64 global AoS, added in FP64 as in tosca_rhs_les_step_farm). */
int tosca_rhs_les_step_mixed_farm(tosca_ctx *ctx, const tosca_mesh *mesh, const tosca_mesh_f32 *mesh32,
                                  const tosca_params *params, const float *ucont, const float *ucat,
                                  const float *theta, const double *farm, double *rhs, double scale, double *cs,
                                  float *nut, double *btheta, double buoy_scale);

#ifdef __cplusplus
}
#endif

#endif /* TOSCA_RHS_H */
