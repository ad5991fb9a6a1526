/*
 * vlm_c.h -- C ABI of the MI355X free-wake hot path (libvlm.so).
 *
 * Drop-in boundary for the unsteady vortex-lattice hot path of
 * cibinjoseph/VOLCANOR.  The reference has no FFI: the path sits behind
 * Fortran type-bound procedures.  Each entry point below names the
 * reference routine it replaces (file:line in cibinjoseph/VOLCANOR);
 * the Fortran host binds them with ISO_C_BINDING interfaces
 * (volcanor_amd/fortran/vlm_c_binding.f90).
 *
 * Conventions
 *   - Every function returns VLM_OK (0) or a VLM_ERR_* code; the message of
 *     the last failure is available from vlm_last_error().  The library
 *     never aborts the process.
 *   - Host buffers are plain double arrays in Fortran (column-major) order:
 *     a point list is 3 x m (x,y,z of point 0 first).
 *   - All device work is ordered on the context's HIP stream; calls that
 *     return host data synchronise that stream.  A context is used by one
 *     host thread at a time.
 *   - Indices (rotor ir, blade b, rows, spans) are 0-based; cursor values
 *     rowNear/rowFar are reported 1-based, as in the reference.
 *
 * Lattice layout of a near wake (device-resident, exported by
 * vlm_rotor_get_wake):
 *   node(b,i,j)   trailing corner 2 of ring (i,j) for j < ns, corner 3 of
 *                 ring (i,ns-1) for j = ns        -> nodes[((b*nNwake+i)*(ns+1)+j)*3+d]
 *   le(b,j)       leading corners 1/4 of the newest row (the shed line)
 *                                                  -> le[(b*(ns+1)+j)*3+d]
 *   gam(b,i,j)    ring circulation                 -> gam[(b*nNwake+i)*ns+j]
 *   rc4(b,i,j,f)  core radius of filament f=0..3   -> rc4[((b*nNwake+i)*ns+j)*4+f]
 * Ring (i,j) has corners c1 = node(i-1,j) (le(j) on the newest row),
 * c2 = node(i,j), c3 = node(i,j+1), c4 = node(i-1,j+1) (le(j+1)).
 * A far-wake filament k: far_fc[(b*nFwake+k)*6 + e*3 + d], e=0 TE end
 * (fc(:,1)), e=1 LE end (fc(:,2)).
 * Velocity arrays: velN[((b*nNwake+i)*(ns+1)+j)*3+d], velF[(b*nFwake+k)*3+d].
 */
#ifndef VLM_C_H
#define VLM_C_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct vlm_ctx vlm_ctx;

enum {
    VLM_OK = 0,
    VLM_ERR_ARG = 1,       /* invalid argument or size */
    VLM_ERR_HIP = 2,       /* HIP runtime failure */
    VLM_ERR_STATE = 3,     /* call out of order */
    VLM_ERR_SINGULAR = 4,  /* zero pivot in the AIC factorisation */
    VLM_ERR_ALLOC = 5      /* device allocation failed */
};

/* Precision of the Biot-Savart pair arithmetic. */
enum {
    VLM_FP64 = 0,   /* FP64 throughout */
    VLM_MIXED = 1   /* FP32 pairs, compensated tile sums, FP64 near field */
};

/* Source sets of a rotor for vlm_rotor_induced_velocity. */
enum {
    VLM_SET_WING = 0,        /* rotor_vind_bywing, classdef.f90:4424-4443 */
    VLM_SET_WING_BOUND = 1,  /* rotor_vind_bywing_boundVortices, 4445-4457 */
    VLM_SET_WAKE = 2,        /* rotor_vind_bywake, 4459-4479 */
    VLM_SET_WAKE_PRED = 3    /* rotor_vind_bywake(P, 'P') */
};

enum { VLM_WAKE_C = 0, VLM_WAKE_P = 1 };     /* current / predicted wake */
enum { VLM_WAKE_A = 2 };                     /* every row of the current wake (mirror) */
enum { VLM_EDGE_LE = 0, VLM_EDGE_TE = 1 };   /* assignshed edge */
enum {                                        /* fdScheme, main.f90:844 */
    VLM_SCHEME_FT1 = 0, VLM_SCHEME_PC2 = 1, VLM_SCHEME_AB2 = 2,
    VLM_SCHEME_AM2 = 3, VLM_SCHEME_AM3 = 4, VLM_SCHEME_AM4 = 5
};
enum {                                        /* vlm_rotor_get_velocity */
    VLM_VEL = 0, VLM_VEL_PREV = 1, VLM_VEL_PRED = 2, VLM_VEL_STEP = 3
};

/* ---- context ---------------------------------------------------------- */
int vlm_create(vlm_ctx **ctx, int device);
int vlm_destroy(vlm_ctx *ctx);
const char *vlm_last_error(const vlm_ctx *ctx);
int vlm_set_precision(vlm_ctx *ctx, int precision);
/* bug_compat=1 (default) reproduces the predicted-wake row loop
 * `do i = 1, rowNear, nNwake` (classdef.f90:1552, 1559); 0 convects every
 * active predicted row. */
int vlm_set_bug_compat(vlm_ctx *ctx, int bug_compat);
int vlm_synchronize(vlm_ctx *ctx);

/* ---- a1/a2: direct sums over explicit filament lists ----------------- */
/* Sum over filaments s of gam[s] * vf_vind(p, a_s -> b_s, rc[s])
 * (classdef.f90:476-503), at m points.  a, b, p, v are 3 x n arrays.
 * accumulate != 0 adds into v. */
int vlm_filament_vind(vlm_ctx *ctx, int64_t nfil, const double *a,
                      const double *b, const double *gam, const double *rc,
                      int64_t m, const double *p, double *v, int accumulate);
/* The same sum on device-resident inputs and output, allocating nothing.
 * No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  a, b, p, v are device pointers to 3 x n arrays (as
 * above), gam and rc device arrays of nfil.  The call only enqueues work on
 * the context's stream; v holds the result once the stream reaches it
 * (vlm_synchronize).  All intermediate data lives in a 64 MiB device array
 * of the library, one per GPU, held until the process ends: about 120 B per
 * filament plus 72 B per point plus 24 B per point for each chunk of
 * partial sums must fit, else VLM_ERR_ARG with a message.  Calls that
 * overlap on one GPU share that array and race: serialise them. */
int vlm_filament_vind_noalloc(vlm_ctx *ctx, int64_t nfil, const double *a,
                              const double *b, const double *gam,
                              const double *rc, int64_t m, const double *p,
                              double *v, int accumulate);

/* Device-resident variant: upload a filament set and a point set once,
 * evaluate repeatedly, download the velocities. */
int vlm_fset_create(vlm_ctx *ctx, int64_t nfil, const double *a,
                    const double *b, const double *gam, const double *rc,
                    int *fset);
int vlm_pset_create(vlm_ctx *ctx, int64_t m, const double *p, int *pset);
int vlm_fset_vind(vlm_ctx *ctx, int fset, int pset);
int vlm_pset_velocity(vlm_ctx *ctx, int pset, double *v);

/* ---- rotor objects ----------------------------------------------------- */
typedef struct {
    int nb, nc, ns;          /* blades, chordwise and spanwise panels */
    int nNwake, nFwake;      /* near-wake rows, far-wake filaments per blade */
    int nbConvect;           /* blades convected (1 with axisymmetrySwitch) */
    int rollupStart, rollupEnd;  /* 1-based span range, classdef.f90:3135-3136 */
    int rollupMin;           /* sign(Omega*theta0) > 0: pick min gam, 4546-4554 */
    int suppressFwake;       /* suppressFwakeSwitch */
} vlm_rotor_desc;

/* rotor_init wake allocation, classdef.f90:3731-3850.  Wake cores start at
 * spanwiseCore (vf2, vf4) and streamwiseCoreVec(j)/(j+1) (vf1/vf3); far
 * filaments at streamwiseCoreVec(ns+1).  streamwiseCoreVec has ns+1
 * entries, already scaled by chord. */
int vlm_rotor_create(vlm_ctx *ctx, const vlm_rotor_desc *desc,
                     double spanwiseCore, const double *streamwiseCoreVec,
                     int *ir);
/* Wing vortex rings after rigid motion (classdef.f90:1092-1281):
 * corners[((b*ns+is)*nc+ic)*12 + k*3 + d], rc[((b*ns+is)*nc+ic)*4 + f],
 * gam[(b*ns+is)*nc+ic].  Any pointer may be NULL to keep the old value. */
int vlm_rotor_set_wing(vlm_ctx *ctx, int ir, const double *corners,
                       const double *rc, const double *gam);
int vlm_rotor_set_axis(vlm_ctx *ctx, int ir, const double *shaftAxis,
                       const double *hubCoords);

/* Cursor bookkeeping: main.f90:227-229 and 412-417. */
int vlm_rotor_init_cursors(vlm_ctx *ctx, int ir);
int vlm_rotor_begin_step(vlm_ctx *ctx, int ir, int iter);
int vlm_rotor_cursors(vlm_ctx *ctx, int ir, int *rowNear, int *rowFar);

/* Wake bookkeeping. */
int vlm_rotor_assignshed(vlm_ctx *ctx, int ir, int edge);    /* classdef.f90:4297-4325 */
int vlm_rotor_age_wake(vlm_ctx *ctx, int ir, double dt);      /* 4331-4354 */
int vlm_rotor_dissipate_wake(vlm_ctx *ctx, int ir, double dt,
                             double kinematicVisc, double apparentViscCoeff,
                             double decayCoeff);              /* 4356-4408 */
int vlm_rotor_strain_wake(vlm_ctx *ctx, int ir);              /* Far-wake burst (rotor_burst_wake, classdef.f90:4911-4917 and
 * blade_burst_wake 2306-2339, called every wakeBurst steps, main.f90:490-496):
 * for every blade, a far filament takes the core radius largeCoreRadius
 * (the reference passes the chord) when the skew |acos(cos) - pi| / pi of
 * the angle between it and a neighbouring far filament, reversed, reaches
 * skewLimit. */
int vlm_rotor_burst_wake(vlm_ctx *ctx, int ir, double skewLimit,
                         double largeCoreRadius);

/* 4410-4422 */
int vlm_rotor_rollup(vlm_ctx *ctx, int ir);                   /* 4515-4605 + shift 4481-4513 */
int vlm_rotor_convectwake(vlm_ctx *ctx, int ir, double dt,
                          int wakeType);                      /* 1515-1702, 4786-4830 */

/* Induced velocity at m host points by one source set of rotor ir
 * (classdef.f90:4424-4479); accumulate != 0 adds into v. */
int vlm_rotor_induced_velocity(vlm_ctx *ctx, int ir, int set, int64_t m,
                               const double *p, double *v, int accumulate);

/* A sum of source sets at the same m host points, applied in order on the
 * device with one upload and one download:
 *   u = accumulate ? v : 0;
 *   for k < nterms: if (flags[k] & VLM_TERM_NEGATE) u = -u;
 *                   u += velocity of set[k] of rotor ir[k];
 *   v = u.
 * Same arithmetic, in the same order, as one vlm_rotor_induced_velocity
 * call per term with accumulate set (negating v on the host where flagged),
 * without a synchronisation per term.  Replaces the sums over rotors of
 * rotor%vind_bywake / vind_bywing at the collocation points (RHS,
 * main.f90:549-561) and the bound-vortex and wing sums of the loads
 * (classdef.f90:1376-1418, 1734). */
#define VLM_TERM_NEGATE 1
int vlm_induced_velocity_terms(vlm_ctx *ctx, int nterms, const int *ir, const int *set,
                               const int *flags, int64_t m, const double *p, double *v,
                               int accumulate);

/* The N-body pass: velNwake/velFwake (wakeType C) or velNwakePredicted /
 * velFwakePredicted (P) of every rotor = sum over all rotors of wing +
 * wake induced velocity (libCommon.f90:114-211, main.f90:800-841 and
 * 1052-1086). */
int vlm_wake_velocity(vlm_ctx *ctx, int wakeType);

/* Initial wake velocity (classdef.f90:394 initWakeVel; libCommon.f90:106
 * initWakeVelNt).  While iter < initWakeVelNt, vlm_convect adds
 * initWakeVel * shaftAxis to the pass-1 near-wake velocity and subtracts it
 * from the pass-1 far-wake velocity (main.f90:829-838), and subtracts it
 * from both pass-2 velocities (902-913, 1070-1081, 1201-1210, 1353-1362):
 * the reference's signs, kept as they are.  The axis is the one of
 * vlm_rotor_set_axis. */
int vlm_rotor_set_init_wake_vel(vlm_ctx *ctx, int ir, double initWakeVel,
                                int initWakeVelNt);

/* Prescribed far wake (pFwake_class, classdef.f90:222-236; pFwake_update
 * 998-1066; rotor_updatePrescribedWake 5172-5217): per blade and wake type
 * a 10-revolution helix of 240 filaments, rebuilt after every convection of
 * that wake type (inside vlm_convect) once the step given to
 * vlm_rotor_begin_step exceeds prescWakeNt, from the far filaments at rowFar
 * and after (prescWakeGenNt = 0) or the last prescWakeGenNt + 1 of them;
 * deltaPsi = omegaSlow * dt of the step (set again whenever it changes).
 * While a far wake exists its filaments join the wake source sets.
 * prescWakeNt = 0 switches it off.  The shaft must lie along z. */
int vlm_rotor_set_prescribed_wake(vlm_ctx *ctx, int ir, int prescWakeNt, int prescWakeGenNt,
                                  double deltaPsi);
/* fc: 6 x 240 x nb end points (TE end, LE end), gam and rc: 240 x nb;
 * *built = 1 once the wake type has been built (geom2file's test). */
int vlm_rotor_get_prescribed_wake(vlm_ctx *ctx, int ir, int wakeType, double *fc, double *gam,
                                  double *rc, int *built);

/* Image rotors (surfaceType = -1; rotor_init 2798-2960, rotor_mirrorGamma
 * 4990-4995, rotor_mirrorWake 5024-5131; main.f90:181-192, 595-599,
 * 844-1406, 1418-1437): rotor ir becomes the mirror image of rotor source
 * in the plane where coordinate imagePlane (1 x, 2 y, 3 z) is zero.  Both
 * rotors must have the same nb, nc, ns, nNwake and nFwake, and the source
 * must not itself be an image.  An image takes no part as a target of the
 * wake passes (its velocities there would go unused); inside vlm_convect,
 * at its place in the rotor order, every convection of a lifting rotor
 * (wake type C or P) is a vlm_rotor_mirror_wake of that type for an image.
 * Its wing, wake and circulation remain sources like any rotor's. */
int vlm_rotor_set_image(vlm_ctx *ctx, int ir, int source, int imagePlane);

/* rotor_mirrorGamma followed by rotor_map_gam: the device wing circulation
 * of image ir becomes the negated wing circulation of its source, enqueued
 * in stream order (after a vlm_rotor_solve_deferred of the source that came
 * before it).  The caller's gamVec is the source's negated (the device
 * holds none for an image). */
int vlm_rotor_mirror_gamma(vlm_ctx *ctx, int ir);

/* rotor_mirrorWake: the image's wake becomes its source's with coordinate
 * imagePlane and the circulation negated, cores, ages and lengths copied.
 * VLM_WAKE_C: current near rows from rowNear and far filaments from rowFar;
 * VLM_WAKE_P: the predicted wake's positions over the same rows, with the
 * circulation, cores and ages (which the predicted wake shares with the
 * current one here, and which the reference copies inside its predicted
 * rings); VLM_WAKE_A:
 * every near row and far filament of the current wake (end of a step, in
 * place of the rollup and the trailing-edge shed). */
int vlm_rotor_mirror_wake(vlm_ctx *ctx, int ir, int wakeType);

/* One whole convection block, main.f90:800-1406: pass 1, the scheme
 * algebra of fdScheme (FT1, PC2, AB2, AM2, AM3 or AM4), pass 2 where the
 * scheme has one, and the final convection of every rotor (the wake
 * mirror for an image rotor).  The learned pass 2 (vlm_set_predictor)
 * replaces pass 2 of AM2 only. */
int vlm_convect(vlm_ctx *ctx, int scheme, int iter, double dt);

/* Restart snapshot of a rotor's device state: every wing, near-wake,
 * far-wake and velocity-history array and the wake cursors, as one opaque
 * byte string of vlm_rotor_state_bytes bytes.  vlm_rotor_set_state takes a
 * snapshot of the same case (sizes are checked) and resumes from it.  The
 * reference writes its whole rotor array for this (main.f90:390-398,
 * 1442-1450); the host writes these bytes beside its own state. */
int vlm_rotor_state_bytes(vlm_ctx *ctx, int ir, int64_t *bytes);
int vlm_rotor_get_state(vlm_ctx *ctx, int ir, void *buf, int64_t bytes);
int vlm_rotor_set_state(vlm_ctx *ctx, int ir, const void *buf, int64_t bytes);

/* State access (tests, restart files, plots). */
int vlm_rotor_get_wake(vlm_ctx *ctx, int ir, int wakeType, double *nodes,
                       double *le, double *gam, double *rc4, double *far_fc,
                       double *far_gam, double *far_rc);
int vlm_rotor_set_wake(vlm_ctx *ctx, int ir, const double *nodes,
                       const double *le, const double *gam, const double *rc4,
                       const double *far_fc, const double *far_gam,
                       const double *far_rc, int rowNear, int rowFar);
int vlm_rotor_get_velocity(vlm_ctx *ctx, int ir, int which, double *velN,
                           double *velF);

/* ---- a7: AIC and the circulation solve -------------------------------- */
/* AIC(row,col) = vr_vind(ring col, CP row) . nCap(row)
 * (rotor_calcAIC, classdef.f90:4151-4179); cp and ncap are 3 x N with
 * N = nb*nc*ns in the reference's row order.  The matrix is assembled,
 * factorised (LU, partial pivoting; blocked with an MFMA-FP64 trailing
 * update above order 256) and kept on the
 * device; A (N x N column-major) is copied out when non-NULL. */
int vlm_rotor_calc_aic(vlm_ctx *ctx, int ir, const double *cp,
                       const double *ncap, double *A);
/* gamVec = AIC^-1 RHS (main.f90:596, matmulAX libMath.f90:105-122) by LU
 * triangular solves plus iterative refinement with an FP64 residual,
 * stopping when ||r||_inf <= tol*||RHS||_inf or after max_iter steps.
 * The solved circulation also becomes the rotor's wing circulation
 * (rotor_map_gam, classdef.f90:4181-4196). */
int vlm_rotor_solve(vlm_ctx *ctx, int ir, const double *rhs, double *gam,
                    int max_iter, double tol, int *iters);

/* ---- deferred results ---------------------------------------------------
 * The _deferred variants enqueue the same work as their plain counterparts
 * and return at once; their host outputs (gam and *iters, v) are written
 * by vlm_wait, which waits only for the device work those results need,
 * or by any call that synchronises the stream (including vlm_destroy).
 * The output buffers must stay valid and unread until then; host inputs
 * are copied at the call.  A caller enqueues the wake passes in between and
 * computes loads while they run (main.f90 computes loads between the solve
 * and the convection; they depend only on the wing and its circulation,
 * which the convection does not change).  No reference counterpart. */
int vlm_rotor_solve_deferred(vlm_ctx *ctx, int ir, const double *rhs, double *gam,
                             int max_iter, double tol, int *iters);
int vlm_induced_velocity_terms_deferred(vlm_ctx *ctx, int nterms, const int *ir,
                                        const int *set, const int *flags, int64_t m,
                                        const double *p, double *v, int accumulate);
int vlm_wait(vlm_ctx *ctx);

/* ---- the step on the device: right-hand side, solve and loads ----------
 * With these a step runs from the wake sums through the circulation solve
 * to the loads without a host round trip (main.f90:521-764): the host
 * uploads the panel geometry after each rigid motion and reads the loads
 * at vlm_wait.  vlm_rotor_solve_deferred with rhs = NULL solves the
 * right-hand side vlm_rotor_rhs left on the device.
 *
 * vlm_rotor_set_panels: geometry after the last rigid motion, in ring
 * order (blade, section, chordwise panel): per panel 15 doubles, CP(3),
 * nCap(3), tauCapChord(3), tauCapSpan(3), panelArea, meanChord, meanSpan;
 * per blade and section 16, secCP(3), corner 1 of the section's leading
 * panel (3), secTauCapChord(3), secTauCapSpan(3), secNormalVec(3), secArea;
 * per blade 9, xAxisAziFlap, yAxisAziFlap, zAxisAziFlap. */
int vlm_rotor_set_panels(vlm_ctx *ctx, int ir, const double *panel, const double *section,
                         const double *blade);

/* Right-hand side of the convected blades (main.f90:121-186 initial,
 * 521-591 every step): velCPm = -velBody - omegaBody x (CP - cgCoords)
 * - omegaSlow shaft x (CP - hub), velCP = velCPm + the ordered term sum at
 * the collocation points (as vlm_induced_velocity_terms), RHS = -(velCP .
 * nCap), copied from blade 1 under axisymmetry.  kin: velBody(3),
 * omegaBody(3), cgCoords(3), omegaSlow; shaft and hub from
 * vlm_rotor_set_axis.  velCP, velCPm and the RHS stay on the device; rhs,
 * when not NULL, receives a host copy at vlm_wait. */
int vlm_rotor_rhs(vlm_ctx *ctx, int ir, int nterms, const int *irs, const int *sets,
                  const int *flags, const double *kin, double *rhs);

/* rotor_mirrorVelCP (classdef.f90:4997-5022): an image's velCP and velCPm
 * are its source's (as left by vlm_rotor_rhs) with coordinate imagePlane
 * negated, on every blade. */
int vlm_rotor_mirror_velcp(vlm_ctx *ctx, int ir);

/* Loads of forceCalcSwitch = 0 (main.f90:247-290 and 632-668;
 * blade_calc_secChordwiseResVel, blade_calc_secAlpha, rotor_dirLiftDrag,
 * blade_calc_force and rotor_calc_force, classdef.f90:1704-1896,
 * 2197-2265, 4607-4671, 4766-4784): velCPTotal = velCP + the ordered term
 * sum at the collocation points (the caller lists every rotor's bound
 * vortices, then the wing's own rings with VLM_TERM_NEGATE; the result is
 * the negated sum, as vlm_induced_velocity_terms), then per convected
 * blade and section the chordwise resultant velocity (quadratic fit over
 * three or more panels), angles, lift direction and the panel pressure
 * loop with gamPrev kept on the device (restart state included), and the
 * blade and rotor sums.  out (host, written at vlm_wait), per convected
 * blade and section 19 doubles: secChordwiseResVel(3), secAlpha, secPhi,
 * secViz, secVix, secTheta, secLift(3), secLiftInPlane(3),
 * secLiftOutPlane(3), secCL, secCLu; then the rotor's forceInertial(3),
 * lift(3), liftUnsteady(3) (drag is zero in this load model). */
int vlm_rotor_loads(vlm_ctx *ctx, int ir, int nterms, const int *irs, const int *sets,
                    const int *flags, double density, double dt, double omegaSign,
                    int spanwiseLiftSwitch, double *out);

/* Dense kernels on explicit matrices (the LU microbenchmark and tests):
 * factorise an n x n column-major matrix in place (ipiv 0-based), and
 * solve with refinement against the unfactorised A. */
int vlm_lu_factor(vlm_ctx *ctx, int n, double *A, int *ipiv);
int vlm_lu_solve_refine(vlm_ctx *ctx, int n, const double *A,
                        const double *LU, const int *ipiv, const double *rhs,
                        double *x, int max_iter, double tol, int *iters);

/* ---- a13: learned pass-2 predictor (no reference counterpart) -------- */
/* MLP 45 -> 64 -> 64 -> 64 -> 3 (tanh), FP32, parameters packed as
 * W1(64x45) b1 W2(64x64) b2 W3(64x64) b3 W4(3x64) b4, row-major W[out][in]. */
int vlm_mlp_num_params(void);
int vlm_mlp_set_params(vlm_ctx *ctx, const float *params);
int vlm_mlp_get_params(vlm_ctx *ctx, float *params);
/* y (m x 3) = MLP(x (m x 45)) */
int vlm_mlp_forward(vlm_ctx *ctx, int64_t m, const float *x, float *y);
/* loss = 0.5/m sum |MLP(x) - t|^2 and its gradient (packed like params) */
int vlm_mlp_loss_grad(vlm_ctx *ctx, int64_t m, const float *x, const float *t,
                      float *grad, float *loss);
/* one Adam step (beta 0.9/0.999) on the batch */
int vlm_mlp_train_step(vlm_ctx *ctx, int64_t m, const float *x, const float *t,
                       float lr, float *loss);
/* mode 0: pass 2 of AM2 is the full Biot-Savart sum (the reference's
 * method, main.f90:1052-1086); mode 1: the MLP proposes the pass-2
 * velocity from the 12 nearest lattice offsets (scaled by 1/lref) and the
 * node's pass-1, previous and predictor velocities; every step the full sum
 * on every sample_stride-th target judges it, and the whole pass is
 * recomputed when RMS(error)/RMS(speed) on the sample exceeds tol.  The
 * first warmup_steps AM2 steps always run the full pass; judged samples and
 * full passes feed train_iters Adam iterations per step. */
int vlm_set_predictor(vlm_ctx *ctx, int mode, double tol, int sample_stride,
                      int warmup_steps, int train_iters, double lref);
int vlm_predictor_stats(vlm_ctx *ctx, int64_t *full_passes, int64_t *learned_passes,
                        int64_t *fallbacks, double *last_err, double *last_loss);
/* Accuracy of the learned pass 2 over all judged steps so far: RMS error
 * relative to the RMS exact speed on the judge samples, for the network and
 * for the network-free substitute of the same cost (the linear
 * extrapolation 2 v_n - v_(n-1) of the pass-1 velocity, which turns the AM2
 * corrector into AB2).  base_rel_rms / mlp_rel_rms is the network's gain. */
int vlm_predictor_errors(vlm_ctx *ctx, double *mlp_rel_rms, double *base_rel_rms,
                         int64_t *judged_steps);

/* ---- kernel timing ----------------------------------------------------- */
/* When enabled, every work-queue Biot-Savart launch (1024 targets or more)
 * is bracketed by HIP events on the context stream.  vlm_profile_get returns the summed kernel time (ms),
 * the launch count and the algorithmic pair count since the last reset. */
int vlm_profile_enable(vlm_ctx *ctx, int enable);
int vlm_profile_reset(vlm_ctx *ctx);
int vlm_profile_get(vlm_ctx *ctx, double *bs_ms, int64_t *bs_launches,
                    double *pairs);
/* Pair count of every Biot-Savart launch since the last reset, the
 * small-target launches (fewer than 1024 targets: collocation points,
 * bound vortices, loads) included, and the number of those launches. */
int vlm_profile_get_all(vlm_ctx *ctx, double *pairs, int64_t *launches);
/* Device time of the LU factorisations (lu_factor on the device, events
 * around the whole factorisation) since the last reset, their count and
 * their nominal flops (2 n^3 / 3 each). */
int vlm_profile_get_lu(vlm_ctx *ctx, double *ms, int64_t *factorisations,
                       double *flops);
/* Calibration of the roofline (no counterpart in the reference): the
 * sustained FP64 rate of this device on a stream of independent FMAs
 * (8 chains per lane, 16 workgroups of 256 per CU, after one untimed round),
 * in TFLOP/s, and the time of the same stream of v_rsq_f64 relative to it.
 * Allocates a scratch buffer of 8 bytes per launched lane for the call. */
int vlm_fp64_issue_rate(vlm_ctx *ctx, double *fma_tflops, double *rsq_over_fma);

#ifdef __cplusplus
}
#endif
#endif /* VLM_C_H */
