/*
 * CPU oracle (test infrastructure only): a plain-C FP64 restatement of the
 * VOLCANOR free-wake hot path, kept in the reference's own ring form
 * (every vortex ring owns four filaments; wakes shift by row copies).
 *
 * It is the checker for the HIP library, never part of the product path:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * it.  Every function cites the reference routine it follows
 * (cibinjoseph/VOLCANOR, file:line).
 *
 * Array conventions follow Fortran column-major order so that the flat
 * buffers read like the reference's arrays:
 *   wing ring corners   corners[((b*ns + is)*nc + ic)*12 + k*3 + d]   (corner k=0..3)
 *   wing ring cores     rc[((b*ns + is)*nc + ic)*4 + f]               (filament f=0..3)
 *   near-wake velocity  vel[((b*(ns+1) + j)*nNwake + i)*3 + d]        (= velNwake(d,i,j))
 *   far-wake velocity   vel[(b*nFwake + i)*3 + d]                     (= velFwake(d,i))
 * Indices i, j, ic, is, b are 0-based here; row i here is row i+1 in the
 * reference.
 */
#ifndef VLM_ORACLE_H
#define VLM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- a1/a2: single filament and ring (classdef.f90:476-503, 527-542) ---- */
void or_vf_vind(const double P[3], const double A[3], const double B[3],
                double rc, double v[3]);
void or_filament_sum(int64_t nfil, const double *a, const double *b,
                     const double *gam, const double *rc, int64_t m,
                     const double *p, double *v);
/* Multi-threaded variant used only for the bench's CPU baseline. */
void or_filament_sum_omp(int64_t nfil, const double *a, const double *b,
                         const double *gam, const double *rc, int64_t m,
                         const double *p, double *v, int nthreads);

/* ---- rotor-system state in ring form ---- */
typedef struct or_sys or_sys;

or_sys *or_sys_create(void);
void or_sys_destroy(or_sys *s);
/* Returns rotor index. rollup_min != 0 picks the minimum circulation in the
 * tip rollup (sign(Omega*theta0) > 0, classdef.f90:4546-4554). */
int or_rotor_create(or_sys *s, int nb, int nc, int ns, int nNwake, int nFwake,
                    int nbConvect, int rollupStart, int rollupEnd,
                    int rollup_min, int suppressFwake);
void or_rotor_set_wing(or_sys *s, int ir, const double *corners,
                       const double *rc, const double *gam);
void or_rotor_set_axis(or_sys *s, int ir, const double shaft[3],
                       const double hub[3]);
void or_rotor_init_wake_cores(or_sys *s, int ir, double spanwiseCore,
                              const double *streamwiseCoreVec);
void or_rotor_init_cursors(or_sys *s, int ir);
void or_rotor_begin_step(or_sys *s, int ir, int iter);
void or_rotor_cursors(const or_sys *s, int ir, int *rowNear, int *rowFar);

/* Wake bookkeeping (classdef.f90:4297-4325, 4331-4422, 4481-4605). */
void or_rotor_assignshed(or_sys *s, int ir, int edge_te);
void or_rotor_age_wake(or_sys *s, int ir, double dt, double omegaSlow);
void or_rotor_dissipate_wake(or_sys *s, int ir, double dt, double kinVisc,
                             double apparentViscCoeff, double decayCoeff);
void or_rotor_strain_wake(or_sys *s, int ir);
void or_rotor_rollup(or_sys *s, int ir);
void or_rotor_convectwake(or_sys *s, int ir, double dt, int predicted,
                          int bug_compat);

/* Induced velocity at arbitrary points by the sources of rotor ir
 * (classdef.f90:4424-4479).  set: 0 wing, 1 wing bound vortices,
 * 2 wake (current), 3 wake (predicted).  Accumulates into v. */
void or_rotor_vind(const or_sys *s, int ir, int set, int64_t m,
                   const double *p, double *v);

/* Pass over all wake nodes of all rotors (main.f90:800-841 / 1052-1086):
 * velNwake (or velNwakePredicted) = sum over rotors of wing + wake. */
void or_wake_velocity(or_sys *s, int predicted);

/* Whole convection block of main.f90:844-1406 for fdScheme 0..3. */
void or_convect(or_sys *s, int scheme, int iter, double dt, int bug_compat);

/* Image rotors (surfaceType -1): rotor ir mirrors rotor src in the plane
 * where coordinate `plane` (1..3) is zero; or_convect then mirrors instead
 * of convecting it.  mirror_gamma: wing circulation = -source's;
 * mirror_wake: mode 0 'C', 1 'P', 2 'A' (rotor_mirrorWake). */
void or_rotor_set_image(or_sys *s, int ir, int src, int plane);
void or_rotor_mirror_gamma(or_sys *s, int ir);
void or_rotor_mirror_wake(or_sys *s, int ir, int mode);

/* Accessors in the lattice form used by the library (see vlm_c.h). */
/* Prescribed far wake (pFwake_update, classdef.f90:998-1066, 5172-5217):
 * rebuilt after every convection of the wake type once the step number
 * of or_rotor_begin_step exceeds presc_nt; dpsi = omegaSlow * dt.
 * or_rotor_get_prescribed: 6 x 240 x nb end points (TE end, LE end), the
 * circulation and core radius of each filament; returns 1 once built. */
void or_rotor_set_prescribed(or_sys *s, int ir, int presc_nt, int presc_gen_nt, double dpsi);
int or_rotor_get_prescribed(const or_sys *s, int ir, int predicted, double *fc, double *gam,
                            double *rc);
void or_rotor_get_wake(const or_sys *s, int ir, int predicted, double *nodes,
                       double *le, double *gam, double *rc4, double *far_fc,
                       double *far_gam, double *far_rc);
void or_rotor_set_wake(or_sys *s, int ir, const double *nodes,
                       const double *le, const double *gam, const double *rc4,
                       const double *far_fc, const double *far_gam,
                       const double *far_rc, int rowNear, int rowFar);
/* which: 0 velNwake/velFwake, 1 *1 (history), 2 *Predicted, 3 *Step */
void or_rotor_get_velocity(const or_sys *s, int ir, int which, double *velN,
                           double *velF);

/* ---- a7: AIC and dense solve (classdef.f90:4151-4179, libMath.f90:48-122) */
void or_rotor_calc_aic(const or_sys *s, int ir, const double *cp,
                       const double *ncap, double *A);
int or_lu_factor(int n, double *A, int *ipiv);
void or_lu_solve(int n, const double *LU, const int *ipiv, double *b);
void or_matvec(int n, const double *A, const double *x, double *y);

/* ---- a13: learned predictor MLP (no reference counterpart) ---- */
void or_mlp_forward(const double *params, int n, const double *x,
                    double *y);
double or_mlp_loss_grad(const double *params, int n, const double *x,
                        const double *target, double *grad);

#ifdef __cplusplus
}
#endif
#endif
