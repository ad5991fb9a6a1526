/* Test infrastructure: the libvlm C ABI (include/vlm_c.h) implemented on the
 * CPU over the oracle's ring-form object model (vlm_oracle.c), for the entry
 * points the Fortran host's run_case calls.  It exists to run the Fortran
 * host on a machine without a GPU, to investigate host-side behaviour (I/O,
 * restart records, plot writers, cursors).  It is linked only into the CPU
 * build of the host under oracle/_cpu (oracle/Makefile, target hostcpu);
 * the product library is libvlm.so and never loads this file.
 *
 * Deferred results are computed at once.  Unsupported options (the learned
 * predictor, the far-wake burst, initWakeVel, AM3/AM4) return an error. */
#include "vlm_oracle.c"

#include <stdio.h>

#include "../include/vlm_c.h"

struct vlm_ctx {
    or_sys *s;
    int bug_compat;
    char err[512];
    double *lu[OR_MAX_ROTORS];
    int *piv[OR_MAX_ROTORS];
    int n[OR_MAX_ROTORS];
    /* right-hand side and loads (vlm_rotor_set_panels records; velCP and
     * velCPm, 6 per panel; RHS; gamPrev of the loads) */
    double *pan[OR_MAX_ROTORS], *sec[OR_MAX_ROTORS], *bla[OR_MAX_ROTORS];
    double *vcp[OR_MAX_ROTORS], *rhs[OR_MAX_ROTORS], *gprev[OR_MAX_ROTORS];
    int image_of[OR_MAX_ROTORS], image_axis[OR_MAX_ROTORS];
};

static int shim_fail(vlm_ctx *c, const char *m)
{
    if (c) snprintf(c->err, sizeof(c->err), "%s", m);
    return VLM_ERR_ARG;
}

#define ROTOR_OK(c, ir) \
    if (!(c) || (ir) < 0 || (ir) >= (c)->s->nr) return shim_fail(c, "unknown rotor")

int vlm_create(vlm_ctx **ctx, int device)
{
    (void)device;
    vlm_ctx *c = (vlm_ctx *)calloc(1, sizeof(vlm_ctx));
    c->s = or_sys_create();
    c->bug_compat = 1;
    *ctx = c;
    return VLM_OK;
}

int vlm_destroy(vlm_ctx *c)
{
    if (!c) return VLM_OK;
    for (int i = 0; i < OR_MAX_ROTORS; ++i) {
        free(c->lu[i]);
        free(c->piv[i]);
        free(c->pan[i]);
        free(c->sec[i]);
        free(c->bla[i]);
        free(c->vcp[i]);
        free(c->rhs[i]);
        free(c->gprev[i]);
    }
    or_sys_destroy(c->s);
    free(c);
    return VLM_OK;
}

const char *vlm_last_error(const vlm_ctx *c) { return c ? c->err : "null context"; }
int vlm_set_precision(vlm_ctx *c, int p) { return p == VLM_FP64 ? VLM_OK : shim_fail(c, "FP64 only"); }
int vlm_set_bug_compat(vlm_ctx *c, int b) { c->bug_compat = b; return VLM_OK; }
int vlm_synchronize(vlm_ctx *c) { (void)c; return VLM_OK; }
int vlm_wait(vlm_ctx *c) { (void)c; return VLM_OK; }

int vlm_set_predictor(vlm_ctx *c, int mode, double tol, int stride, int warmup, int iters,
                      double lref)
{
    (void)tol; (void)stride; (void)warmup; (void)iters; (void)lref;
    return mode == 0 ? VLM_OK : shim_fail(c, "no learned predictor on the CPU");
}

int vlm_rotor_create(vlm_ctx *c, const vlm_rotor_desc *d, double spanwiseCore, const double *scv,
                     int *ir)
{
    const int k = or_rotor_create(c->s, d->nb, d->nc, d->ns, d->nNwake, d->nFwake, d->nbConvect,
                                  d->rollupStart, d->rollupEnd, d->rollupMin, d->suppressFwake);
    if (k < 0) return shim_fail(c, "too many rotors");
    if (d->nNwake > 0) or_rotor_init_wake_cores(c->s, k, spanwiseCore, scv);
    const size_t n = (size_t)d->nb * d->nc * d->ns, nsec = (size_t)d->nb * d->ns;
    c->pan[k] = (double *)calloc(15 * n, sizeof(double));
    c->sec[k] = (double *)calloc(16 * nsec, sizeof(double));
    c->bla[k] = (double *)calloc(9 * (size_t)d->nb, sizeof(double));
    c->vcp[k] = (double *)calloc(6 * n, sizeof(double));
    c->rhs[k] = (double *)calloc(n, sizeof(double));
    c->gprev[k] = (double *)calloc(n, sizeof(double));
    c->image_of[k] = -1;
    *ir = k;
    return VLM_OK;
}

int vlm_rotor_set_wing(vlm_ctx *c, int ir, const double *corners, const double *rc,
                       const double *gam)
{
    ROTOR_OK(c, ir);
    or_rotor_set_wing(c->s, ir, corners, rc, gam);
    return VLM_OK;
}

int vlm_rotor_set_axis(vlm_ctx *c, int ir, const double *shaft, const double *hub)
{
    ROTOR_OK(c, ir);
    or_rotor_set_axis(c->s, ir, shaft, hub);
    return VLM_OK;
}

int vlm_rotor_set_init_wake_vel(vlm_ctx *c, int ir, double v, int nt)
{
    (void)nt;
    ROTOR_OK(c, ir);
    return v == 0.0 ? VLM_OK : shim_fail(c, "no initWakeVel on the CPU");
}

int vlm_rotor_set_prescribed_wake(vlm_ctx *c, int ir, int nt, int gen_nt, double dpsi)
{
    ROTOR_OK(c, ir);
    or_rotor_set_prescribed(c->s, ir, nt, gen_nt, dpsi);
    return VLM_OK;
}

int vlm_rotor_get_prescribed_wake(vlm_ctx *c, int ir, int wakeType, double *fc, double *gam,
                                  double *rc, int *built)
{
    ROTOR_OK(c, ir);
    const int b = or_rotor_get_prescribed(c->s, ir, wakeType, fc, gam, rc);
    if (built) *built = b;
    return VLM_OK;
}

int vlm_rotor_init_cursors(vlm_ctx *c, int ir) { ROTOR_OK(c, ir); or_rotor_init_cursors(c->s, ir); return VLM_OK; }
int vlm_rotor_begin_step(vlm_ctx *c, int ir, int it) { ROTOR_OK(c, ir); or_rotor_begin_step(c->s, ir, it); return VLM_OK; }

int vlm_rotor_cursors(vlm_ctx *c, int ir, int *rn, int *rf)
{
    ROTOR_OK(c, ir);
    or_rotor_cursors(c->s, ir, rn, rf);
    return VLM_OK;
}

int vlm_rotor_assignshed(vlm_ctx *c, int ir, int edge) { ROTOR_OK(c, ir); or_rotor_assignshed(c->s, ir, edge); return VLM_OK; }
int vlm_rotor_age_wake(vlm_ctx *c, int ir, double dt) { ROTOR_OK(c, ir); or_rotor_age_wake(c->s, ir, dt, 0.0); return VLM_OK; }

int vlm_rotor_dissipate_wake(vlm_ctx *c, int ir, double dt, double nu, double avc, double decay)
{
    ROTOR_OK(c, ir);
    or_rotor_dissipate_wake(c->s, ir, dt, nu, avc, decay);
    return VLM_OK;
}

int vlm_rotor_burst_wake(vlm_ctx *c, int ir, double skew, double chord)
{
    (void)ir; (void)skew; (void)chord;
    return shim_fail(c, "no wake burst on the CPU");
}

int vlm_rotor_strain_wake(vlm_ctx *c, int ir) { ROTOR_OK(c, ir); or_rotor_strain_wake(c->s, ir); return VLM_OK; }
int vlm_rotor_rollup(vlm_ctx *c, int ir) { ROTOR_OK(c, ir); or_rotor_rollup(c->s, ir); return VLM_OK; }

int vlm_rotor_set_image(vlm_ctx *c, int ir, int source, int imagePlane)
{
    ROTOR_OK(c, ir);
    ROTOR_OK(c, source);
    if (source == ir || imagePlane < 1 || imagePlane > 3) return shim_fail(c, "bad image rotor");
    or_rotor_set_image(c->s, ir, source, imagePlane);
    c->image_of[ir] = source;
    c->image_axis[ir] = imagePlane - 1;
    return VLM_OK;
}
int vlm_rotor_mirror_gamma(vlm_ctx *c, int ir) { ROTOR_OK(c, ir); or_rotor_mirror_gamma(c->s, ir); return VLM_OK; }
int vlm_rotor_mirror_wake(vlm_ctx *c, int ir, int wakeType)
{
    ROTOR_OK(c, ir);
    if (wakeType < 0 || wakeType > 2) return shim_fail(c, "wakeType must be 0, 1 or 2");
    or_rotor_mirror_wake(c->s, ir, wakeType == VLM_WAKE_A ? 2 : wakeType);
    return VLM_OK;
}

int vlm_convect(vlm_ctx *c, int scheme, int iter, double dt)
{
    if (scheme < 0 || scheme > 3) return shim_fail(c, "schemes FT1, PC2, AB2 and AM2 only");
    or_convect(c->s, scheme, iter, dt, c->bug_compat);
    return VLM_OK;
}

int vlm_rotor_calc_aic(vlm_ctx *c, int ir, const double *cp, const double *ncap, double *A)
{
    ROTOR_OK(c, ir);
    const rotor_t *r = &c->s->rotor[ir];
    const int n = r->nb * r->nc * r->ns;
    free(c->lu[ir]);
    free(c->piv[ir]);
    c->lu[ir] = (double *)malloc(sizeof(double) * (size_t)n * n);
    c->piv[ir] = (int *)malloc(sizeof(int) * (size_t)n);
    c->n[ir] = n;
    or_rotor_calc_aic(c->s, ir, cp, ncap, c->lu[ir]);
    if (A) memcpy(A, c->lu[ir], sizeof(double) * (size_t)n * n);
    if (or_lu_factor(n, c->lu[ir], c->piv[ir]) != 0) return shim_fail(c, "singular AIC");
    return VLM_OK;
}

int vlm_rotor_solve(vlm_ctx *c, int ir, const double *rhs, double *gam, int max_iter, double tol,
                    int *iters)
{
    (void)max_iter; (void)tol;
    ROTOR_OK(c, ir);
    const int n = c->n[ir];
    if (!c->lu[ir]) return shim_fail(c, "no AIC");
    memcpy(gam, rhs ? rhs : c->rhs[ir], sizeof(double) * (size_t)n);
    or_lu_solve(n, c->lu[ir], c->piv[ir], gam);
    /* the solved circulation becomes the wing's (blade 1 copied to all
       blades when only blade 1 is convected) */
    const rotor_t *r = &c->s->rotor[ir];
    double *g = (double *)malloc(sizeof(double) * (size_t)n);
    const int per = r->nc * r->ns;
    for (int b = 0; b < r->nb; ++b)
        memcpy(g + (size_t)b * per, gam + (size_t)(r->nbConvect == 1 ? 0 : b) * per,
               sizeof(double) * (size_t)per);
    or_rotor_set_wing(c->s, ir, NULL, NULL, g);
    free(g);
    if (iters) *iters = 0;
    return VLM_OK;
}

int vlm_rotor_solve_deferred(vlm_ctx *c, int ir, const double *rhs, double *gam, int max_iter,
                             double tol, int *iters)
{
    return vlm_rotor_solve(c, ir, rhs, gam, max_iter, tol, iters);
}

int vlm_induced_velocity_terms(vlm_ctx *c, int nterms, const int *ir, const int *set,
                               const int *flags, int64_t m, const double *p, double *v,
                               int accumulate)
{
    if (!accumulate) memset(v, 0, sizeof(double) * 3 * (size_t)m);
    for (int k = 0; k < nterms; ++k) {
        if (ir[k] < 0 || ir[k] >= c->s->nr) return shim_fail(c, "unknown rotor");
        if (flags[k] & VLM_TERM_NEGATE)
            for (int64_t i = 0; i < 3 * m; ++i) v[i] = -v[i];
        or_rotor_vind(c->s, ir[k], set[k], m, p, v);   /* adds into v */
    }
    return VLM_OK;
}

int vlm_induced_velocity_terms_deferred(vlm_ctx *c, int nterms, const int *ir, const int *set,
                                        const int *flags, int64_t m, const double *p, double *v,
                                        int accumulate)
{
    return vlm_induced_velocity_terms(c, nterms, ir, set, flags, m, p, v, accumulate);
}

int vlm_rotor_get_wake(vlm_ctx *c, int ir, int wakeType, double *nodes, double *le, double *gam,
                       double *rc4, double *ffc, double *fgam, double *frc)
{
    ROTOR_OK(c, ir);
    or_rotor_get_wake(c->s, ir, wakeType, nodes, le, gam, rc4, ffc, fgam, frc);
    return VLM_OK;
}

/* Restart state: the rotor's cursors, every wake and velocity array and
 * the prescribed far wake. */
static size_t state_size(const rotor_t *r)
{
    const size_t npan = (size_t)r->nb * r->nc * r->ns;
    const size_t nw = (size_t)r->nNwake * r->ns, nf = (size_t)(r->nFwake > 0 ? r->nFwake : 1);
    const size_t nvn = 3u * (size_t)r->nNwake * (r->ns + 1), nvf = 3u * nf;
    return 4 * sizeof(int) + npan * sizeof(double) +
           (size_t)r->nb * (2 * nw * sizeof(ring_t) + 2 * nf * sizeof(fil_t) +
                            4 * (nvn + nvf) * sizeof(double) + 2 * OR_NPF * sizeof(fil_t) +
                            4 * sizeof(double));
}

static void state_io(rotor_t *r, double *gprev, char *buf, int out)
{
    const size_t nw = (size_t)r->nNwake * r->ns, nf = (size_t)(r->nFwake > 0 ? r->nFwake : 1);
    const size_t nvn = 3u * (size_t)r->nNwake * (r->ns + 1), nvf = 3u * nf;
#define IO(ptr, bytes)                                     \
    do {                                                   \
        if (out) memcpy(buf, (ptr), (bytes));              \
        else memcpy((ptr), buf, (bytes));                  \
        buf += (bytes);                                    \
    } while (0)
    IO(&r->rowNear, sizeof(int));
    IO(&r->rowFar, sizeof(int));
    IO(&r->presc_present[0], sizeof(int));
    IO(&r->presc_present[1], sizeof(int));
    IO(gprev, (size_t)r->nb * r->nc * r->ns * sizeof(double));
    for (int t = 0; t < 2; ++t) {
        IO(r->pf[t], (size_t)r->nb * OR_NPF * sizeof(fil_t));
        IO(r->presc_pitch[t], (size_t)r->nb * sizeof(double));
        IO(r->presc_radius[t], (size_t)r->nb * sizeof(double));
    }
    for (int b = 0; b < r->nb; ++b) {
        blade_t *bl = &r->blade[b];
        IO(bl->waN, nw * sizeof(ring_t));
        IO(bl->waNP, nw * sizeof(ring_t));
        IO(bl->waF, nf * sizeof(fil_t));
        IO(bl->waFP, nf * sizeof(fil_t));
        IO(bl->velN, nvn * sizeof(double));
        IO(bl->velN1, nvn * sizeof(double));
        IO(bl->velNP, nvn * sizeof(double));
        IO(bl->velNS, nvn * sizeof(double));
        IO(bl->velF, nvf * sizeof(double));
        IO(bl->velF1, nvf * sizeof(double));
        IO(bl->velFP, nvf * sizeof(double));
        IO(bl->velFS, nvf * sizeof(double));
    }
#undef IO
}

int vlm_rotor_state_bytes(vlm_ctx *c, int ir, int64_t *bytes)
{
    ROTOR_OK(c, ir);
    *bytes = (int64_t)state_size(&c->s->rotor[ir]);
    return VLM_OK;
}

int vlm_rotor_get_state(vlm_ctx *c, int ir, void *buf, int64_t bytes)
{
    ROTOR_OK(c, ir);
    if ((size_t)bytes != state_size(&c->s->rotor[ir])) return shim_fail(c, "state size");
    state_io(&c->s->rotor[ir], c->gprev[ir], (char *)buf, 1);
    return VLM_OK;
}

int vlm_rotor_set_state(vlm_ctx *c, int ir, const void *buf, int64_t bytes)
{
    ROTOR_OK(c, ir);
    if ((size_t)bytes != state_size(&c->s->rotor[ir])) return shim_fail(c, "state size");
    state_io(&c->s->rotor[ir], c->gprev[ir], (char *)buf, 0);
    return VLM_OK;
}

int vlm_profile_enable(vlm_ctx *c, int e) { (void)c; (void)e; return VLM_OK; }
int vlm_profile_reset(vlm_ctx *c) { (void)c; return VLM_OK; }

int vlm_profile_get(vlm_ctx *c, double *ms, int64_t *n, double *pairs)
{
    (void)c;
    *ms = 0.0;
    *n = 0;
    *pairs = 0.0;
    return VLM_OK;
}

int vlm_profile_get_all(vlm_ctx *c, double *pairs, int64_t *n)
{
    (void)c;
    *pairs = 0.0;
    *n = 0;
    return VLM_OK;
}

int vlm_predictor_stats(vlm_ctx *c, int64_t *full, int64_t *learned, int64_t *fallbacks,
                        double *last_err, double *last_loss)
{
    (void)c;
    *full = *learned = *fallbacks = 0;
    *last_err = *last_loss = 0.0;
    return VLM_OK;
}

int vlm_predictor_errors(vlm_ctx *c, double *mlp, double *base, int64_t *judged)
{
    (void)c;
    *mlp = *base = 0.0;
    *judged = 0;
    return VLM_OK;
}

int vlm_profile_get_lu(vlm_ctx *c, double *ms, int64_t *n, double *flops)
{
    (void)c;
    *ms = 0.0;
    *n = 0;
    *flops = 0.0;
    return VLM_OK;
}

/* ---- right-hand side and loads (vlm_rotor_set_panels, vlm_rotor_rhs,
 * vlm_rotor_mirror_velcp, vlm_rotor_loads): a serial restatement of
 * main.f90:521-591 and 632-668 with blade_calc_secChordwiseResVel,
 * blade_calc_secAlpha, rotor_dirLiftDrag, blade_calc_force and
 * rotor_calc_force (classdef.f90:1704-1896, 2197-2265, 4607-4671,
 * 4766-4784), records as in include/vlm_c.h. */
static double sh_dot(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static void sh_cross(const double *a, const double *b, double *c)
{
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}
static void sh_unit(const double *a, double *u)
{
    const double n = sqrt(sh_dot(a, a));
    for (int i = 0; i < 3; ++i) u[i] = n > OR_EPS ? a[i] / n : 0.0;
}
static void sh_proj(const double *a, const double *d, double *p)
{
    const double n2 = sh_dot(d, d), t = sh_dot(a, d);
    for (int i = 0; i < 3; ++i) p[i] = n2 > OR_EPS ? t * d[i] / n2 : 0.0;
}
static void sh_noproj(const double *a, const double *d, double *p)
{
    const double n2 = sh_dot(d, d), t = sh_dot(a, d);
    for (int i = 0; i < 3; ++i) p[i] = n2 > OR_EPS ? a[i] - t * d[i] / n2 : a[i];
}
static double sh_angle(const double *a, const double *b)
{
    const double d = sh_dot(a, b);
    return atan2(sqrt(sh_dot(a, a) * sh_dot(b, b) - d * d), d);
}
/* quadratic least-squares fit (libMath lsq2, 577-605) from its normal sums */
static double sh_lsq2(double xq, const double *s, double b0, double b1, double b2)
{
    double A[3][3] = {{s[0], s[1], s[2]}, {s[1], s[2], s[3]}, {s[2], s[3], s[4]}};
    double b[3] = {b0, b1, b2}, c[3];
    for (int k = 0; k < 3; ++k) {
        int p = k;
        for (int i = k + 1; i < 3; ++i)
            if (fabs(A[i][k]) > fabs(A[p][k])) p = i;
        if (p != k) {
            for (int j = 0; j < 3; ++j) {
                const double t = A[k][j];
                A[k][j] = A[p][j];
                A[p][j] = t;
            }
            const double t = b[k];
            b[k] = b[p];
            b[p] = t;
        }
        for (int i = k + 1; i < 3; ++i) {
            const double f = A[i][k] / A[k][k];
            for (int j = k; j < 3; ++j) A[i][j] = A[i][j] - f * A[k][j];
            b[i] = b[i] - f * b[k];
        }
    }
    for (int i = 2; i >= 0; --i) {
        c[i] = b[i];
        for (int j = i + 1; j < 3; ++j) c[i] = c[i] - A[i][j] * c[j];
        c[i] = c[i] / A[i][i];
    }
    return c[0] + c[1] * xq + c[2] * xq * xq;
}

int vlm_rotor_set_panels(vlm_ctx *c, int ir, const double *panel, const double *section,
                         const double *blade)
{
    ROTOR_OK(c, ir);
    const rotor_t *r = &c->s->rotor[ir];
    const size_t n = (size_t)r->nb * r->nc * r->ns, nsec = (size_t)r->nb * r->ns;
    memcpy(c->pan[ir], panel, 15 * n * sizeof(double));
    memcpy(c->sec[ir], section, 16 * nsec * sizeof(double));
    memcpy(c->bla[ir], blade, 9 * (size_t)r->nb * sizeof(double));
    return VLM_OK;
}

/* the term sum at every collocation point of rotor ir, AoS */
static int sh_cp_terms(vlm_ctx *c, int ir, int nterms, const int *irs, const int *sets,
                       const int *flags, double *v)
{
    const rotor_t *r = &c->s->rotor[ir];
    const size_t n = (size_t)r->nb * r->nc * r->ns;
    double *cp = (double *)malloc(3 * n * sizeof(double));
    for (size_t e = 0; e < n; ++e) memcpy(cp + 3 * e, c->pan[ir] + 15 * e, 3 * sizeof(double));
    const int rc = vlm_induced_velocity_terms(c, nterms, irs, sets, flags, (int64_t)n, cp, v, 0);
    free(cp);
    return rc;
}

int vlm_rotor_rhs(vlm_ctx *c, int ir, int nterms, const int *irs, const int *sets,
                  const int *flags, const double *kin, double *rhs)
{
    ROTOR_OK(c, ir);
    const rotor_t *r = &c->s->rotor[ir];
    const size_t n = (size_t)r->nb * r->nc * r->ns, per = (size_t)r->nc * r->ns;
    const size_t nconv = per * r->nbConvect;
    double *v = (double *)malloc(3 * n * sizeof(double));
    int rc = sh_cp_terms(c, ir, nterms, irs, sets, flags, v);
    if (rc != VLM_OK) {
        free(v);
        return rc;
    }
    for (size_t e = 0; e < nconv; ++e) {
        const double *P = c->pan[ir] + 15 * e;
        double r1[3], r2[3], w[3], c1[3], c2[3], vv[3];
        for (int i = 0; i < 3; ++i) {
            r1[i] = P[i] - kin[6 + i];
            r2[i] = P[i] - r->hub[i];
            w[i] = kin[9] * r->shaft[i];
        }
        sh_cross(kin + 3, r1, c1);
        sh_cross(w, r2, c2);
        for (int i = 0; i < 3; ++i) {
            const double vm = -1.0 * kin[i] - c1[i] - c2[i];
            c->vcp[ir][6 * e + 3 + i] = vm;
            vv[i] = vm + v[3 * e + i];
            c->vcp[ir][6 * e + i] = vv[i];
        }
        c->rhs[ir][e] = -sh_dot(vv, P + 3);
    }
    for (size_t e = nconv; e < n; ++e) c->rhs[ir][e] = c->rhs[ir][e % per];
    if (rhs) memcpy(rhs, c->rhs[ir], n * sizeof(double));
    free(v);
    return VLM_OK;
}

int vlm_rotor_mirror_velcp(vlm_ctx *c, int ir)
{
    ROTOR_OK(c, ir);
    if (c->image_of[ir] < 0) return shim_fail(c, "rotor is not an image");
    const rotor_t *r = &c->s->rotor[ir];
    const size_t n = (size_t)r->nb * r->nc * r->ns;
    const double *src = c->vcp[c->image_of[ir]];
    for (size_t e = 0; e < 6 * n; ++e)
        c->vcp[ir][e] = ((int)(e % 6) % 3 == c->image_axis[ir]) ? -src[e] : src[e];
    return VLM_OK;
}

int vlm_rotor_loads(vlm_ctx *c, int ir, int nterms, const int *irs, const int *sets,
                    const int *flags, double density, double dt, double omegaSign,
                    int spanwiseLiftSwitch, double *out)
{
    ROTOR_OK(c, ir);
    const rotor_t *r = &c->s->rotor[ir];
    const int nb = r->nb, nbc = r->nbConvect, nc = r->nc, ns = r->ns;
    const size_t n = (size_t)nb * nc * ns;
    const double sgn = copysign(1.0, omegaSign), inv = -1.0 * sgn;
    double *vld = (double *)malloc(3 * n * sizeof(double));
    int rc = sh_cp_terms(c, ir, nterms, irs, sets, flags, vld);
    if (rc != VLM_OK) {
        free(vld);
        return rc;
    }
    const double *pan = c->pan[ir], *vcp = c->vcp[ir];
    double *gprev = c->gprev[ir];
    double *fi_s = (double *)calloc(6 * (size_t)nbc * ns, sizeof(double));
    double *bs = (double *)malloc(9 * (size_t)nbc * sizeof(double));
    for (int t = 0; t < nbc * ns; ++t) {
        const int b = t / ns, is = t - b * ns;
        const double *S = c->sec[ir] + 16 * ((size_t)b * ns + is);
        const double *B = c->bla[ir] + 9 * (size_t)b;
        const size_t e0 = ((size_t)b * ns + is) * nc;
        double d0[3];
        for (int i = 0; i < 3; ++i) d0[i] = S[i] - S[3 + i];
        const double xq = sh_dot(d0, S + 6);
        double sx[5] = {(double)nc, 0, 0, 0, 0}, by[3] = {0, 0, 0}, bxy[3] = {0, 0, 0},
               bx2y[3] = {0, 0, 0};
        for (int ic = 0; ic < nc; ++ic) {
            const size_t e = e0 + ic;
            const double *P = pan + 15 * e;
            double vt[3], crv[3], dp[3];
            for (int i = 0; i < 3; ++i) {
                vt[i] = vcp[6 * e + i] + vld[3 * e + i];
                dp[i] = P[i] - S[3 + i];
            }
            sh_noproj(vt, P + 9, crv);
            const double x = sh_dot(dp, S + 6), x2 = x * x;
            sx[1] += x;
            sx[2] += x2;
            sx[3] += x2 * x;
            sx[4] += x2 * x2;
            for (int i = 0; i < 3; ++i) {
                by[i] += crv[i];
                bxy[i] += crv[i] * x;
                bx2y[i] += crv[i] * x2;
            }
        }
        double crv[3];
        for (int i = 0; i < 3; ++i)
            crv[i] = nc >= 3 ? sh_lsq2(xq, sx, by[i], bxy[i], bx2y[i]) : by[i] / nc;
        const double alpha = atan2(sh_dot(crv, S + 12), sh_dot(crv, S + 6));
        const double *vm1 = vcp + 6 * e0 + 3;
        double w[3], vi[3];
        sh_noproj(vm1, S + 9, w);
        const double phi = sh_angle(crv, w);
        for (int i = 0; i < 3; ++i) vi[i] = crv[i] - vm1[i];
        sh_noproj(vi, S + 9, w);
        w[2] = w[2] - vi[2];
        const double viz = vi[2], vix = sqrt(sh_dot(w, w));
        double mtc[3] = {-1.0 * S[6], -1.0 * S[7], -1.0 * S[8]};
        const double theta = OR_PI * 0.5 - sh_angle(mtc, B + 6);
        double ddir[3], ldir[3], cr[3];
        sh_unit(crv, ddir);
        sh_cross(ddir, B + 3, cr);
        sh_unit(cr, ldir);
        for (int i = 0; i < 3; ++i) ldir[i] = sgn * ldir[i];
        double fi[3] = {0, 0, 0}, lift[3] = {0, 0, 0}, liftu[3] = {0, 0, 0};
        for (int ic = 0; ic < nc; ++ic) {
            const size_t e = e0 + ic;
            const double *P = pan + 15 * e;
            const double *vc = vcp + 6 * e;
            const ring_t *W = r->blade[b].wiP;
            const double g = W[WI(r, ic, is)].gam;
            const double gm = ic > 0 ? W[WI(r, ic - 1, is)].gam : 0.0;
            const double gs = is > 0 ? W[WI(r, ic, is - 1)].gam : 0.0;
            const double gec = inv * (ic == 0 ? g : g - gm);
            const double ges = inv * (is == 0 ? g : g - gs);
            const double vtc = sh_dot(vc, P + 6);
            const double trapz = ic > 0 ? inv * 0.5 * (g + gm) : inv * 0.5 * g;
            const double dpu = density * (trapz - gprev[e]) / dt;
            double dpr = dpu + density * vtc * gec / P[13];
            if (spanwiseLiftSwitch) dpr = dpr + density * sh_dot(vc, P + 9) * ges / P[14];
            gprev[e] = trapz;
            double nf[3], nfu[3], q[3];
            for (int i = 0; i < 3; ++i) {
                nf[i] = dpr * P[12] * P[3 + i];
                nfu[i] = dpu * P[12] * P[3 + i];
                fi[i] += nf[i];
            }
            sh_proj(nf, ldir, q);
            for (int i = 0; i < 3; ++i) lift[i] += q[i];
            sh_proj(nfu, ldir, q);
            for (int i = 0; i < 3; ++i) liftu[i] += q[i];
        }
        double lin[3], lo[3];
        sh_proj(lift, B, lin);
        for (int i = 0; i < 3; ++i) lin[i] = inv * lin[i];
        sh_proj(lift, B + 6, lo);
        const double vmag = sqrt(sh_dot(crv, crv));
        const double q = 0.5 * density * (vmag * vmag);
        double cl = 0.0, clu = 0.0;
        if (fabs(q) > OR_EPS) {
            const double sg = copysign(1.0, sh_dot(lift, B + 6));
            cl = sqrt(sh_dot(lift, lift)) * sg / (q * S[15]);
            clu = sqrt(sh_dot(liftu, liftu)) * sg / (q * S[15]);
        }
        double *o = out + 19 * (size_t)t;
        for (int i = 0; i < 3; ++i) {
            o[i] = crv[i];
            o[8 + i] = lift[i];
            o[11 + i] = lin[i];
            o[14 + i] = lo[i];
            fi_s[6 * t + i] = fi[i];
            fi_s[6 * t + 3 + i] = liftu[i];
        }
        o[3] = alpha;
        o[4] = phi;
        o[5] = viz;
        o[6] = vix;
        o[7] = theta;
        o[17] = cl;
        o[18] = clu;
    }
    for (int b = 0; b < nbc; ++b) {
        double sm[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (int is = 0; is < ns; ++is) {
            const int t = b * ns + is;
            for (int i = 0; i < 3; ++i) {
                sm[i] += fi_s[6 * t + i];
                sm[3 + i] += out[19 * (size_t)t + 8 + i];
                sm[6 + i] += fi_s[6 * t + 3 + i];
            }
        }
        for (int k = 0; k < 9; ++k) bs[9 * b + k] = sm[k];
    }
    double rr[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (nbc == 1 && nb > 1)
        for (int k = 0; k < 9; ++k) rr[k] = (double)nb * bs[k];
    else
        for (int b = 0; b < nbc; ++b)
            for (int k = 0; k < 9; ++k) rr[k] = rr[k] + bs[9 * b + k];
    memcpy(out + 19 * (size_t)nbc * ns, rr, 9 * sizeof(double));
    free(bs);
    free(fi_s);
    free(vld);
    return VLM_OK;
}
