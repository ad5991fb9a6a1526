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
    memcpy(gam, rhs, sizeof(double) * (size_t)n);
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
    const size_t nw = (size_t)r->nNwake * r->ns, nf = (size_t)(r->nFwake > 0 ? r->nFwake : 1);
    const size_t nvn = 3u * (size_t)r->nNwake * (r->ns + 1), nvf = 3u * nf;
    return 4 * sizeof(int) + (size_t)r->nb * (2 * nw * sizeof(ring_t) + 2 * nf * sizeof(fil_t) +
                                              4 * (nvn + nvf) * sizeof(double) +
                                              2 * OR_NPF * sizeof(fil_t) + 4 * sizeof(double));
}

static void state_io(rotor_t *r, char *buf, int out)
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
    state_io(&c->s->rotor[ir], (char *)buf, 1);
    return VLM_OK;
}

int vlm_rotor_set_state(vlm_ctx *c, int ir, const void *buf, int64_t bytes)
{
    ROTOR_OK(c, ir);
    if ((size_t)bytes != state_size(&c->s->rotor[ir])) return shim_fail(c, "state size");
    state_io(&c->s->rotor[ir], (char *)buf, 0);
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
