/*
 * CPU oracle for the VOLCANOR free-wake hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C FP64 restatement of the reference routines, in the reference's
 * ring form: every ring owns four filaments (vf1 c1->c2, vf2 c2->c3,
 * vf3 c3->c4, vf4 c4->c1), the near wake is shifted by row copies, and the
 * velocity history is indexed by row exactly as the Fortran arrays are.
 * The HIP library stores the same state as a node lattice with unique
 * segments; comparing the two checks that re-layout as well as the
 * arithmetic.  Nothing in the product path calls this file.
 *
 * Pins: closed forms and reference fixtures in tests/ (see
 * tests/test_oracle.py), and per-node velocities dumped by oracle/_ref
 * (the reference compiled from its own sources) in tests/golden/.
 */
#include "vlm_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* libMath.f90:9-11 */
static const double OR_PI = 3.141592653589793238462643383279502884;
static const double OR_EPS = DBL_EPSILON;

/* ------------------------------------------------------------------ */
/* a1: vf_vind, classdef.f90:476-503; unitVec, libMath.f90:249-262     */
/* ------------------------------------------------------------------ */
static void unit_vec(const double a[3], double u[3])
{
    double n = sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
    if (n > OR_EPS) {
        u[0] = a[0] / n;
        u[1] = a[1] / n;
        u[2] = a[2] / n;
    } else {
        u[0] = u[1] = u[2] = 0.0;
    }
}

void or_vf_vind(const double P[3], const double A[3], const double B[3],
                double rc, double v[3])
{
    double r1[3], r2[3], r0[3], c[3], u1[3], u2[3];
    for (int d = 0; d < 3; ++d) {
        r1[d] = P[d] - A[d];
        r2[d] = P[d] - B[d];
        r0[d] = r1[d] - r2[d];
    }
    c[0] = r1[1] * r2[2] - r1[2] * r2[1];
    c[1] = r1[2] * r2[0] - r1[0] * r2[2];
    c[2] = r1[0] * r2[1] - r1[1] * r2[0];
    double c2 = c[0] * c[0] + c[1] * c[1] + c[2] * c[2];
    v[0] = v[1] = v[2] = 0.0;
    if (c2 > OR_EPS * OR_EPS) {
        unit_vec(r1, u1);
        unit_vec(r2, u2);
        double dot = r0[0] * (u1[0] - u2[0]) + r0[1] * (u1[1] - u2[1]) +
                     r0[2] * (u1[2] - u2[2]);
        double r0n = sqrt(r0[0] * r0[0] + r0[1] * r0[1] + r0[2] * r0[2]);
        double q = rc * r0n;
        double den = sqrt(q * q * q * q + c2 * c2);
        double s = (0.25 / OR_PI) * dot / den;
        v[0] = c[0] * s;
        v[1] = c[1] * s;
        v[2] = c[2] * s;
    }
}

void or_filament_sum(int64_t nfil, const double *a, const double *b,
                     const double *gam, const double *rc, int64_t m,
                     const double *p, double *v)
{
    for (int64_t t = 0; t < m; ++t) {
        double acc[3] = {0, 0, 0}, w[3];
        for (int64_t s = 0; s < nfil; ++s) {
            or_vf_vind(p + 3 * t, a + 3 * s, b + 3 * s, rc[s], w);
            acc[0] += w[0] * gam[s];
            acc[1] += w[1] * gam[s];
            acc[2] += w[2] * gam[s];
        }
        v[3 * t + 0] = acc[0];
        v[3 * t + 1] = acc[1];
        v[3 * t + 2] = acc[2];
    }
}

void or_filament_sum_omp(int64_t nfil, const double *a, const double *b,
                         const double *gam, const double *rc, int64_t m,
                         const double *p, double *v, int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(static)
#else
    (void)nthreads;
#endif
    for (int64_t t = 0; t < m; ++t) {
        double acc[3] = {0, 0, 0}, w[3];
        for (int64_t s = 0; s < nfil; ++s) {
            or_vf_vind(p + 3 * t, a + 3 * s, b + 3 * s, rc[s], w);
            acc[0] += w[0] * gam[s];
            acc[1] += w[1] * gam[s];
            acc[2] += w[2] * gam[s];
        }
        v[3 * t + 0] = acc[0];
        v[3 * t + 1] = acc[1];
        v[3 * t + 2] = acc[2];
    }
}

/* ------------------------------------------------------------------ */
/* Object model in ring form (classdef.f90:57-467)                     */
/* ------------------------------------------------------------------ */
typedef struct {
    double c[4][3];    /* corners 1..4 */
    double gam;
    double rc[4];      /* vf(1..4)%rVc */
    double rc0[4];
    double l0[4], lc[4];
    double age;        /* all four filaments age together (classdef.f90:4331-4354) */
} ring_t;

typedef struct {
    double fc[2][3];   /* fc(:,1) TE end, fc(:,2) LE end */
    double gam, rc, rc0, l0, lc, age;
} fil_t;

typedef struct {
    ring_t *wiP;                  /* nc*ns: (ic,is) -> ic + nc*is */
    ring_t *waN, *waNP;           /* nNwake*ns: (i,j) -> i + nNwake*j */
    fil_t *waF, *waFP;            /* nFwake */
    double *velN, *velN1, *velNP, *velNS;   /* 3*nNwake*(ns+1) */
    double *velF, *velF1, *velFP, *velFS;   /* 3*nFwake */
} blade_t;

#define OR_NPF 240  /* prescribed far-wake filaments per blade, classdef.f90:225 */

typedef struct {
    int nb, nc, ns, nNwake, nFwake, nbConvect;
    int rollupStart, rollupEnd, rollupMin, suppressFwake;
    int rowNear, rowFar;
    double shaft[3], hub[3];
    blade_t *blade;
    /* prescribed far wake (pFwake_class, classdef.f90:222-236): per wake
     * type (0 'C', 1 'P') the blades' OR_NPF filaments and their relaxed
     * helix pitch and radius; the step number of the last begin_step */
    int presc_nt, presc_gen_nt, iter;
    double presc_dpsi;
    int presc_present[2];
    fil_t *pf[2];            /* nb * OR_NPF */
    double *presc_pitch[2];  /* nb */
    double *presc_radius[2]; /* nb */
    /* image rotor (surfaceType -1): the rotor it mirrors (-1: none) and
     * the negated coordinate (0 x, 1 y, 2 z) */
    int image_of, image_axis;
} rotor_t;

#define OR_MAX_ROTORS 16
struct or_sys {
    int nr;
    rotor_t rotor[OR_MAX_ROTORS];
};

#define WI(r, ic, is) ((ic) + (r)->nc * (is))
#define WN(r, i, j) ((i) + (r)->nNwake * (j))
#define VN(r, d, i, j) ((d) + 3 * ((i) + (r)->nNwake * (j)))

or_sys *or_sys_create(void)
{
    return (or_sys *)calloc(1, sizeof(or_sys));
}

static void free_blade(blade_t *b)
{
    free(b->wiP); free(b->waN); free(b->waNP); free(b->waF); free(b->waFP);
    free(b->velN); free(b->velN1); free(b->velNP); free(b->velNS);
    free(b->velF); free(b->velF1); free(b->velFP); free(b->velFS);
}

void or_sys_destroy(or_sys *s)
{
    if (!s) return;
    for (int ir = 0; ir < s->nr; ++ir) {
        for (int b = 0; b < s->rotor[ir].nb; ++b) free_blade(&s->rotor[ir].blade[b]);
        free(s->rotor[ir].blade);
        for (int t = 0; t < 2; ++t) {
            free(s->rotor[ir].pf[t]);
            free(s->rotor[ir].presc_pitch[t]);
            free(s->rotor[ir].presc_radius[t]);
        }
    }
    free(s);
}

int or_rotor_create(or_sys *s, int nb, int nc, int ns, int nNwake, int nFwake,
                    int nbConvect, int rollupStart, int rollupEnd,
                    int rollup_min, int suppressFwake)
{
    if (s->nr >= OR_MAX_ROTORS) return -1;
    rotor_t *r = &s->rotor[s->nr];
    memset(r, 0, sizeof(*r));
    r->nb = nb; r->nc = nc; r->ns = ns; r->nNwake = nNwake; r->nFwake = nFwake;
    r->nbConvect = nbConvect;
    r->rollupStart = rollupStart; r->rollupEnd = rollupEnd;
    r->rollupMin = rollup_min; r->suppressFwake = suppressFwake;
    r->shaft[2] = 1.0;
    r->image_of = -1;
    r->blade = (blade_t *)calloc((size_t)nb, sizeof(blade_t));
    size_t nvn = 3u * (size_t)nNwake * (size_t)(ns + 1);
    size_t nvf = 3u * (size_t)(nFwake > 0 ? nFwake : 1);
    for (int b = 0; b < nb; ++b) {
        blade_t *bl = &r->blade[b];
        bl->wiP = (ring_t *)calloc((size_t)nc * ns, sizeof(ring_t));
        bl->waN = (ring_t *)calloc((size_t)nNwake * ns, sizeof(ring_t));
        bl->waNP = (ring_t *)calloc((size_t)nNwake * ns, sizeof(ring_t));
        bl->waF = (fil_t *)calloc((size_t)(nFwake > 0 ? nFwake : 1), sizeof(fil_t));
        bl->waFP = (fil_t *)calloc((size_t)(nFwake > 0 ? nFwake : 1), sizeof(fil_t));
        bl->velN = (double *)calloc(nvn, sizeof(double));
        bl->velN1 = (double *)calloc(nvn, sizeof(double));
        bl->velNP = (double *)calloc(nvn, sizeof(double));
        bl->velNS = (double *)calloc(nvn, sizeof(double));
        bl->velF = (double *)calloc(nvf, sizeof(double));
        bl->velF1 = (double *)calloc(nvf, sizeof(double));
        bl->velFP = (double *)calloc(nvf, sizeof(double));
        bl->velFS = (double *)calloc(nvf, sizeof(double));
    }
    for (int t = 0; t < 2; ++t) {
        r->pf[t] = (fil_t *)calloc((size_t)nb * OR_NPF, sizeof(fil_t));
        r->presc_pitch[t] = (double *)calloc((size_t)nb, sizeof(double));
        r->presc_radius[t] = (double *)calloc((size_t)nb, sizeof(double));
    }
    return s->nr++;
}

void or_rotor_set_wing(or_sys *s, int ir, const double *corners,
                       const double *rc, const double *gam)
{
    rotor_t *r = &s->rotor[ir];
    for (int b = 0; b < r->nb; ++b)
        for (int is = 0; is < r->ns; ++is)
            for (int ic = 0; ic < r->nc; ++ic) {
                size_t k = ((size_t)b * r->ns + is) * r->nc + ic;
                ring_t *w = &r->blade[b].wiP[WI(r, ic, is)];
                if (corners) memcpy(w->c, corners + 12 * k, 12 * sizeof(double));
                if (rc) memcpy(w->rc, rc + 4 * k, 4 * sizeof(double));
                if (gam) w->gam = gam[k];
            }
}

void or_rotor_set_axis(or_sys *s, int ir, const double shaft[3],
                       const double hub[3])
{
    memcpy(s->rotor[ir].shaft, shaft, 3 * sizeof(double));
    memcpy(s->rotor[ir].hub, hub, 3 * sizeof(double));
}

/* Wake core radii at start-up, classdef.f90:3826-3850. */
void or_rotor_init_wake_cores(or_sys *s, int ir, double spanwiseCore,
                              const double *scv)
{
    rotor_t *r = &s->rotor[ir];
    for (int b = 0; b < r->nb; ++b) {
        for (int j = 0; j < r->ns; ++j)
            for (int i = 0; i < r->nNwake; ++i) {
                ring_t *w = &r->blade[b].waN[WN(r, i, j)];
                w->rc[1] = w->rc0[1] = spanwiseCore;
                w->rc[3] = w->rc0[3] = spanwiseCore;
                w->rc[0] = w->rc0[0] = scv[j];
                w->rc[2] = w->rc0[2] = scv[j + 1];
                w->gam = 0.0;
            }
        for (int i = 0; i < r->nFwake; ++i) {
            fil_t *f = &r->blade[b].waF[i];
            f->rc = f->rc0 = scv[r->ns];
            f->gam = 0.0;
        }
    }
}

/* main.f90:227-229 (0-based storage, 1-based cursor values kept). */
void or_rotor_init_cursors(or_sys *s, int ir)
{
    rotor_t *r = &s->rotor[ir];
    r->rowFar = r->nFwake + 1;
    r->rowNear = r->nNwake + 1;
}

/* main.f90:412-417 */
void or_rotor_begin_step(or_sys *s, int ir, int iter)
{
    rotor_t *r = &s->rotor[ir];
    r->iter = iter;
    r->rowNear = r->rowNear - 1 > 1 ? r->rowNear - 1 : 1;
    if (iter > r->nNwake) r->rowFar = r->rowFar - 1 > 1 ? r->rowFar - 1 : 1;
}

void or_rotor_cursors(const or_sys *s, int ir, int *rowNear, int *rowFar)
{
    *rowNear = s->rotor[ir].rowNear;
    *rowFar = s->rotor[ir].rowFar;
}

/* ------------------------------------------------------------------ */
/* Source sums, classdef.f90:527-542, 1342-1395, 1437-1513, 4424-4479 */
/* ------------------------------------------------------------------ */
static void ring_vind(const ring_t *w, const double P[3], double v[3])
{
    /* vr_vind: vf1 c1->c2, vf2 c2->c3, vf3 c3->c4, vf4 c4->c1 */
    double t[3];
    v[0] = v[1] = v[2] = 0.0;
    for (int f = 0; f < 4; ++f) {
        or_vf_vind(P, w->c[f], w->c[(f + 1) & 3], w->rc[f], t);
        v[0] += t[0]; v[1] += t[1]; v[2] += t[2];
    }
}

static void fil_vind(const ring_t *w, int f, const double P[3], double v[3])
{
    or_vf_vind(P, w->c[f], w->c[(f + 1) & 3], w->rc[f], v);
}

static void blade_vind_bywing(const rotor_t *r, const blade_t *bl,
                              const double P[3], double v[3])
{
    double t[3];
    for (int is = 0; is < r->ns; ++is)
        for (int ic = 0; ic < r->nc; ++ic) {
            const ring_t *w = &bl->wiP[WI(r, ic, is)];
            ring_vind(w, P, t);
            v[0] += t[0] * w->gam; v[1] += t[1] * w->gam; v[2] += t[2] * w->gam;
        }
}

/* classdef.f90:1376-1395: vf2 + vf4 of every ring, minus vf2 of the last row */
static void blade_vind_bound(const rotor_t *r, const blade_t *bl,
                             const double P[3], double v[3])
{
    double t2[3], t4[3];
    for (int is = 0; is < r->ns; ++is)
        for (int ic = 0; ic < r->nc; ++ic) {
            const ring_t *w = &bl->wiP[WI(r, ic, is)];
            fil_vind(w, 1, P, t2);
            fil_vind(w, 3, P, t4);
            for (int d = 0; d < 3; ++d) v[d] += (t2[d] + t4[d]) * w->gam;
        }
    for (int is = 0; is < r->ns; ++is) {
        const ring_t *w = &bl->wiP[WI(r, r->nc - 1, is)];
        fil_vind(w, 1, P, t2);
        for (int d = 0; d < 3; ++d) v[d] -= t2[d] * w->gam;
    }
}

static void blade_vind_bywake(const rotor_t *r, const blade_t *bl, int pred,
                              const double P[3], double v[3])
{
    const ring_t *N = pred ? bl->waNP : bl->waN;
    const fil_t *F = pred ? bl->waFP : bl->waF;
    double t[3];
    int nN = r->nNwake, nF = r->nFwake;
    for (int j = 0; j < r->ns; ++j)
        for (int i = r->rowNear - 1; i < nN; ++i) {
            const ring_t *w = &N[WN(r, i, j)];
            if (fabs(w->gam) > OR_EPS) {
                ring_vind(w, P, t);
                v[0] += t[0] * w->gam; v[1] += t[1] * w->gam; v[2] += t[2] * w->gam;
            }
        }
    if (r->rowFar <= nF) {
        /* open horseshoe: last near row loses its trailing filament */
        for (int j = 0; j < r->ns; ++j) {
            const ring_t *w = &N[WN(r, nN - 1, j)];
            fil_vind(w, 1, P, t);
            v[0] -= t[0] * w->gam; v[1] -= t[1] * w->gam; v[2] -= t[2] * w->gam;
        }
        for (int i = r->rowFar - 1; i < nF; ++i) {
            const fil_t *f = &F[i];
            if (fabs(f->gam) > OR_EPS) {
                or_vf_vind(P, f->fc[0], f->fc[1], f->rc, t);
                v[0] += t[0] * f->gam; v[1] += t[1] * f->gam; v[2] += t[2] * f->gam;
            }
        }
        /* prescribed far wake, classdef.f90:1471-1476 and 1501-1507 */
        const fil_t *W = r->pf[pred] + (size_t)(bl - r->blade) * OR_NPF;
        for (int i = 0; i < OR_NPF; ++i)
            if (fabs(W[i].gam) > OR_EPS) {
                or_vf_vind(P, W[i].fc[0], W[i].fc[1], W[i].rc, t);
                v[0] += t[0] * W[i].gam; v[1] += t[1] * W[i].gam; v[2] += t[2] * W[i].gam;
            }
    }
}

void or_rotor_vind(const or_sys *s, int ir, int set, int64_t m,
                   const double *p, double *v)
{
    const rotor_t *r = &s->rotor[ir];
    for (int64_t t = 0; t < m; ++t) {
        double acc[3] = {0, 0, 0};
        for (int b = 0; b < r->nb; ++b) {
            const blade_t *bl = &r->blade[b];
            if (set == 0) blade_vind_bywing(r, bl, p + 3 * t, acc);
            else if (set == 1) blade_vind_bound(r, bl, p + 3 * t, acc);
            else if (r->nNwake > 0) blade_vind_bywake(r, bl, set == 3, p + 3 * t, acc);
        }
        v[3 * t + 0] += acc[0];
        v[3 * t + 1] += acc[1];
        v[3 * t + 2] += acc[2];
    }
}

/* ------------------------------------------------------------------ */
/* Wake-node pass, libCommon.f90:114-211 and main.f90:800-841          */
/* ------------------------------------------------------------------ */
void or_wake_velocity(or_sys *s, int pred)
{
    for (int ir = 0; ir < s->nr; ++ir) {
        rotor_t *r = &s->rotor[ir];
        /* an image rotor's wake velocities are never used: it mirrors its
         * source instead of convecting (main.f90:851-856, 884-915) */
        if (r->nNwake <= 0 || r->image_of >= 0) continue;
        for (int b = 0; b < r->nbConvect; ++b) {
            blade_t *bl = &r->blade[b];
            double *VN_ = pred ? bl->velNP : bl->velN;
            double *VF_ = pred ? bl->velFP : bl->velF;
            const ring_t *N = pred ? bl->waNP : bl->waN;
            const fil_t *F = pred ? bl->waFP : bl->waF;
            int rn = r->rowNear - 1, rf = r->rowFar - 1;
            /* zero the active range (main.f90:806-807, 876-877) */
            for (int j = 0; j <= r->ns; ++j)
                for (int i = rn; i < r->nNwake; ++i)
                    for (int d = 0; d < 3; ++d) VN_[VN(r, d, i, j)] = 0.0;
            for (int i = rf; i < r->nFwake; ++i)
                for (int d = 0; d < 3; ++d) VF_[3 * i + d] = 0.0;
#pragma omp parallel for schedule(static) collapse(2)
            for (int j = 0; j <= r->ns; ++j)
                for (int i = rn; i < r->nNwake; ++i) {
                    const double *P = (j < r->ns) ? N[WN(r, i, j)].c[1]
                                                  : N[WN(r, i, r->ns - 1)].c[2];
                    for (int jr = 0; jr < s->nr; ++jr) {
                        const rotor_t *q = &s->rotor[jr];
                        double acc[3] = {0, 0, 0};
                        for (int qb = 0; qb < q->nb; ++qb)
                            blade_vind_bywing(q, &q->blade[qb], P, acc);
                        if (q->nNwake > 0)
                            for (int qb = 0; qb < q->nb; ++qb)
                                blade_vind_bywake(q, &q->blade[qb], pred, P, acc);
                        for (int d = 0; d < 3; ++d) VN_[VN(r, d, i, j)] += acc[d];
                    }
                }
#pragma omp parallel for schedule(static)
            for (int i = rf; i < r->nFwake; ++i) {
                const double *P = F[i].fc[0];
                for (int jr = 0; jr < s->nr; ++jr) {
                    const rotor_t *q = &s->rotor[jr];
                    double acc[3] = {0, 0, 0};
                    for (int qb = 0; qb < q->nb; ++qb)
                        blade_vind_bywing(q, &q->blade[qb], P, acc);
                    if (q->nNwake > 0)
                        for (int qb = 0; qb < q->nb; ++qb)
                            blade_vind_bywake(q, &q->blade[qb], pred, P, acc);
                    for (int d = 0; d < 3; ++d) VF_[3 * i + d] += acc[d];
                }
            }
        }
    }
}

/* ------------------------------------------------------------------ */
/* Wake bookkeeping                                                    */
/* ------------------------------------------------------------------ */
static double norm3(const double a[3])
{
    return sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
}

/* vf_calclength, classdef.f90:505-515 */
static void ring_calclength(ring_t *w, int original)
{
    for (int f = 0; f < 4; ++f) {
        double d[3];
        for (int k = 0; k < 3; ++k) d[k] = w->c[f][k] - w->c[(f + 1) & 3][k];
        w->lc[f] = norm3(d);
        if (original) w->l0[f] = w->lc[f];
    }
}

/* rotor_assignshed, classdef.f90:4297-4325 */
void or_rotor_assignshed(or_sys *s, int ir, int edge_te)
{
    rotor_t *r = &s->rotor[ir];
    for (int b = 0; b < r->nb; ++b) {
        blade_t *bl = &r->blade[b];
        for (int j = 0; j < r->ns; ++j) {
            const ring_t *w = &bl->wiP[WI(r, r->nc - 1, j)];
            if (!edge_te) {
                ring_t *x = &bl->waN[WN(r, r->rowNear - 1, j)];
                memcpy(x->c[0], w->c[1], 3 * sizeof(double));
                memcpy(x->c[3], w->c[2], 3 * sizeof(double));
                ring_calclength(x, 1);
                x->gam = w->gam;
            } else {
                int row = (r->rowNear - 1 > 1 ? r->rowNear - 1 : 1) - 1;
                ring_t *x = &bl->waN[WN(r, row, j)];
                memcpy(x->c[1], w->c[1], 3 * sizeof(double));
                memcpy(x->c[2], w->c[2], 3 * sizeof(double));
            }
        }
    }
}

/* rotor_age_wake, classdef.f90:4331-4354 (ageAzimuthal is output-only) */
void or_rotor_age_wake(or_sys *s, int ir, double dt, double omegaSlow)
{
    (void)omegaSlow;
    rotor_t *r = &s->rotor[ir];
    for (int b = 0; b < r->nb; ++b) {
        blade_t *bl = &r->blade[b];
        for (int j = 0; j < r->ns; ++j)
            for (int i = r->rowNear - 1; i < r->nNwake; ++i) bl->waN[WN(r, i, j)].age += dt;
        for (int i = r->rowFar - 1; i < r->nFwake; ++i) bl->waF[i].age += dt;
    }
}

/* rotor_dissipate_wake, classdef.f90:4356-4408 */
void or_rotor_dissipate_wake(or_sys *s, int ir, double dt, double kinVisc,
                             double apparentViscCoeff, double decayCoeff)
{
    rotor_t *r = &s->rotor[ir];
    const double oseen = 1.2564;
    double inc = 4.0 * oseen * apparentViscCoeff * kinVisc * dt;
    double decay = exp(-decayCoeff * dt);
    for (int b = 0; b < r->nb; ++b) {
        blade_t *bl = &r->blade[b];
        for (int j = 0; j < r->ns; ++j) {
            for (int i = r->rowNear - 1; i < r->nNwake; ++i) {
                ring_t *w = &bl->waN[WN(r, i, j)];
                w->rc[0] = sqrt(w->rc[0] * w->rc[0] + inc);
                w->rc[2] = w->rc[0];
                w->gam = w->gam * decay;
            }
            for (int i = r->rowNear - 1; i < r->nNwake; ++i) {
                ring_t *w = &bl->waN[WN(r, i, j)];
                w->rc[1] = sqrt(w->rc[1] * w->rc[1] + inc);
            }
            if (r->rowNear != r->nNwake)
                for (int i = r->rowNear; i < r->nNwake; ++i)
                    bl->waN[WN(r, i, j)].rc[3] = bl->waN[WN(r, i - 1, j)].rc[1];
        }
        for (int i = r->rowFar - 1; i < r->nFwake; ++i) {
            fil_t *f = &bl->waF[i];
            f->rc = sqrt(f->rc * f->rc + inc);
            f->gam = f->gam * decay;
        }
    }
}

/* rotor_strain_wake, classdef.f90:4410-4422 with vf_strain 517-521 */
void or_rotor_strain_wake(or_sys *s, int ir)
{
    rotor_t *r = &s->rotor[ir];
    for (int b = 0; b < r->nb; ++b)
        for (int i = r->rowFar - 1; i < r->nFwake; ++i) {
            fil_t *f = &r->blade[b].waF[i];
            double d[3];
            for (int k = 0; k < 3; ++k) d[k] = f->fc[0][k] - f->fc[1][k];
            f->lc = norm3(d);
            f->rc = f->rc0 * sqrt(f->l0 / f->lc);
        }
}

/* rotor_shiftwake / rotor_shiftFwake, classdef.f90:4481-4513 */
static void shiftwake(rotor_t *r)
{
    for (int b = 0; b < r->nb; ++b) {
        blade_t *bl = &r->blade[b];
        for (int i = r->nNwake - 1; i >= 1; --i)
            for (int j = 0; j < r->ns; ++j) bl->waN[WN(r, i, j)] = bl->waN[WN(r, i - 1, j)];
        for (int j = 0; j < r->ns; ++j) bl->waN[WN(r, 0, j)].age = 0.0;
    }
}

static void shiftFwake(rotor_t *r)
{
    for (int b = 0; b < r->nb; ++b) {
        blade_t *bl = &r->blade[b];
        for (int i = r->nFwake - 1; i >= 1; --i) bl->waF[i] = bl->waF[i - 1];
        bl->waF[0].age = 0.0;
    }
}

/* rotor_rollup, classdef.f90:4515-4605 */
void or_rotor_rollup(or_sys *s, int ir)
{
    rotor_t *r = &s->rotor[ir];
    int rowFarNext = r->rowFar - 1;  /* 1-based */
    int last = r->nNwake - 1;
    for (int b = 0; b < r->nb; ++b) {
        blade_t *bl = &r->blade[b];
        double gamRollup = bl->waN[WN(r, last, r->ns - 1)].gam;
        double cLE[3] = {0, 0, 0}, cTE[3] = {0, 0, 0};
        double radius = 0.0, gamSum = 0.0;
        for (int is = r->rollupStart - 1; is <= r->rollupEnd - 1; ++is) {
            const ring_t *w = &bl->waN[WN(r, last, is)];
            for (int k = 0; k < 3; ++k) {
                cLE[k] += w->c[3][k] * w->gam;   /* vf(4)%fc(:,1) = corner 4 */
                cTE[k] += w->c[2][k] * w->gam;   /* vf(3)%fc(:,1) = corner 3 */
            }
            gamSum += w->gam;
            if (r->rollupMin) {
                if (w->gam < gamRollup) gamRollup = w->gam;
            } else {
                if (w->gam > gamRollup) gamRollup = w->gam;
            }
            radius += w->rc[2] * w->gam;
        }
        double ageRollup = bl->waN[WN(r, last, r->ns - 1)].age;
        if (fabs(gamSum) > OR_EPS) {
            for (int k = 0; k < 3; ++k) { cLE[k] /= gamSum; cTE[k] /= gamSum; }
            radius /= gamSum;
        } else {
            const ring_t *w = &bl->waN[WN(r, last, r->rollupEnd - 1)];
            memcpy(cLE, w->c[1], sizeof cLE);   /* vf(2)%fc(:,1) = corner 2 */
            memcpy(cTE, w->c[2], sizeof cTE);
            radius = w->rc[2];
        }
        if (r->suppressFwake) gamRollup = 0.0;
        if (r->nFwake > 0) {
            if (rowFarNext == 0) {
                /* shiftFwake shifts every blade; done once, for blade 1 */
                if (b == 0) shiftFwake(r);
                rowFarNext = 1;
            }
            fil_t *f = &bl->waF[rowFarNext - 1];
            memcpy(f->fc[1], cLE, sizeof cLE);
            memcpy(f->fc[0], cTE, sizeof cTE);
            f->gam = gamRollup;
            f->age = ageRollup;
            f->rc0 = radius;
            f->rc = radius;
            double d[3];
            for (int k = 0; k < 3; ++k) d[k] = f->fc[0][k] - f->fc[1][k];
            f->lc = f->l0 = norm3(d);
            if (rowFarNext < r->nFwake) memcpy(bl->waF[rowFarNext].fc[1], cTE, sizeof cTE);
        }
    }
    shiftwake(r);
}

/* getTransformAxis, libMath.f90:695-726 (column-major T, x' = T x) */
static void transform_axis(double theta, const double axv[3], double T[9])
{
    double n = norm3(axv);
    double a[3] = {axv[0] / n, axv[1] / n, axv[2] / n};
    double ct = cos(theta), st = sin(theta), om = 1.0 - ct;
    T[0] = ct + a[0] * a[0] * om;
    T[1] = a[2] * st + a[1] * a[0] * om;
    T[2] = -a[1] * st + a[2] * a[0] * om;
    T[3] = -a[2] * st + a[0] * a[1] * om;
    T[4] = ct + a[1] * a[1] * om;
    T[5] = a[0] * st + a[2] * a[1] * om;
    T[6] = a[1] * st + a[0] * a[2] * om;
    T[7] = -a[0] * st + a[1] * a[2] * om;
    T[8] = ct + a[2] * a[2] * om;
}

static void rot_point(const double T[9], const double o[3], double p[3])
{
    double q[3] = {p[0] - o[0], p[1] - o[1], p[2] - o[2]};
    for (int i = 0; i < 3; ++i)
        p[i] = T[i] * q[0] + T[i + 3] * q[1] + T[i + 6] * q[2] + o[i];
}

/* rotor_updatePrescribedWake (classdef.f90:5172-5217) with pFwake_update
 * (998-1066): per convected blade, a 10-revolution helix of OR_NPF
 * filaments anchored at the TE end of the oldest far filament, its radius
 * the mean distance from the z axis of the far filaments' TE ends (from
 * rowFar, or the last presc_gen_nt + 1 of them) and its pitch their mean
 * z step per revolution, each relaxed by half against the previous update
 * of the same wake type.  The helix winds clockwise whatever the rotor
 * (isClockwiseRotor is never set, classdef.f90:231, 1048).  Every filament
 * takes the oldest far filament's circulation and core radius. */
static void presc_update(rotor_t *r, int pred)
{
    const int nF = r->nFwake;
    const int start = r->presc_gen_nt == 0 ? r->rowFar - 1 : nF - r->presc_gen_nt - 1;
    const int n = nF - start;
    if (start < 0 || n < 1) return;  /* no far filament to anchor to */
    const double twoPi = 2.0 * OR_PI;
    const double dx = (twoPi * 10.0 - 0.0) / (OR_NPF + 1 - 1);
    for (int b = 0; b < r->nbConvect; ++b) {
        const fil_t *F = (pred ? r->blade[b].waFP : r->blade[b].waF) + start;
        const double *anchor = F[n - 1].fc[0];
        double pitch = 0.0, radius = 0.0;
        for (int i = 0; i < n; ++i) {
            radius += sqrt(F[i].fc[0][1] * F[i].fc[0][1] + F[i].fc[0][0] * F[i].fc[0][0]);
            if (i < n - 1) pitch += F[i].fc[0][2] - F[i + 1].fc[0][2];
        }
        pitch = fabs(pitch) * (-twoPi / r->presc_dpsi) / (n - 1);
        radius = radius / n;
        double *hp = &r->presc_pitch[pred][b], *hr = &r->presc_radius[pred][b];
        *hp = 0.5 * pitch + (1 - 0.5) * *hp;
        *hr = 0.5 * radius + (1 - 0.5) * *hr;
        const double dth = atan2(anchor[1], anchor[0]);
        const double dz = anchor[2] - r->hub[2];
        double c[OR_NPF + 1][3];
        for (int k = 0; k <= OR_NPF; ++k) {
            const double th = -(k * dx + 0.0);
            c[k][0] = *hr * cos(th + dth);
            c[k][1] = *hr * sin(th + dth);
            c[k][2] = *hp * fabs(th) / twoPi + dz;
        }
        fil_t *W = r->pf[pred] + (size_t)b * OR_NPF;
        for (int i = 0; i < OR_NPF; ++i)
            for (int d = 0; d < 3; ++d) {
                W[i].fc[1][d] = r->hub[d] + c[i][d];
                W[i].fc[0][d] = r->hub[d] + c[i + 1][d];
            }
        memcpy(W[0].fc[1], anchor, 3 * sizeof(double));
        for (int i = 0; i < OR_NPF; ++i) {
            W[i].gam = F[n - 1].gam;
            W[i].rc = F[n - 1].rc;
        }
    }
    r->presc_present[pred] = 1;
    if (r->nbConvect == 1 && r->nb > 1)
        for (int b = 1; b < r->nb; ++b) {
            fil_t *W = r->pf[pred] + (size_t)b * OR_NPF;
            memcpy(W, r->pf[pred], OR_NPF * sizeof(fil_t));
            r->presc_pitch[pred][b] = r->presc_pitch[pred][0];
            r->presc_radius[pred][b] = r->presc_radius[pred][0];
            const double theta = 2.0 * OR_PI / r->nb * b;
            if (fabs(theta) > OR_EPS) {
                double T[9];
                transform_axis(theta, r->shaft, T);
                for (int i = 0; i < OR_NPF; ++i) {
                    rot_point(T, r->hub, W[i].fc[0]);
                    rot_point(T, r->hub, W[i].fc[1]);
                }
            }
        }
}

void or_rotor_set_prescribed(or_sys *s, int ir, int presc_nt, int presc_gen_nt, double dpsi)
{
    rotor_t *r = &s->rotor[ir];
    r->presc_nt = presc_nt;
    r->presc_gen_nt = presc_gen_nt;
    r->presc_dpsi = dpsi;
}

int or_rotor_get_prescribed(const or_sys *s, int ir, int pred, double *fc, double *gam,
                            double *rc)
{
    const rotor_t *r = &s->rotor[ir];
    for (int b = 0; b < r->nb; ++b)
        for (int i = 0; i < OR_NPF; ++i) {
            const fil_t *W = r->pf[pred] + (size_t)b * OR_NPF + i;
            const size_t q = (size_t)b * OR_NPF + i;
            if (fc) memcpy(fc + 6 * q, W->fc, 6 * sizeof(double));
            if (gam) gam[q] = W->gam;
            if (rc) rc[q] = W->rc;
        }
    return r->presc_present[pred];
}

/* blade_convectwake + blade_wake_continuity, classdef.f90:1515-1702;
 * rotor_convectwake with the axisymmetric copy, 4786-4830, and the
 * prescribed far wake after it (4826-4829) */
void or_rotor_convectwake(or_sys *s, int ir, double dt, int pred, int bug_compat)
{
    rotor_t *r = &s->rotor[ir];
    int nN = r->nNwake, ns = r->ns, rn = r->rowNear - 1, rf = r->rowFar - 1;
    for (int b = 0; b < r->nbConvect; ++b) {
        blade_t *bl = &r->blade[b];
        ring_t *N = pred ? bl->waNP : bl->waN;
        fil_t *F = pred ? bl->waFP : bl->waF;
        /* 'P' near-wake loops run `do i = 1, rowNear, nNwake`: only row 1 */
        int i0 = rn, i1 = nN;
        if (pred && bug_compat) { i0 = 0; i1 = 1; }
        for (int j = 0; j < ns; ++j)
            for (int i = i0; i < i1; ++i)
                for (int k = 0; k < 3; ++k) N[WN(r, i, j)].c[1][k] += bl->velN[VN(r, k, i, j)] * dt;
        for (int i = i0; i < i1; ++i)
            for (int k = 0; k < 3; ++k) N[WN(r, i, ns - 1)].c[2][k] += bl->velN[VN(r, k, i, ns)] * dt;
        for (int i = rf; i < r->nFwake; ++i)
            for (int k = 0; k < 3; ++k) F[i].fc[0][k] += bl->velF[3 * i + k] * dt;
        /* continuity */
        for (int j = 0; j < ns - 1; ++j)
            for (int i = rn + 1; i < nN; ++i) {
                memcpy(N[WN(r, i, j)].c[0], N[WN(r, i - 1, j)].c[1], 3 * sizeof(double));
                memcpy(N[WN(r, i, j)].c[2], N[WN(r, i, j + 1)].c[1], 3 * sizeof(double));
                memcpy(N[WN(r, i, j)].c[3], N[WN(r, i - 1, j + 1)].c[1], 3 * sizeof(double));
            }
        for (int j = 0; j < ns - 1; ++j)
            memcpy(N[WN(r, rn, j)].c[2], N[WN(r, rn, j + 1)].c[1], 3 * sizeof(double));
        for (int i = rn + 1; i < nN; ++i) {
            memcpy(N[WN(r, i, ns - 1)].c[0], N[WN(r, i - 1, ns - 1)].c[1], 3 * sizeof(double));
            memcpy(N[WN(r, i, ns - 1)].c[3], N[WN(r, i - 1, ns - 1)].c[2], 3 * sizeof(double));
        }
        for (int i = rf + 1; i < r->nFwake; ++i)
            memcpy(F[i].fc[1], F[i - 1].fc[0], 3 * sizeof(double));
    }
    if (r->nbConvect == 1 && r->nb > 1) {
        const blade_t *b0 = &r->blade[0];
        for (int b = 1; b < r->nb; ++b) {
            blade_t *bl = &r->blade[b];
            ring_t *N = pred ? bl->waNP : bl->waN;
            fil_t *F = pred ? bl->waFP : bl->waF;
            const ring_t *N0 = pred ? b0->waNP : b0->waN;
            const fil_t *F0 = pred ? b0->waFP : b0->waF;
            for (int j = 0; j < ns; ++j)
                for (int i = rn; i < nN; ++i) N[WN(r, i, j)] = N0[WN(r, i, j)];
            for (int i = rf; i < r->nFwake; ++i) F[i] = F0[i];
            double theta = 2.0 * OR_PI / r->nb * b;
            if (fabs(theta) > OR_EPS) {
                double T[9];
                transform_axis(theta, r->shaft, T);
                for (int j = 0; j < ns; ++j)
                    for (int i = rn; i < nN; ++i)
                        for (int c = 0; c < 4; ++c) rot_point(T, r->hub, N[WN(r, i, j)].c[c]);
                for (int i = rf; i < r->nFwake; ++i) {
                    rot_point(T, r->hub, F[i].fc[0]);
                    rot_point(T, r->hub, F[i].fc[1]);
                }
            }
        }
    }
    if (r->presc_nt > 0 && r->iter > r->presc_nt) presc_update(r, pred);
}

/* ------------------------------------------------------------------ */
/* Image rotors, classdef.f90:722-735, 982-993, 4990-5131              */
/* ------------------------------------------------------------------ */
void or_rotor_set_image(or_sys *s, int ir, int src, int plane)
{
    s->rotor[ir].image_of = src;
    s->rotor[ir].image_axis = plane - 1;
}

/* rotor_mirrorGamma + map_gam: every wing ring's circulation is the
 * source's negated. */
void or_rotor_mirror_gamma(or_sys *s, int ir)
{
    rotor_t *r = &s->rotor[ir];
    const rotor_t *q = &s->rotor[r->image_of];
    for (int b = 0; b < r->nb; ++b)
        for (int k = 0; k < r->nc * r->ns; ++k)
            r->blade[b].wiP[k].gam = -q->blade[b].wiP[k].gam;
}

/* rotor_mirrorWake: mode 0 'C' (waN from rowNear, waF from rowFar), 1 'P'
 * (waNPredicted, waFPredicted over the same rows), 2 'A' (every waN and
 * waF); rings and filaments are copied whole, then vr_mirror /
 * Fwake_mirror negate coordinate image_axis of every end and gam. */
void or_rotor_mirror_wake(or_sys *s, int ir, int mode)
{
    rotor_t *r = &s->rotor[ir];
    const rotor_t *q = &s->rotor[r->image_of];
    const int ax = r->image_axis;
    const int i0 = mode == 2 ? 0 : r->rowNear - 1, f0 = mode == 2 ? 0 : r->rowFar - 1;
    for (int b = 0; b < r->nb; ++b) {
        ring_t *N = mode == 1 ? r->blade[b].waNP : r->blade[b].waN;
        fil_t *F = mode == 1 ? r->blade[b].waFP : r->blade[b].waF;
        const ring_t *NQ = mode == 1 ? q->blade[b].waNP : q->blade[b].waN;
        const fil_t *FQ = mode == 1 ? q->blade[b].waFP : q->blade[b].waF;
        for (int j = 0; j < r->ns; ++j)
            for (int i = i0; i < r->nNwake; ++i) {
                ring_t *w = &N[WN(r, i, j)];
                *w = NQ[WN(r, i, j)];
                for (int c = 0; c < 4; ++c) w->c[c][ax] = -w->c[c][ax];
                w->gam = -w->gam;
            }
        for (int i = f0; i < r->nFwake; ++i) {
            fil_t *f = &F[i];
            *f = FQ[i];
            f->fc[0][ax] = -f->fc[0][ax];
            f->fc[1][ax] = -f->fc[1][ax];
            f->gam = -f->gam;
        }
    }
}

/* ------------------------------------------------------------------ */
/* Scheme algebra, main.f90:844-1115; libCommon.f90:213-258            */
/* ------------------------------------------------------------------ */
static void copy_predicted(rotor_t *r)
{
    for (int b = 0; b < r->nbConvect; ++b) {
        blade_t *bl = &r->blade[b];
        for (int j = 0; j < r->ns; ++j)
            for (int i = r->rowNear - 1; i < r->nNwake; ++i)
                bl->waNP[WN(r, i, j)] = bl->waN[WN(r, i, j)];
        for (int i = r->rowFar - 1; i < r->nFwake; ++i) bl->waFP[i] = bl->waF[i];
    }
}

static void vel_order2(double *vn, const double *vnp1, int rows, int stride, int cols,
                       int base)
{
    /* vn(:, rows, cols) <- mixed average over the active row window */
    double *tmp = (double *)malloc(sizeof(double) * 3 * (size_t)rows * cols);
    for (int j = 0; j < cols; ++j)
        for (int i = 0; i < rows; ++i)
            for (int d = 0; d < 3; ++d) {
                size_t k = (size_t)d + 3 * ((size_t)(base + i) + (size_t)stride * j);
                size_t km = k - 3, kp = k + 3;
                double val;
                if (i == 0 || i == rows - 1) {
                    val = (vnp1[k] + vn[k]) * 0.5;
                } else {
                    val = (vnp1[k] + vnp1[km] + vn[kp] + vn[k]) * 0.25;
                }
                tmp[d + 3 * (i + (size_t)rows * j)] = val;
            }
    for (int j = 0; j < cols; ++j)
        for (int i = 0; i < rows; ++i)
            for (int d = 0; d < 3; ++d)
                vn[(size_t)d + 3 * ((size_t)(base + i) + (size_t)stride * j)] =
                    tmp[d + 3 * (i + (size_t)rows * j)];
    free(tmp);
}

void or_convect(or_sys *s, int scheme, int iter, double dt, int bug_compat)
{
    size_t nr = (size_t)s->nr;
    or_wake_velocity(s, 0);
    /* an image rotor (main.f90:851-856 and every later scheme block)
     * mirrors its source's wake where a lifting rotor convects */
#define OR_IMAGE(r, ir, mode)                                \
    if ((r)->image_of >= 0) {                                \
        if ((r)->nNwake > 0) or_rotor_mirror_wake(s, (int)(ir), (mode)); \
        continue;                                            \
    }
    if (scheme == 0) {
        for (size_t ir = 0; ir < nr; ++ir) {
            OR_IMAGE(&s->rotor[ir], ir, 0)
            if (s->rotor[ir].nNwake > 0) or_rotor_convectwake(s, (int)ir, dt, 0, bug_compat);
        }
        return;
    }
    if (scheme == 1) {  /* PC2, main.f90:861-949 */
        for (size_t ir = 0; ir < nr; ++ir) {
            rotor_t *r = &s->rotor[ir];
            OR_IMAGE(r, ir, 1)
            if (r->nNwake <= 0) continue;
            copy_predicted(r);
            or_rotor_convectwake(s, (int)ir, dt, 1, bug_compat);
        }
        or_wake_velocity(s, 1);
        for (size_t ir = 0; ir < nr; ++ir) {
            rotor_t *r = &s->rotor[ir];
            OR_IMAGE(r, ir, 0)
            if (r->nNwake <= 0) continue;
            for (int b = 0; b < r->nbConvect; ++b) {
                blade_t *bl = &r->blade[b];
                int rows = r->nNwake - (r->rowNear - 1);
                vel_order2(bl->velN, bl->velNP, rows, r->nNwake, r->ns + 1, r->rowNear - 1);
                int frows = r->nFwake - (r->rowFar - 1);
                if (frows > 0) vel_order2(bl->velF, bl->velFP, frows, r->nFwake, 1, r->rowFar - 1);
            }
            or_rotor_convectwake(s, (int)ir, dt, 0, bug_compat);
        }
        return;
    }
    size_t nvn, nvf;
    if (iter == 1) {  /* AB2 / AM2 first step */
        for (size_t ir = 0; ir < nr; ++ir) {
            rotor_t *r = &s->rotor[ir];
            OR_IMAGE(r, ir, 0)
            if (r->nNwake <= 0) continue;
            or_rotor_convectwake(s, (int)ir, dt, 0, bug_compat);
            nvn = 3u * (size_t)r->nNwake * (size_t)(r->ns + 1);
            nvf = 3u * (size_t)r->nFwake;
            for (int b = 0; b < r->nbConvect; ++b) {
                memcpy(r->blade[b].velN1, r->blade[b].velN, nvn * sizeof(double));
                memcpy(r->blade[b].velF1, r->blade[b].velF, nvf * sizeof(double));
            }
        }
        return;
    }
    if (scheme == 2) {  /* AB2, main.f90:968-1000 */
        for (size_t ir = 0; ir < nr; ++ir) {
            rotor_t *r = &s->rotor[ir];
            OR_IMAGE(r, ir, 0)
            if (r->nNwake <= 0) continue;
            nvn = 3u * (size_t)r->nNwake * (size_t)(r->ns + 1);
            nvf = 3u * (size_t)r->nFwake;
            for (int b = 0; b < r->nbConvect; ++b) {
                blade_t *bl = &r->blade[b];
                for (size_t k = 0; k < nvn; ++k) {
                    bl->velNS[k] = 0.5 * (3.0 * bl->velN[k] - bl->velN1[k]);
                    bl->velN1[k] = bl->velNS[k];
                    bl->velN[k] = bl->velNS[k];
                }
                for (size_t k = 0; k < nvf; ++k) {
                    bl->velFS[k] = 0.5 * (3.0 * bl->velF[k] - bl->velF1[k]);
                    bl->velF1[k] = bl->velFS[k];
                    bl->velF[k] = bl->velFS[k];
                }
            }
            or_rotor_convectwake(s, (int)ir, dt, 0, bug_compat);
        }
        return;
    }
    /* AM2, main.f90:1019-1115 */
    for (size_t ir = 0; ir < nr; ++ir) {
        rotor_t *r = &s->rotor[ir];
        OR_IMAGE(r, ir, 1)
        if (r->nNwake <= 0) continue;
        copy_predicted(r);
        nvn = 3u * (size_t)r->nNwake * (size_t)(r->ns + 1);
        nvf = 3u * (size_t)r->nFwake;
        for (int b = 0; b < r->nbConvect; ++b) {
            blade_t *bl = &r->blade[b];
            for (size_t k = 0; k < nvn; ++k) {
                bl->velNS[k] = bl->velN[k];
                bl->velN[k] = 0.5 * (3.0 * bl->velN[k] - bl->velN1[k]);
            }
            for (size_t k = 0; k < nvf; ++k) {
                bl->velFS[k] = bl->velF[k];
                bl->velF[k] = 0.5 * (3.0 * bl->velF[k] - bl->velF1[k]);
            }
        }
        or_rotor_convectwake(s, (int)ir, dt, 1, bug_compat);
    }
    or_wake_velocity(s, 1);
    for (size_t ir = 0; ir < nr; ++ir) {
        rotor_t *r = &s->rotor[ir];
        OR_IMAGE(r, ir, 0)
        if (r->nNwake <= 0) continue;
        nvn = 3u * (size_t)r->nNwake * (size_t)(r->ns + 1);
        nvf = 3u * (size_t)r->nFwake;
        for (int b = 0; b < r->nbConvect; ++b) {
            blade_t *bl = &r->blade[b];
            for (size_t k = 0; k < nvn; ++k) bl->velN[k] = (bl->velNP[k] + bl->velNS[k]) * 0.5;
            for (size_t k = 0; k < nvf; ++k) bl->velF[k] = (bl->velFP[k] + bl->velFS[k]) * 0.5;
        }
        or_rotor_convectwake(s, (int)ir, dt, 0, bug_compat);
        for (int b = 0; b < r->nbConvect; ++b) {
            blade_t *bl = &r->blade[b];
            memcpy(bl->velN1, bl->velNS, nvn * sizeof(double));
            memcpy(bl->velF1, bl->velFS, nvf * sizeof(double));
        }
    }
}

/* ------------------------------------------------------------------ */
/* Lattice accessors (layout shared with include/vlm_c.h)              */
/* ------------------------------------------------------------------ */
void or_rotor_get_wake(const or_sys *s, int ir, int pred, double *nodes,
                       double *le, double *gam, double *rc4, double *far_fc,
                       double *far_gam, double *far_rc)
{
    const rotor_t *r = &s->rotor[ir];
    int nN = r->nNwake, ns = r->ns, nF = r->nFwake;
    for (int b = 0; b < r->nb; ++b) {
        const blade_t *bl = &r->blade[b];
        const ring_t *N = pred ? bl->waNP : bl->waN;
        const fil_t *F = pred ? bl->waFP : bl->waF;
        for (int i = 0; i < nN; ++i) {
            for (int j = 0; j <= ns; ++j) {
                const double *P = (j < ns) ? N[WN(r, i, j)].c[1] : N[WN(r, i, ns - 1)].c[2];
                size_t k = ((size_t)b * nN + i) * (ns + 1) + j;
                if (nodes) memcpy(nodes + 3 * k, P, 3 * sizeof(double));
            }
            for (int j = 0; j < ns; ++j) {
                size_t k = ((size_t)b * nN + i) * ns + j;
                if (gam) gam[k] = N[WN(r, i, j)].gam;
                if (rc4) memcpy(rc4 + 4 * k, N[WN(r, i, j)].rc, 4 * sizeof(double));
            }
        }
        if (le && nN > 0) {
            int rn = r->rowNear - 1;
            if (rn >= nN) rn = nN - 1;
            for (int j = 0; j <= ns; ++j) {
                const double *P = (j < ns) ? N[WN(r, rn, j)].c[0] : N[WN(r, rn, ns - 1)].c[3];
                memcpy(le + 3 * ((size_t)b * (ns + 1) + j), P, 3 * sizeof(double));
            }
        }
        for (int i = 0; i < nF; ++i) {
            size_t k = (size_t)b * nF + i;
            if (far_fc) memcpy(far_fc + 6 * k, F[i].fc, 6 * sizeof(double));
            if (far_gam) far_gam[k] = F[i].gam;
            if (far_rc) far_rc[k] = F[i].rc;
        }
    }
}

void or_rotor_set_wake(or_sys *s, int ir, const double *nodes,
                       const double *le, const double *gam, const double *rc4,
                       const double *far_fc, const double *far_gam,
                       const double *far_rc, int rowNear, int rowFar)
{
    rotor_t *r = &s->rotor[ir];
    int nN = r->nNwake, ns = r->ns, nF = r->nFwake;
    r->rowNear = rowNear;
    r->rowFar = rowFar;
    for (int b = 0; b < r->nb; ++b) {
        blade_t *bl = &r->blade[b];
        for (int i = 0; i < nN; ++i)
            for (int j = 0; j < ns; ++j) {
                ring_t *w = &bl->waN[WN(r, i, j)];
                size_t kg = ((size_t)b * nN + i) * ns + j;
                size_t n2 = ((size_t)b * nN + i) * (ns + 1) + j;
                memcpy(w->c[1], nodes + 3 * n2, 3 * sizeof(double));
                memcpy(w->c[2], nodes + 3 * (n2 + 1), 3 * sizeof(double));
                if (i == rowNear - 1 || i == 0) {
                    const double *L = le + 3 * ((size_t)b * (ns + 1) + j);
                    memcpy(w->c[0], L, 3 * sizeof(double));
                    memcpy(w->c[3], L + 3, 3 * sizeof(double));
                }
                if (i > rowNear - 1 && i > 0) {
                    size_t n1 = ((size_t)b * nN + i - 1) * (ns + 1) + j;
                    memcpy(w->c[0], nodes + 3 * n1, 3 * sizeof(double));
                    memcpy(w->c[3], nodes + 3 * (n1 + 1), 3 * sizeof(double));
                }
                w->gam = gam[kg];
                memcpy(w->rc, rc4 + 4 * kg, 4 * sizeof(double));
                memcpy(w->rc0, rc4 + 4 * kg, 4 * sizeof(double));
                ring_calclength(w, 1);
            }
        for (int i = 0; i < nF; ++i) {
            fil_t *f = &bl->waF[i];
            size_t k = (size_t)b * nF + i;
            memcpy(f->fc, far_fc + 6 * k, 6 * sizeof(double));
            f->gam = far_gam[k];
            f->rc = f->rc0 = far_rc[k];
            double d[3];
            for (int q = 0; q < 3; ++q) d[q] = f->fc[0][q] - f->fc[1][q];
            f->lc = f->l0 = norm3(d);
        }
    }
}

void or_rotor_get_velocity(const or_sys *s, int ir, int which, double *velN,
                           double *velF)
{
    const rotor_t *r = &s->rotor[ir];
    size_t nvn = 3u * (size_t)r->nNwake * (size_t)(r->ns + 1);
    size_t nvf = 3u * (size_t)r->nFwake;
    for (int b = 0; b < r->nb; ++b) {
        const blade_t *bl = &r->blade[b];
        const double *sn = which == 0 ? bl->velN : which == 1 ? bl->velN1
                         : which == 2 ? bl->velNP : bl->velNS;
        const double *sf = which == 0 ? bl->velF : which == 1 ? bl->velF1
                         : which == 2 ? bl->velFP : bl->velFS;
        /* library layout: ((b*nNwake + i)*(ns+1) + j)*3 + d */
        for (int i = 0; i < r->nNwake; ++i)
            for (int j = 0; j <= r->ns; ++j)
                for (int d = 0; d < 3; ++d)
                    velN[(((size_t)b * r->nNwake + i) * (r->ns + 1) + j) * 3 + d] =
                        sn[VN(r, d, i, j)];
        if (velF) memcpy(velF + (size_t)b * nvf, sf, nvf * sizeof(double));
    }
    (void)nvn;
}

/* ------------------------------------------------------------------ */
/* a7: AIC (classdef.f90:4151-4179) and LU with partial pivoting       */
/* ------------------------------------------------------------------ */
void or_rotor_calc_aic(const or_sys *s, int ir, const double *cp,
                       const double *ncap, double *A)
{
    const rotor_t *r = &s->rotor[ir];
    int nc = r->nc, ns = r->ns, nb = r->nb;
    int n = nc * ns * nb;
    for (int ib = 0; ib < nb; ++ib)
        for (int is = 0; is < ns; ++is)
            for (int ic = 0; ic < nc; ++ic) {
                int row = ic + nc * is + ns * nc * ib;
                for (int jb = 0; jb < nb; ++jb)
                    for (int j = 0; j < ns; ++j)
                        for (int i = 0; i < nc; ++i) {
                            int col = i + nc * j + ns * nc * jb;
                            double v[3];
                            ring_vind(&r->blade[jb].wiP[WI(r, i, j)], cp + 3 * row, v);
                            A[row + (size_t)n * col] = v[0] * ncap[3 * row] +
                                                       v[1] * ncap[3 * row + 1] +
                                                       v[2] * ncap[3 * row + 2];
                        }
            }
}

/* Doolittle LU with partial pivoting, column-major; returns 0 on success */
int or_lu_factor(int n, double *A, int *ipiv)
{
    for (int k = 0; k < n; ++k) {
        int p = k;
        double amax = fabs(A[k + (size_t)n * k]);
        for (int i = k + 1; i < n; ++i)
            if (fabs(A[i + (size_t)n * k]) > amax) { amax = fabs(A[i + (size_t)n * k]); p = i; }
        ipiv[k] = p;
        if (amax == 0.0) return k + 1;
        if (p != k)
            for (int j = 0; j < n; ++j) {
                double t = A[k + (size_t)n * j];
                A[k + (size_t)n * j] = A[p + (size_t)n * j];
                A[p + (size_t)n * j] = t;
            }
        double inv = 1.0 / A[k + (size_t)n * k];
        for (int i = k + 1; i < n; ++i) A[i + (size_t)n * k] *= inv;
        for (int j = k + 1; j < n; ++j) {
            double a = A[k + (size_t)n * j];
            for (int i = k + 1; i < n; ++i) A[i + (size_t)n * j] -= A[i + (size_t)n * k] * a;
        }
    }
    return 0;
}

void or_lu_solve(int n, const double *LU, const int *ipiv, double *b)
{
    for (int k = 0; k < n; ++k)
        if (ipiv[k] != k) { double t = b[k]; b[k] = b[ipiv[k]]; b[ipiv[k]] = t; }
    for (int j = 0; j < n; ++j)
        for (int i = j + 1; i < n; ++i) b[i] -= LU[i + (size_t)n * j] * b[j];
    for (int j = n - 1; j >= 0; --j) {
        b[j] /= LU[j + (size_t)n * j];
        for (int i = 0; i < j; ++i) b[i] -= LU[i + (size_t)n * j] * b[j];
    }
}

/* matmulAX, libMath.f90:105-122 */
void or_matvec(int n, const double *A, const double *x, double *y)
{
    for (int i = 0; i < n; ++i) y[i] = 0.0;
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) y[i] += A[i + (size_t)n * j] * x[j];
}

/* ------------------------------------------------------------------ */
/* a13: MLP 45 -> 64 -> 64 -> 64 -> 3, tanh hidden layers, FP64        */
/* Parameter packing: W1(64x45) b1 W2(64x64) b2 W3(64x64) b3 W4(3x64) b4 */
/* (row-major W[out][in]).                                             */
/* ------------------------------------------------------------------ */
#define MLP_IN 45
#define MLP_H 64
#define MLP_OUT 3

static void mlp_layer(const double *W, const double *b, int nin, int nout,
                      const double *x, double *y, int act)
{
    for (int o = 0; o < nout; ++o) {
        double a = b[o];
        for (int i = 0; i < nin; ++i) a += W[o * nin + i] * x[i];
        y[o] = act ? tanh(a) : a;
    }
}

void or_mlp_forward(const double *params, int n, const double *x, double *y)
{
    const double *W1 = params, *b1 = W1 + MLP_H * MLP_IN;
    const double *W2 = b1 + MLP_H, *b2 = W2 + MLP_H * MLP_H;
    const double *W3 = b2 + MLP_H, *b3 = W3 + MLP_H * MLP_H;
    const double *W4 = b3 + MLP_H, *b4 = W4 + MLP_OUT * MLP_H;
    for (int s = 0; s < n; ++s) {
        double h1[MLP_H], h2[MLP_H], h3[MLP_H];
        mlp_layer(W1, b1, MLP_IN, MLP_H, x + (size_t)s * MLP_IN, h1, 1);
        mlp_layer(W2, b2, MLP_H, MLP_H, h1, h2, 1);
        mlp_layer(W3, b3, MLP_H, MLP_H, h2, h3, 1);
        mlp_layer(W4, b4, MLP_H, MLP_OUT, h3, y + (size_t)s * MLP_OUT, 0);
    }
}

/* Mean-squared-error loss 0.5/n * sum |y - target|^2 and its gradient
 * with respect to every parameter (same packing). */
double or_mlp_loss_grad(const double *params, int n, const double *x,
                        const double *target, double *grad)
{
    const int nW1 = MLP_H * MLP_IN, nW2 = MLP_H * MLP_H, nW4 = MLP_OUT * MLP_H;
    const double *W1 = params, *b1 = W1 + nW1;
    const double *W2 = b1 + MLP_H, *b2 = W2 + nW2;
    const double *W3 = b2 + MLP_H, *b3 = W3 + nW2;
    const double *W4 = b3 + MLP_H, *b4 = W4 + nW4;
    double *gW1 = grad, *gb1 = gW1 + nW1, *gW2 = gb1 + MLP_H, *gb2 = gW2 + nW2;
    double *gW3 = gb2 + MLP_H, *gb3 = gW3 + nW2, *gW4 = gb3 + MLP_H, *gb4 = gW4 + nW4;
    int np = nW1 + MLP_H + 2 * (nW2 + MLP_H) + nW4 + MLP_OUT;
    memset(grad, 0, sizeof(double) * (size_t)np);
    double loss = 0.0;
    for (int s = 0; s < n; ++s) {
        const double *x0 = x + (size_t)s * MLP_IN;
        double h1[MLP_H], h2[MLP_H], h3[MLP_H], y[MLP_OUT];
        mlp_layer(W1, b1, MLP_IN, MLP_H, x0, h1, 1);
        mlp_layer(W2, b2, MLP_H, MLP_H, h1, h2, 1);
        mlp_layer(W3, b3, MLP_H, MLP_H, h2, h3, 1);
        mlp_layer(W4, b4, MLP_H, MLP_OUT, h3, y, 0);
        double dy[MLP_OUT], d3[MLP_H], d2[MLP_H], d1[MLP_H];
        for (int o = 0; o < MLP_OUT; ++o) {
            double e = y[o] - target[(size_t)s * MLP_OUT + o];
            loss += 0.5 * e * e / n;
            dy[o] = e / n;
        }
        for (int o = 0; o < MLP_OUT; ++o) {
            gb4[o] += dy[o];
            for (int i = 0; i < MLP_H; ++i) gW4[o * MLP_H + i] += dy[o] * h3[i];
        }
        for (int i = 0; i < MLP_H; ++i) {
            double a = 0.0;
            for (int o = 0; o < MLP_OUT; ++o) a += W4[o * MLP_H + i] * dy[o];
            d3[i] = a * (1.0 - h3[i] * h3[i]);
        }
        for (int o = 0; o < MLP_H; ++o) {
            gb3[o] += d3[o];
            for (int i = 0; i < MLP_H; ++i) gW3[o * MLP_H + i] += d3[o] * h2[i];
        }
        for (int i = 0; i < MLP_H; ++i) {
            double a = 0.0;
            for (int o = 0; o < MLP_H; ++o) a += W3[o * MLP_H + i] * d3[o];
            d2[i] = a * (1.0 - h2[i] * h2[i]);
        }
        for (int o = 0; o < MLP_H; ++o) {
            gb2[o] += d2[o];
            for (int i = 0; i < MLP_H; ++i) gW2[o * MLP_H + i] += d2[o] * h1[i];
        }
        for (int i = 0; i < MLP_H; ++i) {
            double a = 0.0;
            for (int o = 0; o < MLP_H; ++o) a += W2[o * MLP_H + i] * d2[o];
            d1[i] = a * (1.0 - h1[i] * h1[i]);
        }
        for (int o = 0; o < MLP_H; ++o) {
            gb1[o] += d1[o];
            for (int i = 0; i < MLP_IN; ++i) gW1[o * MLP_IN + i] += d1[o] * x0[i];
        }
    }
    return loss;
}
