/* Error return of the Fortran host.
 *
 * The reference stops the process on every error (`error stop`, e.g.
 * libCommon.f90:75-77).  Inside a caller's process (Python through ctypes,
 * a C driver) the host must return a status instead: volcanor_run arms a
 * jump buffer and calls the Fortran entry volcanor_run_impl; host_fail
 * (every error site of the host and vlm_check) records the message and
 * jumps back, and volcanor_run returns 1 after destroying the run's libvlm
 * context.  Outside an armed call (the volcanor_gpu and gridgen_gpu
 * programs) host_fail prints the message and exits with status 1, the
 * reference's `error stop` behaviour.
 *
 * The jump skips the Fortran frames between the error site and
 * volcanor_run: their allocatable locals are not freed (a leak bounded by
 * one run's host arrays), and files they opened stay open.
 */
#include <setjmp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* Fortran side (volcanor_host.f90); weak, so that programs that link only
 * host_fail (gridgen_gpu) need not carry the host. */
extern void volcanor_run_impl(const char *cdir, int precision, int nt_override, int verbose,
                              void *ct_hist, int ct_cap, int hist_len, void *stats,
                              int tail_steps, int predictor, double pred_tol)
    __attribute__((weak));
extern void volcanor_host_cleanup(void) __attribute__((weak));

static jmp_buf g_env;
static volatile int g_armed = 0;
static char g_msg[1024];

void volcanor_host_fail(const char *msg, int len)
{
    if (len < 0) len = 0;
    if (len > (int)sizeof(g_msg) - 1) len = (int)sizeof(g_msg) - 1;
    memcpy(g_msg, msg, (size_t)len);
    g_msg[len] = '\0';
    if (g_armed) {
        g_armed = 0;
        longjmp(g_env, 1);
    }
    fprintf(stderr, "%s\n", g_msg);
    fflush(stderr);
    exit(1);
}

/* Status of the run: 0 done, 1 stopped by an error (volcanor_last_error). */
int volcanor_run(const char *cdir, int precision, int nt_override, int verbose, void *ct_hist,
                 int ct_cap, int hist_len, void *stats, int tail_steps, int predictor,
                 double pred_tol)
{
    if (g_armed) {
        snprintf(g_msg, sizeof(g_msg), "ERROR: volcanor_run is not re-entrant");
        return 1;
    }
    g_msg[0] = '\0';
    if (!volcanor_run_impl || !volcanor_host_cleanup) {
        snprintf(g_msg, sizeof(g_msg), "ERROR: the Fortran host is not linked");
        return 1;
    }
    if (setjmp(g_env)) {
        volcanor_host_cleanup();
        return 1;
    }
    g_armed = 1;
    volcanor_run_impl(cdir, precision, nt_override, verbose, ct_hist, ct_cap, hist_len, stats,
                      tail_steps, predictor, pred_tol);
    g_armed = 0;
    return 0;
}

const char *volcanor_last_error(void) { return g_msg; }
