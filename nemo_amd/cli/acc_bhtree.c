/*
 * bhtree.so -- NEMO acceleration(5) plugin: self-gravity of the bodies
 * passed in, from the GPU tree of libnemo_bh.
 *
 * Interface of inc/acceleration.h:46-69 and the loader of
 * src/orbit/potential/acceleration.cc:385-391,426-503: NEMO finds
 * "bhtree.so" in $ACCPATH / $POTPATH / $NEMO/obj/potential and calls
 * iniacceleration(pars, npar, file, &acc, &need_mass, &need_vel).
 * pars: tol (default 1.0), eps (0.05), usequad (0), as hackforce's
 * defaults, and criterion (0 hackcode1's, 1-3 treecode1's default / bh86 /
 * sw94, bh_criterion); file is not used.  acc(ndim, t, n, m, x, v, f, pot, a,
 * indicator, type): all n bodies are the mass set, each body skips itself;
 * with flags f only bodies with f[i] & 1 receive pot/acc; indicator bit 0
 * adds the potential, bit 1 adds the acceleration (otherwise assigned);
 * type 'f' or 'd' gives the array type.  Failures print
 * "### Fatal error [bhtree]: ..." and exit with status 255, as NEMO's
 * error() does.
 */
#include <stdbool.h>
#include <stdio.h>
#include <stdlib.h>

#include "../../include/bh.h"

typedef void (*acc_pter)(int, double, int, const void *, const void *, const void *, const int *, void *, void *,
                         int, char);

static bh_ctx *g_ctx;
static bh_params g_par;
static double *g_buf;
static size_t g_cap;

static void die(const char *msg, const char *detail)
{
    fflush(stdout);
    fprintf(stderr, "### Fatal error [bhtree]: %s%s%s\n", msg, detail ? ": " : "", detail ? detail : "");
    exit(255);
}

static double *scratch(size_t n)
{
    if (n > g_cap) {
        free(g_buf);
        g_buf = (double *)malloc(n * sizeof(double));
        if (!g_buf) die("out of memory", NULL);
        g_cap = n;
    }
    return g_buf;
}

static void bh_acceleration(int ndim, double t, int n, const void *m, const void *x, const void *v, const int *f,
                            void *pot, void *acc, int add, char type)
{
    (void)t;
    (void)v;
    if (ndim != 3) die("bhtree needs ndim = 3", NULL);
    if (type != 'f' && type != 'd') die("array type must be 'f' or 'd'", NULL);
    if (n < 1) return;
    if (!m || !x) die("masses and positions are required", NULL);
    /* FP64 copies: mass[n], pos[3n], phi[n], acc[3n] */
    double *buf = scratch(8 * (size_t)n);
    double *dm = buf, *dx = buf + n, *dphi = buf + 4 * (size_t)n, *dacc = buf + 5 * (size_t)n;
    for (int i = 0; i < n; i++) {
        dm[i] = type == 'd' ? ((const double *)m)[i] : (double)((const float *)m)[i];
        for (int k = 0; k < 3; k++)
            dx[3 * i + k] = type == 'd' ? ((const double *)x)[3 * i + k] : (double)((const float *)x)[3 * i + k];
    }
    if (!g_ctx && bh_create(&g_ctx, 0) != BH_OK) die("cannot open GPU device 0", NULL);
    bh_stats st;
    if (bh_force(g_ctx, &g_par, n, dm, dx, 0, NULL, dphi, dacc, &st) != BH_OK)
        die("force calculation failed", bh_last_error(g_ctx));
    for (int i = 0; i < n; i++) {
        if (f && !(f[i] & 1)) continue;
        if (pot) {
            if (type == 'd') {
                double *p = (double *)pot;
                p[i] = (add & 1) ? p[i] + dphi[i] : dphi[i];
            } else {
                float *p = (float *)pot;
                p[i] = (add & 1) ? p[i] + (float)dphi[i] : (float)dphi[i];
            }
        }
        if (acc) {
            for (int k = 0; k < 3; k++) {
                if (type == 'd') {
                    double *a = (double *)acc;
                    a[3 * i + k] = (add & 2) ? a[3 * i + k] + dacc[3 * i + k] : dacc[3 * i + k];
                } else {
                    float *a = (float *)acc;
                    a[3 * i + k] = (add & 2) ? a[3 * i + k] + (float)dacc[3 * i + k] : (float)dacc[3 * i + k];
                }
            }
        }
    }
}

void iniacceleration(const double *pars, int npar, const char *file, acc_pter *accf, bool *need_mass,
                     bool *need_vel)
{
    (void)file;
    bh_default_params(&g_par);
    if (npar > 0) g_par.tol = pars[0];
    if (npar > 1) g_par.eps = pars[1];
    if (npar > 2) g_par.usequad = pars[2] != 0.0;
    if (npar > 3) g_par.criterion = (int)pars[3];
    if (npar > 4) fprintf(stderr, "bhtree: parameters beyond (tol, eps, usequad, criterion) are ignored\n");
    if (g_par.criterion < BH_CRIT_HACK || g_par.criterion > BH_CRIT_SW94) die("criterion must be 0..3", NULL);
    if (!(g_par.tol >= 0.0) || !(g_par.eps >= 0.0)) die("tol and eps must be >= 0", NULL);
    if (accf) *accf = bh_acceleration;
    if (need_mass) *need_mass = true;
    if (need_vel) *need_vel = false;
}
