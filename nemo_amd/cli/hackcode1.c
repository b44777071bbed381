/*
 * hackcode1 -- Barnes & Hut hierarchical N-body code with the force pass and
 * the leapfrog on the GPU (bh_lf_* of libnemo_bh).
 *
 * Same keywords, run loop, log table, snapshot frames and state files as
 * NEMO's hackcode1 (src/nbody/evolve/hackcode/hackcode1/code.c:21-48
 * keywords, 57-64 main loop, 70-121 startrun, 159-191 stepsystem;
 * code_io.c:33-52 inputdata, 60-77 initoutput, 111-158 output, 160-197
 * diagnostics, 203-216 snapshot diagnostics, 222-284 savestate and
 * restorestate).  A step is: force pass (tree + walk), second-order velocity
 * correction after the first step, diagnostics and output, drift.
 *
 * State files hold the reference's items (program, version, headline,
 * parameters, tnow, tout, minor_tout, nstep, rmin, rsize, nbody, bodytab)
 * with bodytab as the reference's 96-byte body records (type, mass, pos,
 * vel, acc, phi on an LP64 host), so a saved run resumes with continue= or
 * restart=.  Without in=, restart= and continue=, the initial system is the
 * reference's test model (uniform ball, or M(r) = r with cencon=) drawn
 * from this program's own generator seeded by seed=.  Keywords added:
 * usequad= (default: true when invoked as hackcode1_qp), mode=fast|strict,
 * group=, nleaf=, local=, criterion=hack|rcrit|bh86|sw94 (treecode1's opening
 * tests), device=.  fcells= is accepted and ignored.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <time.h>

#include "../../include/bh.h"
#include "filestruct.h"
#include "getparam.h"

#define VERSION "1.5b-mi355x"
#define PROGRAM "hackcode1"

static gp_keyword keys[] = {
    {"in", "", "snapshot of initial conditions", NULL},
    {"out", "", "stream of output snapshots", NULL},
    {"restart", "", "input state and set controls", NULL},
    {"continue", "", "input state and continue run", NULL},
    {"save", "", "output state as code runs", NULL},
    {"nbody", "128", "number of particles to generate", NULL},
    {"seed", "123", "random number generator seed", NULL},
    {"cencon", "false", "centrally concentrated system (M(r) = r) instead of a uniform ball", NULL},
    {"freq", "32.0", "fundamental integration frequency", NULL},
    {"eps", "0.05", "usual potential softening", NULL},
    {"tol", "1.0", "cell subdivision tolerence", NULL},
    {"fcells", "1.0", "cell allocation parameter (accepted, not used)", NULL},
    {"options", "mass,phase", "misc. control options", NULL},
    {"tstop", "2.0", "time to stop integration", NULL},
    {"freqout", "4.0", "major data-output frequency", NULL},
    {"minor_freqout", "32.0", "minor data-output frequency", NULL},
    {"log", "-", "logging output", NULL},
    {"usequad", "", "quadrupole cells (default: true when invoked as hackcode1_qp)", NULL},
    {"mode", "fast", "fast (FP32 group walk) or strict (FP64, reference order)", NULL},
    {"group", "0", "particles per group in fast mode (0 = default, 64, 128, 192 or 256)", NULL},
    {"nleaf", "-1", "cells with <= nleaf bodies interact body by body in fast mode (-1 = default)", NULL},
    {"local", "0", "fast mode, monopole cells: far cells (group radius < local x distance) by group local expansion (0 = off, <= 0.5)", NULL},
    {"criterion", "hack", "Cell opening: hack (hackcode1), rcrit, bh86 or sw94 (treecode1's criteria)", NULL},
    {"device", "0", "GPU ordinal", NULL},
    {"VERSION", VERSION, "", NULL},
};
#define NKEYS ((int)(sizeof keys / sizeof keys[0]))
static const char *const criteria[] = {"hack", "rcrit", "bh86", "sw94"};   /* bh_criterion order */

/* run state (code.h globals) */
static double freq, eps, tol, fcells, tstop, freqout, minor_freqout, tnow, tout, minor_tout;
static int nstep, nbody;
static double rmin[3], rsize;
static char *options, *headline;
static double *mass, *pos, *vel, *acc, *phi;
static int restored;
static FILE *logstr, *outstr;
static const char *savefile;
static int firstmass = 1;

static double cputime(void) { return (double)clock() / CLOCKS_PER_SEC / 60.0; }

static char *dupstr(const char *s)
{
    char *d = (char *)malloc(strlen(s) + 1);
    strcpy(d, s);
    return d;
}

static void alloc_bodies(int n)
{
    nbody = n;
    mass = (double *)calloc((size_t)n, sizeof(double));
    pos = (double *)calloc(3 * (size_t)n, sizeof(double));
    vel = (double *)calloc(3 * (size_t)n, sizeof(double));
    acc = (double *)calloc(3 * (size_t)n, sizeof(double));
    phi = (double *)calloc((size_t)n, sizeof(double));
    if (!mass || !pos || !vel || !acc || !phi) gp_fatal("out of memory for %d bodies", n);
}

/* ---- input (inputdata, make_testdata) ---- */

static char *hist;
static size_t histlen;
static int nhist;

static void inputdata(const char *file)
{
    FILE *f = strcmp(file, "-") ? fopen(file, "rb") : stdin;
    if (!f) gp_fatal("stropen: cannot open file \"%s\" for reading", file);
    fs_reader r = {f, -1, ""};
    fs_snapshot s;
    const int st = fs_read_snapshot(&r, &s, &hist, &histlen, &nhist);
    if (st < 0) gp_fatal("inputdata: %s: %s", file, r.err);
    if (st == 0) gp_fatal("inputdata: essential data missing in %s", file);
    if (f != stdin) fclose(f);
    alloc_bodies((int)s.n);
    memcpy(mass, s.mass, (size_t)nbody * sizeof(double));
    memcpy(pos, s.pos, 3 * (size_t)nbody * sizeof(double));
    memcpy(vel, s.vel, 3 * (size_t)nbody * sizeof(double));
    tnow = gp_scanopt(options, "reset_time") ? 0.0 : s.time;
    fs_free_snapshot(&s);
}

/* Uniform deviates in [0, 1) from a counter-based generator (splitmix64 of
 * seed and a running counter): the test model's own RNG. */
static uint64_t td_state;

static double td_uniform(void)
{
    uint64_t z = (td_state += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    z ^= z >> 31;
    return (double)(z >> 11) * 0x1.0p-53;
}

/* A point uniform in the ball of radius r (rejection from the cube). */
static void td_ball(double *x, double r)
{
    double s;
    do {
        for (int k = 0; k < 3; k++) x[k] = 2.0 * td_uniform() - 1.0;
        s = x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
    } while (s > 1.0 || s == 0.0);
    for (int k = 0; k < 3; k++) x[k] *= r;
}

/* A point on the sphere of radius r (uniform direction). */
static void td_shell(double *x, double r)
{
    td_ball(x, 1.0);
    const double n = sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
    for (int k = 0; k < 3; k++) x[k] *= r / n;
}

/* The reference's test model (make_testdata, code.c:129-153, and pickvec,
 * util.c:15-27): equal masses 1/nbody; positions uniform in the unit ball,
 * or with cencon on shells of radius uniform in [0, 1) (M(r) = r);
 * velocities uniform in the unit ball; centre-of-mass position and velocity
 * removed.  Drawn from this program's own generator seeded by seed=. */
static void make_testdata(void)
{
    headline = dupstr("Hack code: test data");
    tnow = 0.0;
    alloc_bodies(nbody);
    td_state = (uint64_t)gp_iparam("seed") * 0xd1342543de82ef95ULL + 1;
    const int cencon = gp_bparam("cencon", 0);
    double cmr[3] = {0, 0, 0}, cmv[3] = {0, 0, 0};
    for (int i = 0; i < nbody; i++) {
        mass[i] = 1.0 / nbody;
        if (cencon) td_shell(pos + 3 * i, td_uniform());
        else td_ball(pos + 3 * i, 1.0);
        td_ball(vel + 3 * i, 1.0);
        for (int k = 0; k < 3; k++) {
            cmr[k] += pos[3 * i + k];
            cmv[k] += vel[3 * i + k];
        }
    }
    for (int k = 0; k < 3; k++) {
        cmr[k] /= nbody;
        cmv[k] /= nbody;
    }
    for (int i = 0; i < nbody; i++)
        for (int k = 0; k < 3; k++) {
            pos[3 * i + k] -= cmr[k];
            vel[3 * i + k] -= cmv[k];
        }
}

/* ---- state files (savestate, restorestate) ---- */

enum { kBodyBytes = 96 };   /* short type; pad; double mass, pos[3], vel[3], acc[3], phi */

static void put_str(FILE *f, const char *tag, const char *s)
{
    int d = (int)strlen(s) + 1;
    fs_put_data(f, tag, "c", s, 1, &d);
}

static void savestate(const char *file)
{
    FILE *f = fopen(file, "wb");
    if (!f) gp_fatal("savestate: cannot open %s", file);
    unsigned char *bt = (unsigned char *)calloc((size_t)nbody, kBodyBytes);
    for (int i = 0; i < nbody; i++) {
        unsigned char *b = bt + (size_t)i * kBodyBytes;
        const short type = 1;
        double rec[11];
        rec[0] = mass[i];
        for (int k = 0; k < 3; k++) {
            rec[1 + k] = pos[3 * i + k];
            rec[4 + k] = vel[3 * i + k];
            rec[7 + k] = acc[3 * i + k];
        }
        rec[10] = phi[i];
        memcpy(b, &type, sizeof type);
        memcpy(b + 8, rec, sizeof rec);
    }
    const int n3 = 3, dims[2] = {nbody, kBodyBytes};
    int err = 0;
    put_str(f, "program", PROGRAM);
    put_str(f, "version", VERSION);
    put_str(f, "headline", headline);
    err |= fs_put_data(f, "freq", "d", &freq, 0, NULL);
    err |= fs_put_data(f, "tol", "d", &tol, 0, NULL);
    err |= fs_put_data(f, "eps", "d", &eps, 0, NULL);
    err |= fs_put_data(f, "fcells", "d", &fcells, 0, NULL);
    put_str(f, "options", options);
    err |= fs_put_data(f, "tstop", "d", &tstop, 0, NULL);
    err |= fs_put_data(f, "freqout", "d", &freqout, 0, NULL);
    err |= fs_put_data(f, "minor_freqout", "d", &minor_freqout, 0, NULL);
    err |= fs_put_data(f, "tnow", "d", &tnow, 0, NULL);
    err |= fs_put_data(f, "tout", "d", &tout, 0, NULL);
    err |= fs_put_data(f, "minor_tout", "d", &minor_tout, 0, NULL);
    err |= fs_put_data(f, "nstep", "i", &nstep, 0, NULL);
    err |= fs_put_data(f, "rmin", "d", rmin, 1, &n3);
    err |= fs_put_data(f, "rsize", "d", &rsize, 0, NULL);
    err |= fs_put_data(f, "nbody", "i", &nbody, 0, NULL);
    err |= fs_put_data(f, "bodytab", "a", bt, 2, dims);
    free(bt);
    if (err || fclose(f)) gp_fatal("savestate: write error on %s", file);
}

static const fs_item *state_item(const fs_item *items, int n, const char *tag, const char *type, size_t count)
{
    for (int k = 0; k < n; k++)
        if (!strcmp(items[k].tag, tag)) {
            if (strcmp(items[k].type, type) || (count && items[k].count != count))
                gp_fatal("restorestate: item %s has the wrong type or size", tag);
            return &items[k];
        }
    gp_fatal("restorestate: item %s missing", tag);
}

static double sd(const fs_item *it) { return *(const double *)it->data; }

static void restorestate(const char *file)
{
    FILE *f = fopen(file, "rb");
    if (!f) gp_fatal("stropen: cannot open file \"%s\" for reading", file);
    fs_reader r = {f, -1, ""};
    fs_item items[64];
    int n = 0;
    for (;;) {
        if (n == 64) gp_fatal("restorestate: too many items in %s", file);
        const int st = fs_read_item(&r, &items[n]);
        if (st < 0) gp_fatal("restorestate: %s: %s", file, r.err);
        if (st == 0) break;
        n++;
    }
    fclose(f);
    const char *prog = (const char *)state_item(items, n, "program", "c", 0)->data;
    const char *ver = (const char *)state_item(items, n, "version", "c", 0)->data;
    if (strcmp(prog, PROGRAM) || strcmp(ver, VERSION))
        fprintf(logstr, "warning: state file may be outdated\n\n");
    headline = dupstr((const char *)state_item(items, n, "headline", "c", 0)->data);
    freq = sd(state_item(items, n, "freq", "d", 1));
    tol = sd(state_item(items, n, "tol", "d", 1));
    eps = sd(state_item(items, n, "eps", "d", 1));
    fcells = sd(state_item(items, n, "fcells", "d", 1));
    options = dupstr((const char *)state_item(items, n, "options", "c", 0)->data);
    tstop = sd(state_item(items, n, "tstop", "d", 1));
    freqout = sd(state_item(items, n, "freqout", "d", 1));
    minor_freqout = sd(state_item(items, n, "minor_freqout", "d", 1));
    tnow = sd(state_item(items, n, "tnow", "d", 1));
    tout = sd(state_item(items, n, "tout", "d", 1));
    minor_tout = sd(state_item(items, n, "minor_tout", "d", 1));
    nstep = *(const int32_t *)state_item(items, n, "nstep", "i", 1)->data;
    const double *rm = (const double *)state_item(items, n, "rmin", "d", 3)->data;
    for (int k = 0; k < 3; k++) rmin[k] = rm[k];
    rsize = sd(state_item(items, n, "rsize", "d", 1));
    const int nb = *(const int32_t *)state_item(items, n, "nbody", "i", 1)->data;
    if (nb < 1) gp_fatal("restorestate: nbody = %d", nb);
    const unsigned char *bt =
        (const unsigned char *)state_item(items, n, "bodytab", "a", (size_t)nb * kBodyBytes)->data;
    alloc_bodies(nb);
    for (int i = 0; i < nb; i++) {
        double rec[11];
        memcpy(rec, bt + (size_t)i * kBodyBytes + 8, sizeof rec);
        mass[i] = rec[0];
        for (int k = 0; k < 3; k++) {
            pos[3 * i + k] = rec[1 + k];
            vel[3 * i + k] = rec[4 + k];
            acc[3 * i + k] = rec[7 + k];
        }
        phi[i] = rec[10];
    }
    for (int k = 0; k < n; k++) fs_free_item(&items[k]);
    restored = 1;
}

/* ---- output (initoutput, output) ---- */

static void put_frame(const double *diag, int bits_phase, int bits_mass, int bits_phi, int bits_acc)
{
    const int cs = FS_CS_CARTESIAN_3D_PHASE, d1[1] = {nbody}, d2[2] = {nbody, 3}, d3[3] = {nbody, 2, 3};
    const int e3[1] = {3}, t33[2] = {3, 3}, c23[2] = {2, 3};
    int err = 0;
    err |= fs_put_set(outstr, "SnapShot");
    err |= fs_put_set(outstr, "Parameters");
    err |= fs_put_data(outstr, "Nobj", "i", &nbody, 0, NULL);
    err |= fs_put_data(outstr, "Time", "d", &tnow, 0, NULL);
    err |= fs_put_tes(outstr);
    if (bits_phase || bits_mass || bits_phi || bits_acc) {
        err |= fs_put_set(outstr, "Particles");
        err |= fs_put_data(outstr, "CoordSystem", "i", &cs, 0, NULL);
        if (bits_mass) err |= fs_put_data(outstr, "Mass", "d", mass, 1, d1);
        if (bits_phase) {
            double *ps = (double *)malloc((size_t)nbody * 6 * sizeof(double));
            if (!ps) gp_fatal("out of memory");
            for (int i = 0; i < nbody; i++)
                for (int k = 0; k < 3; k++) {
                    ps[6 * i + k] = pos[3 * i + k];
                    ps[6 * i + 3 + k] = vel[3 * i + k];
                }
            err |= fs_put_data(outstr, "PhaseSpace", "d", ps, 3, d3);
            free(ps);
        }
        if (bits_phi) err |= fs_put_data(outstr, "Potential", "d", phi, 1, d1);
        if (bits_acc) err |= fs_put_data(outstr, "Acceleration", "d", acc, 2, d2);
        err |= fs_put_tes(outstr);
    }
    const double cput = cputime();
    err |= fs_put_set(outstr, "Diagnostics");
    err |= fs_put_data(outstr, "Energy", "d", diag, 1, e3);
    err |= fs_put_data(outstr, "KETensor", "d", diag + 3, 2, t33);
    err |= fs_put_data(outstr, "PETensor", "d", diag + 12, 2, t33);
    err |= fs_put_data(outstr, "AMTensor", "d", diag + 21, 2, t33);
    err |= fs_put_data(outstr, "CMPhaseSpace", "d", diag + 30, 2, c23);
    err |= fs_put_data(outstr, "cputime", "d", &cput, 0, NULL);
    err |= fs_put_tes(outstr);
    err |= fs_put_tes(outstr);
    fflush(outstr);
    if (err) gp_fatal("write error on the output snapshot");
}

static void output(bh_ctx *ctx, const double *diag, long long n2b, long long nbc)
{
    const long long nttot = n2b + nbc;
    const int nbavg = (int)((double)n2b / (double)nbody), ncavg = (int)((double)nbc / (double)nbody);
    fprintf(logstr, "\n  %10s%10s%10s%10s%10s%10s%10s\n", "tnow", "T+U", "T/U", "nttot", "nbavg", "ncavg", "cputime");
    fprintf(logstr, "  %10.3f%10.4f%10.4f%10lld%10d%10d%10.2f\n\n", tnow, diag[0], diag[1] / diag[2], nttot, nbavg,
            ncavg, cputime());
    fprintf(logstr, "\t    %10s", "cm pos");
    for (int k = 0; k < 3; k++) fprintf(logstr, "%10.4f", diag[30 + k]);
    fprintf(logstr, "\n\t    %10s", "cm vel");
    for (int k = 0; k < 3; k++) fprintf(logstr, "%10.4f", diag[33 + k]);
    fprintf(logstr, "\n");
    int timebit = 0, phase = 0, wmass = 0, wphi = 0, wacc = 0;
    if (minor_freqout > 0.0 && (minor_tout - 0.01 / freq) <= tnow) {
        minor_tout += 1.0 / minor_freqout;
        timebit = 1;
    }
    if (freqout > 0.0 && (tout - 0.01 / freq) <= tnow) {
        tout += 1.0 / freqout;
        timebit = phase = 1;
        if (gp_scanopt(options, "mass") || firstmass) {
            wmass = 1;
            firstmass = 0;
        }
        wphi = gp_scanopt(options, "phi");
        wacc = gp_scanopt(options, "acc");
    }
    const int want = (timebit && outstr) || *savefile;
    if (want && (phase || *savefile)) {
        int64_t ns;
        if (bh_lf_get(ctx, pos, vel, acc, phi, &ns) != BH_OK) gp_fatal("bh_lf_get: %s", bh_last_error(ctx));
        if (bh_lf_box(ctx, rmin, &rsize) != BH_OK) gp_fatal("bh_lf_box: %s", bh_last_error(ctx));
    }
    if (timebit && outstr) {
        put_frame(diag, phase, wmass, wphi, wacc);
        if (phase) fprintf(logstr, "\n\tparticle data written\n");
    }
    if (*savefile) savestate(savefile);
    fflush(logstr);
}

int main(int argc, char **argv)
{
    const char *base = strrchr(argv[0], '/');
    base = base ? base + 1 : argv[0];
    const int qp_name = strstr(base, "_qp") != NULL;
    gp_init(PROGRAM, keys, NKEYS, argc, argv);
    const char *infile = gp_param("in"), *outfile = gp_param("out");
    const char *restfile = gp_param("restart"), *contfile = gp_param("continue");
    const char *logfile = gp_param("log");
    savefile = gp_param("save");
    options = dupstr(gp_param("options"));
    headline = dupstr("Hack code");
    logstr = strcmp(logfile, "-") ? fopen(logfile, "w") : stdout;
    if (!logstr) gp_fatal("stropen: cannot open log file \"%s\"", logfile);
    if (*outfile && strcmp(outfile, "-")) {
        struct stat sb;
        if (stat(outfile, &sb) == 0) gp_fatal("stropen: file \"%s\" already exists", outfile);
    }

    /* startrun() */
    if (*contfile) {
        restorestate(contfile);
    } else if (*restfile) {
        restorestate(restfile);
        eps = gp_dparam("eps");
        tol = gp_dparam("tol");
        options = dupstr(gp_param("options"));
        fcells = gp_dparam("fcells");
        tstop = gp_dparam("tstop");
        freqout = gp_dparam("freqout");
        minor_freqout = gp_dparam("minor_freqout");
        if (gp_scanopt(options, "new_tout")) {
            tout = tnow + 1 / freqout;
            minor_tout = tnow + 1 / minor_freqout;
        }
    } else {
        if (*infile) {
            inputdata(infile);
        } else {
            nbody = gp_iparam("nbody");
            if (nbody < 1) gp_fatal("startrun: absurd nbody");
            make_testdata();
        }
        freq = gp_dparam("freq");
        eps = gp_dparam("eps");
        tol = gp_dparam("tol");
        fcells = gp_dparam("fcells");
        tstop = gp_dparam("tstop");
        freqout = gp_dparam("freqout");
        minor_freqout = gp_dparam("minor_freqout");
        nstep = 0;
        minor_tout = tout = tnow;
        for (int k = 0; k < 3; k++) rmin[k] = -2.0;
        rsize = -2.0 * rmin[0];
    }
    if (!(freq > 0.0)) gp_fatal("freq must be > 0");

    bh_params p;
    bh_default_params(&p);
    p.tol = tol;
    p.eps = eps;
    p.usequad = gp_bparam("usequad", qp_name);
    const char *mode = gp_param("mode");
    if (!strcmp(mode, "fast")) p.mode = BH_MODE_FAST;
    else if (!strcmp(mode, "strict")) p.mode = BH_MODE_STRICT;
    else gp_fatal("Parameter mode=%s: fast or strict", mode);
    p.group_size = gp_iparam("group");
    p.nleaf = gp_iparam("nleaf");
    p.local_alpha = gp_dparam("local");
    p.criterion = gp_choice("criterion", criteria, 4);
    p.rsize = rsize;
    p.rmin_given = 1;
    for (int k = 0; k < 3; k++) p.rmin[k] = rmin[k];
    bh_ctx *ctx = NULL;
    if (bh_create(&ctx, gp_iparam("device")) != BH_OK) gp_fatal("cannot open GPU device %s", gp_param("device"));
    if (bh_lf_init(ctx, &p, nbody, mass, pos, vel, 1.0 / freq) != BH_OK)
        gp_fatal("bh_lf_init: %s", bh_last_error(ctx));
    if (restored && bh_lf_restore(ctx, acc, nstep, rmin, rsize) != BH_OK)
        gp_fatal("bh_lf_restore: %s", bh_last_error(ctx));

    /* initoutput() */
    fprintf(logstr, "\n%s\n\n", headline);
    fprintf(logstr, "%12s%12s%12s%12s\n", "nbody", "freq", "eps", "tol");
    fprintf(logstr, "%12d%12.2f%12.4f%12.4f\n\n", nbody, freq, eps, tol);
    if (*options) fprintf(logstr, "\toptions: %s\n", options);
    if (*outfile) {
        outstr = strcmp(outfile, "-") ? fopen(outfile, "wb") : stdout;
        if (!outstr) gp_fatal("stropen: cannot open file \"%s\" for writing", outfile);
        const char *h = hist;
        for (int k = 0; k < nhist; k++, h += strlen(h) + 1) fs_put_history(outstr, h);
        char *cmd = (char *)malloc(strlen(gp_cmdline()) + 32);
        sprintf(cmd, "%s VERSION=%s", gp_cmdline(), VERSION);
        fs_put_history(outstr, cmd);
        free(cmd);
    }
    if (*savefile) savestate(savefile);

    /* while not past tstop: stepsystem() */
    const double dt = 1.0 / freq;
    double diag[36];
    while (tnow < tstop + 0.1 / freq) {
        int64_t n2b, nbc;
        if (bh_lf_force(ctx, diag, &n2b, &nbc) != BH_OK) gp_fatal("force pass: %s", bh_last_error(ctx));
        output(ctx, diag, n2b, nbc);
        if (bh_lf_drift(ctx) != BH_OK) gp_fatal("drift: %s", bh_last_error(ctx));
        nstep++;
        tnow = tnow + dt;
    }
    /* stopoutput() */
    if (outstr && outstr != stdout && fclose(outstr)) gp_fatal("write error on %s", outfile);
    if (logstr != stdout) fclose(logstr);
    bh_destroy(ctx);
    return 0;
}
