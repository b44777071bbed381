/*
 * hackforce -- hackcode force calculation on the GPU: reads snapshots, adds
 * Potential and Acceleration, writes snapshots.
 *
 * Same keywords, defaults, input rules and output layout as NEMO's
 * hackforce (src/nbody/evolve/hackcode/hackcode1/hackforce.c:38-50 keywords,
 * 86-110 input_data, 111-169 read_snapshot, 177-215 force_calc, 231-272
 * write_snapshot); the force pass is bh_force of libnemo_bh.  Invoked as
 * hackforce_qp (Makefile:86-91 of the reference) it uses quadrupole cells.
 *
 * Keywords added here: usequad= (default: by program name),
 * mode=fast|strict|direct, group=, nleaf=, local=, criterion=hack|rcrit|bh86|sw94
 * (bh_criterion: treecode1's opening tests, treeload.c:232-252), device=.  fcells= is accepted and
 * ignored: the device tree sizes itself.  Errors print
 * "### Fatal error [hackforce]: ..." and exit with status 255
 * (error.c:100-103,171-180); an existing out= file is refused
 * (stropen.c:174-176).
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>

#include "../../include/bh.h"
#include "filestruct.h"
#include "getparam.h"

#define VERSION "1.6b-mi355x"

static gp_keyword keys[] = {
    {"in", NULL, "Input (snapshot) file with mass coordinates", NULL},
    {"out", NULL, "Output (snapshot) file with f.c. results", NULL},
    {"test", "", "Input file with test coordinates (if blank, in= coords are used)", NULL},
    {"tol", "1.0", "Cell subdivision tolerence", NULL},
    {"eps", "0.05", "Standard softening parameter", NULL},
    {"rsize", "4.0", "Side-length of initial box", NULL},
    {"rmin", "", "Lower left corner of initial box [default is -rsize/2 (centered)]", NULL},
    {"options", "mass,phase", "Output options: phase and/or mass", NULL},
    {"fcells", "0.75", "Cell/body allocation ratio (accepted, not used: the device tree sizes itself)", NULL},
    {"usequad", "", "Quadrupole cells (default: true when invoked as hackforce_qp)", NULL},
    {"mode", "fast", "fast (FP32 group walk), strict (FP64, reference order), direct (FP64 direct sum)", NULL},
    {"group", "0", "Particles per group in fast mode (0 = default, 64, 128, 192 or 256)", NULL},
    {"nleaf", "-1", "Cells with <= nleaf bodies interact body by body in fast mode (-1 = default)", NULL},
    {"local", "0", "Fast mode, monopole cells: far cells (group radius < local x distance) by group local expansion (0 = off, <= 0.5)", NULL},
    {"criterion", "hack", "Cell opening: hack (hackcode1), rcrit, bh86 or sw94 (treecode1's criteria)", NULL},
    {"device", "0", "GPU ordinal", NULL},
    {"VERSION", VERSION, "", NULL},
};
#define NKEYS ((int)(sizeof keys / sizeof keys[0]))

static const char *const criteria[] = {"hack", "rcrit", "bh86", "sw94"};   /* bh_criterion order */

#define fatal(...) gp_fatal(__VA_ARGS__)
#define param gp_param
#define dparam gp_dparam
#define scanopt gp_scanopt

static FILE *open_in(const char *name)
{
    if (!strcmp(name, "-")) return stdin;
    FILE *f = fopen(name, "rb");
    if (!f) fatal("stropen: cannot open file \"%s\" for reading", name);
    return f;
}

int main(int argc, char **argv)
{
    const char *base = strrchr(argv[0], '/');
    base = base ? base + 1 : argv[0];
    const int qp_name = strstr(base, "_qp") != NULL;
    gp_init("hackforce", keys, NKEYS, argc, argv);
    const char *in = param("in"), *out = param("out"), *test = param("test");
    const char *options = param("options");
    bh_params p;
    bh_default_params(&p);
    p.tol = dparam("tol");
    p.eps = dparam("eps");
    p.rsize = dparam("rsize");
    (void)dparam("fcells");
    p.usequad = gp_bparam("usequad", qp_name);
    p.group_size = (int)dparam("group");
    p.nleaf = (int)dparam("nleaf");
    p.local_alpha = dparam("local");
    p.criterion = gp_choice("criterion", criteria, 4);
    const char *mode = param("mode");
    int direct = 0;
    if (!strcmp(mode, "fast")) p.mode = BH_MODE_FAST;
    else if (!strcmp(mode, "strict")) p.mode = BH_MODE_STRICT;
    else if (!strcmp(mode, "direct")) direct = 1;
    else fatal("Parameter mode=%s: fast, strict or direct", mode);
    if (direct && *test) fatal("mode=direct evaluates the mass set only (no test=)");
    {
        /* burststring(rmin, ", "): fewer than 3 values -> -rsize/2 */
        const char *s = param("rmin");
        double v[3];
        int k = 0;
        while (*s && k < 3) {
            while (*s == ',' || *s == ' ') s++;
            if (!*s) break;
            char *end;
            v[k] = strtod(s, &end);
            if (end == s) fatal("Parameter rmin=%s is not a list of numbers", param("rmin"));
            k++;
            s = end;
        }
        if (k == 3) {
            p.rmin_given = 1;
            for (int d = 0; d < 3; d++) p.rmin[d] = v[d];
        }
    }
    if (strcmp(out, "-")) {
        struct stat sb;
        if (stat(out, &sb) == 0) fatal("stropen: file \"%s\" already exists", out);
    }

    fs_reader rin = {open_in(in), -1, ""}, rtest = {NULL, -1, ""};
    if (*test) rtest.f = open_in(test);
    char *hist = NULL;
    size_t histlen = 0;
    int nhist = 0;

    bh_ctx *ctx = NULL;
    int dev = (int)dparam("device");
    FILE *fo = NULL;
    fs_snapshot mass = {0}, tst = {0};
    int have_mass = 0;
    for (;;) {
        /* input_data() */
        int st;
        if (!rtest.f) {
            fs_free_snapshot(&mass);
            st = fs_read_snapshot(&rin, &mass, &hist, &histlen, &nhist);
            if (st < 0) fatal("%s: %s", in, rin.err);
            if (st == 0) break;
        } else {
            if (!have_mass) {
                st = fs_read_snapshot(&rin, &mass, &hist, &histlen, &nhist);
                if (st < 0) fatal("%s: %s", in, rin.err);
                if (st == 0) break;
                have_mass = 1;
            }
            fs_free_snapshot(&tst);
            st = fs_read_snapshot(&rtest, &tst, &hist, &histlen, &nhist);
            if (st < 0) fatal("%s: %s", test, rtest.err);
            if (st == 0) break;
        }
        const fs_snapshot *ts = rtest.f ? &tst : &mass;
        const int64_t nt = ts->n;

        /* force_calc() */
        if (!ctx) {
            bh_status s = bh_create(&ctx, dev);
            if (s != BH_OK) fatal("cannot open GPU device %s (bh_create failed)", param("device"));
        }
        double *phi = (double *)malloc((size_t)nt * sizeof(double));
        double *acc = (double *)malloc((size_t)nt * 3 * sizeof(double));
        if (!phi || !acc) fatal("out of memory");
        bh_stats stt;
        memset(&stt, 0, sizeof stt);
        bh_status s;
        if (direct) {
            int64_t *idx = (int64_t *)malloc((size_t)nt * sizeof(int64_t));
            for (int64_t i = 0; i < nt; i++) idx[i] = i;
            s = bh_upload(ctx, mass.n, mass.mass, mass.pos, 0, NULL);
            if (s == BH_OK) s = bh_direct(ctx, p.eps, nt, idx, phi, acc);
            free(idx);
        } else {
            s = bh_force(ctx, &p, mass.n, mass.mass, mass.pos, rtest.f ? nt : 0, rtest.f ? ts->pos : NULL, phi, acc,
                         &stt);
        }
        if (s != BH_OK) fatal("force calculation failed: %s", bh_last_error(ctx));

        /* out_result() */
        if (!fo) {
            if (!strcmp(out, "-")) fo = stdout;
            else {
                struct stat sb;
                if (stat(out, &sb) == 0) fatal("stropen: file \"%s\" already exists", out);
                fo = fopen(out, "wb");
                if (!fo) fatal("stropen: cannot open file \"%s\" for writing", out);
            }
            const char *h = hist;
            for (int k = 0; k < nhist; k++, h += strlen(h) + 1) fs_put_history(fo, h);
            char *cmdline = (char *)malloc(strlen(gp_cmdline()) + 32);
            sprintf(cmdline, "%s VERSION=%s", gp_cmdline(), VERSION);
            fs_put_history(fo, cmdline);
            free(cmdline);
        }
        const int n32 = (int)nt, cs = FS_CS_CARTESIAN_3D_PHASE;
        int n2b = stt.n2btot > INT32_MAX ? INT32_MAX : (int)stt.n2btot;
        int nbc = stt.nbctot > INT32_MAX ? INT32_MAX : (int)stt.nbctot;
        double cputree = stt.t_tree_ms / 60000.0, cpufcal = stt.t_walk_ms / 60000.0;
        double tsnap = ts->time;
        int d1[1] = {n32}, d2[2] = {n32, 3}, d3[3] = {n32, 2, 3};
        int werr = 0;
        werr |= fs_put_set(fo, "SnapShot");
        werr |= fs_put_set(fo, "Parameters");
        werr |= fs_put_data(fo, "Nobj", "i", &n32, 0, NULL);
        werr |= fs_put_data(fo, "Time", "d", &tsnap, 0, NULL);
        werr |= fs_put_data(fo, "tol", "d", &p.tol, 0, NULL);
        werr |= fs_put_data(fo, "eps", "d", &p.eps, 0, NULL);
        werr |= fs_put_tes(fo);
        werr |= fs_put_set(fo, "Particles");
        werr |= fs_put_data(fo, "CoordSystem", "i", &cs, 0, NULL);
        if (scanopt(options, "mass")) {
            /* Mass of the test points (the mass set's masses when test= is blank) */
            werr |= fs_put_data(fo, "Mass", "d", ts->mass, 1, d1);
        }
        if (scanopt(options, "phase")) {
            double *ps = (double *)malloc((size_t)nt * 6 * sizeof(double));
            if (!ps) fatal("out of memory");
            for (int64_t i = 0; i < nt; i++)
                for (int k = 0; k < 3; k++) {
                    ps[6 * i + k] = ts->pos[3 * i + k];
                    ps[6 * i + 3 + k] = ts->vel[3 * i + k];
                }
            werr |= fs_put_data(fo, "PhaseSpace", "d", ps, 3, d3);
            free(ps);
        }
        werr |= fs_put_data(fo, "Potential", "d", phi, 1, d1);
        werr |= fs_put_data(fo, "Acceleration", "d", acc, 2, d2);
        werr |= fs_put_tes(fo);
        werr |= fs_put_set(fo, "Diagnostics");
        werr |= fs_put_data(fo, "n2btot", "i", &n2b, 0, NULL);
        werr |= fs_put_data(fo, "nbctot", "i", &nbc, 0, NULL);
        werr |= fs_put_data(fo, "cputree", "d", &cputree, 0, NULL);
        werr |= fs_put_data(fo, "cpufcal", "d", &cpufcal, 0, NULL);
        werr |= fs_put_tes(fo);
        werr |= fs_put_tes(fo);
        if (werr) fatal("write error on %s", out);
        free(phi);
        free(acc);
    }
    if (fo && fo != stdout && fclose(fo)) fatal("write error on %s", out);
    if (fo == stdout) fflush(stdout);
    fs_free_snapshot(&mass);
    fs_free_snapshot(&tst);
    free(hist);
    if (ctx) bh_destroy(ctx);
    return 0;
}
