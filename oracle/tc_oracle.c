/*
 * tc_oracle.c -- CPU restatement of NEMO's treecode1 force calculation
 * (Barnes' 2001 treecode) over a given root box.
 *
 * TEST INFRASTRUCTURE ONLY.  The tests compare the engine's treecode1
 * opening criteria (bh_params.criterion) against this file.  It is never
 * linked into, loaded by or called from the product library.
 *
 * Semantics restated (file:line under src/nbody/evolve/hackcode/treecode1):
 *   loadbody        treeload.c:139-166 (insertion; new cell midpoint = parent
 *                   midpoint +- parent size / 4; identical positions are an error)
 *   subindex        treeload.c:172-182 (upper half when midpoint <= x)
 *   hackcofm        treeload.c:189-225 (c.o.m. over subnodes in index order,
 *                   geometric centre when massless, containment check)
 *   setrcrit        treeload.c:232-252 (theta = 0: (2 rsize)^2; sw94: max
 *                   corner distance^2 / theta^2; bh86: (size/theta)^2;
 *                   default: (size/theta + |cm - centre|)^2)
 *   threadtree      treeload.c:259-276 (Next / More links)
 *   hackquad        treeload.c:284-313 (traceless quadrupole about the c.o.m.)
 *   gravcalc        treegrav.c:37-57
 *   walktree        treegrav.c:64-107 (level-by-level active list; accepted
 *                   cells append to the cell list, bodies other than the sink
 *                   to the body list, which grows downwards)
 *   accept          treegrav.c:137-156 (min distance^2 from the c.o.m. to the
 *                   sink cell > Rcrit2 and max |dk| > 1.5 psize)
 *   walksub         treegrav.c:164-188 (fan out over the sink's subnodes; a
 *                   body sink descends a virtual tree towards its position)
 *   gravsum         treegrav.c:194-214 (cells first, then bodies from the
 *                   last appended one)
 *   sumnode/sumcell treegrav.c:220-264
 *
 * Differences, on purpose: the root box is an argument (rmin, rsize) instead
 * of treecode1's expandbox around the origin (treeload.c:115-133), so that
 * the tree matches the engine's final box; every body is walked (Update
 * true); test points outside the tree are walked as body sinks that are not
 * in the tree.  Arithmetic follows inc/vectmath.h's loop forms with
 * real = double; build with -ffp-contract=off.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

enum { TC_BODY = 1, TC_CELL = 2 };

typedef struct {
    int n, nt, ncell, maxcell;
    int crit;             /* 0 default rcrit, 1 bh86, 2 sw94 */
    double theta, eps, rsize;
    int usequad;
    /* nodes: bodies [0,n), cells [n, n+maxcell), test points [n+maxcell, ...) */
    double *mass, (*pos)[3];
    int *next, *more, (*sub)[8];
    double *rcrit2, (*quad)[3][3];
    int *level;
    /* outputs per target (bodies, then test points) */
    double *phi, (*acc)[3];
    int64_t *n2b, *nbc;
    /* lists */
    int actlen;
    int *active;
    double *imass, (*ipos)[3], (*iquad)[3][3];
    int err;
} tc_tree;

static int tc_type(const tc_tree *t, int r) { return (r >= t->n && r < t->n + t->maxcell) ? TC_CELL : TC_BODY; }

static int tc_makecell(tc_tree *t)
{
    if (t->ncell >= t->maxcell) {
        t->err = 1;
        return -1;
    }
    int c = t->n + t->ncell++;
    t->mass[c] = 0.0;
    for (int i = 0; i < 8; i++) t->sub[c][i] = -1;
    return c;
}

static int tc_subindex(const tc_tree *t, const double *x, int q)
{
    int ind = 0;
    for (int k = 0; k < 3; k++)
        if (t->pos[q][k] <= x[k]) ind += 8 >> (k + 1);
    return ind;
}

static void tc_loadbody(tc_tree *t, int root, int p)
{
    int q = root;
    int qind = tc_subindex(t, t->pos[p], q);
    double qsize = t->rsize;
    int lev = 0;
    while (t->sub[q][qind] != -1) {
        int o = t->sub[q][qind];
        if (tc_type(t, o) == TC_BODY) {
            double d0 = t->pos[p][0] - t->pos[o][0], d1 = t->pos[p][1] - t->pos[o][1],
                   d2 = t->pos[p][2] - t->pos[o][2];
            double dist2 = d0 * d0;
            dist2 += d1 * d1;
            dist2 += d2 * d2;
            if (dist2 == 0.0) {
                t->err = 2;
                return;
            }
            int c = tc_makecell(t);
            if (c < 0) return;
            for (int k = 0; k < 3; k++)
                t->pos[c][k] = t->pos[q][k] + (t->pos[p][k] < t->pos[q][k] ? -qsize : qsize) / 4;
            t->level[c] = lev + 1;
            t->sub[c][tc_subindex(t, t->pos[o], c)] = o;
            t->sub[q][qind] = c;
        }
        q = t->sub[q][qind];
        qind = tc_subindex(t, t->pos[p], q);
        qsize = qsize / 2;
        lev++;
    }
    t->sub[q][qind] = p;
}

static double sqr(double x) { return x * x; }

static void tc_setrcrit(tc_tree *t, int p, const double cm[3], double psize)
{
    if (t->theta == 0.0) {
        t->rcrit2[p] = sqr(2 * t->rsize);
    } else if (t->crit == 2) {
        double bmax2 = 0.0;
        for (int k = 0; k < 3; k++) {
            double d = cm[k] - t->pos[p][k] + psize / 2;
            bmax2 += sqr(d > psize - d ? d : psize - d);
        }
        t->rcrit2[p] = bmax2 / sqr(t->theta);
    } else if (t->crit == 1) {
        t->rcrit2[p] = sqr(psize / t->theta);
    } else {
        double s = 0.0;
        for (int k = 0; k < 3; k++) s += (cm[k] - t->pos[p][k]) * (cm[k] - t->pos[p][k]);
        double d = sqrt(s);
        t->rcrit2[p] = sqr(psize / t->theta + d);
    }
}

static void tc_hackcofm(tc_tree *t, int p, double psize)
{
    double cm[3] = {0.0, 0.0, 0.0};
    t->mass[p] = 0.0;
    for (int i = 0; i < 8; i++) {
        int q = t->sub[p][i];
        if (q == -1) continue;
        if (tc_type(t, q) == TC_CELL) tc_hackcofm(t, q, psize / 2);
        t->mass[p] += t->mass[q];
        for (int k = 0; k < 3; k++) cm[k] += t->pos[q][k] * t->mass[q];
    }
    if (t->mass[p] > 0.0) {
        for (int k = 0; k < 3; k++) cm[k] = cm[k] / t->mass[p];
    } else {
        for (int k = 0; k < 3; k++) cm[k] = t->pos[p][k];
    }
    for (int k = 0; k < 3; k++)
        if (cm[k] < t->pos[p][k] - psize / 2 || t->pos[p][k] + psize / 2 <= cm[k]) t->err = 3;
    tc_setrcrit(t, p, cm, psize);
    for (int k = 0; k < 3; k++) t->pos[p][k] = cm[k];
}

static void tc_threadtree(tc_tree *t, int p, int nxt)
{
    t->next[p] = nxt;
    if (tc_type(t, p) == TC_CELL) {
        int desc[9], nd = 0;
        for (int i = 0; i < 8; i++)
            if (t->sub[p][i] != -1) desc[nd++] = t->sub[p][i];
        t->more[p] = desc[0];
        desc[nd] = nxt;
        for (int i = 0; i < nd; i++) tc_threadtree(t, desc[i], desc[i + 1]);
    }
}

/* hackquad: the moments are kept in their own array, the subnode table
 * stays intact (the reference overlays the two). */
static void tc_hackquad(tc_tree *t, int p)
{
    int desc[8], nd = 0;
    for (int i = 0; i < 8; i++)
        if (t->sub[p][i] != -1) desc[nd++] = t->sub[p][i];
    memset(t->quad[p], 0, sizeof t->quad[p]);
    for (int i = 0; i < nd; i++) {
        int q = desc[i];
        if (tc_type(t, q) == TC_CELL) tc_hackquad(t, q);
        double dr[3], drsq = 0.0, tmp[3][3];
        for (int k = 0; k < 3; k++) dr[k] = t->pos[q][k] - t->pos[p][k];
        for (int k = 0; k < 3; k++) drsq += dr[k] * dr[k];
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) {
                double idr = (a == b ? 1.0 : 0.0) * drsq;
                double v = (dr[a] * dr[b]) * 3.0;
                v = v - idr;
                v = v * t->mass[q];
                if (tc_type(t, q) == TC_CELL) v = v + t->quad[q][a][b];
                tmp[a][b] = v;
            }
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) t->quad[p][a][b] = t->quad[p][a][b] + tmp[a][b];
    }
}

static int tc_accept(const tc_tree *t, int c, double psize, const double pmid[3])
{
    double dmax = psize, dsq = 0.0;
    for (int k = 0; k < 3; k++) {
        double dk = t->pos[c][k] - pmid[k];
        if (dk < 0) dk = -dk;
        if (dk > dmax) dmax = dk;
        dk -= 0.5 * psize;
        if (dk > 0) dsq += dk * dk;
    }
    return dsq > t->rcrit2[c] && dmax > 1.5 * psize;
}

static void tc_sumnode(const tc_tree *t, int lo, int hi, const double pos0[3], double *phi0, double acc0[3])
{
    double eps2 = t->eps * t->eps;
    for (int i = lo; i < hi; i++) {
        double dr[3];
        dr[0] = t->ipos[i][0] - pos0[0];
        double dr2 = dr[0] * dr[0];
        dr[1] = t->ipos[i][1] - pos0[1];
        dr2 += dr[1] * dr[1];
        dr[2] = t->ipos[i][2] - pos0[2];
        dr2 += dr[2] * dr[2];
        dr2 += eps2;
        double drab = sqrt(dr2);
        double phi_p = t->imass[i] / drab;
        *phi0 -= phi_p;
        double mr3i = phi_p / dr2;
        for (int k = 0; k < 3; k++) acc0[k] += dr[k] * mr3i;
    }
}

static void tc_sumcell(const tc_tree *t, int lo, int hi, const double pos0[3], double *phi0, double acc0[3])
{
    double eps2 = t->eps * t->eps;
    for (int i = lo; i < hi; i++) {
        double dr[3], qdr[3];
        dr[0] = t->ipos[i][0] - pos0[0];
        double dr2 = dr[0] * dr[0];
        dr[1] = t->ipos[i][1] - pos0[1];
        dr2 += dr[1] * dr[1];
        dr[2] = t->ipos[i][2] - pos0[2];
        dr2 += dr[2] * dr[2];
        dr2 += eps2;
        double drab = sqrt(dr2);
        double phi_p = t->imass[i] / drab;
        double mr3i = phi_p / dr2;
        /* DOTPMULMV: qdr = Q dr row by row, drqdr = sum dr[k] qdr[k] */
        double drqdr = 0.0;
        for (int a = 0; a < 3; a++) {
            qdr[a] = 0.0;
            for (int b = 0; b < 3; b++) qdr[a] += t->iquad[i][a][b] * dr[b];
            drqdr += dr[a] * qdr[a];
        }
        double dr5i = 1.0 / (dr2 * dr2 * drab);
        double phi_q = 0.5 * dr5i * drqdr;
        *phi0 -= phi_p + phi_q;
        mr3i += 5.0 * phi_q / dr2;
        for (int k = 0; k < 3; k++) acc0[k] += dr[k] * mr3i + qdr[k] * -dr5i;
    }
}

static void tc_gravsum(tc_tree *t, int p, int ctop, int bbot)
{
    double phi0 = 0.0, acc0[3] = {0.0, 0.0, 0.0};
    const double *pos0 = t->pos[p];
    if (t->usequad)
        tc_sumcell(t, 0, ctop, pos0, &phi0, acc0);
    else
        tc_sumnode(t, 0, ctop, pos0, &phi0, acc0);
    tc_sumnode(t, bbot, t->actlen, pos0, &phi0, acc0);
    int o = p < t->n ? p : t->n + (p - t->n - t->maxcell);
    t->phi[o] = phi0;
    for (int k = 0; k < 3; k++) t->acc[o][k] = acc0[k];
    t->n2b[o] = t->actlen - bbot;
    t->nbc[o] = ctop;
}

static void tc_walksub(tc_tree *t, int a0, int a1, int ctop, int bbot, int p, double psize, const double pmid[3]);

/* aptr..nptr: this level's active list; cells accumulate in [0, ctop), bodies
 * in [bbot, actlen). */
static void tc_walktree(tc_tree *t, int a0, int a1, int ctop, int bbot, int p, double psize, const double pmid[3])
{
    if (t->err) return;
    int np = a1;
    for (int ap = a0; ap < a1; ap++) {
        int q = t->active[ap];
        if (tc_type(t, q) == TC_CELL) {
            if (tc_accept(t, q, psize, pmid)) {
                if (ctop >= bbot) {
                    t->err = 4;
                    return;
                }
                t->imass[ctop] = t->mass[q];
                memcpy(t->ipos[ctop], t->pos[q], sizeof t->ipos[ctop]);
                memcpy(t->iquad[ctop], t->quad[q], sizeof t->iquad[ctop]);
                ctop++;
            } else {
                for (int r = t->more[q]; r != t->next[q]; r = t->next[r]) {
                    if (np >= t->actlen) {
                        t->err = 4;
                        return;
                    }
                    t->active[np++] = r;
                }
            }
        } else if (q != p) {
            if (bbot <= ctop) {
                t->err = 4;
                return;
            }
            --bbot;
            t->imass[bbot] = t->mass[q];
            memcpy(t->ipos[bbot], t->pos[q], sizeof t->ipos[bbot]);
        }
    }
    if (np != a1)
        tc_walksub(t, a1, np, ctop, bbot, p, psize, pmid);
    else if (tc_type(t, p) == TC_CELL)
        t->err = 5;
    else
        tc_gravsum(t, p, ctop, bbot);
}

static void tc_walksub(tc_tree *t, int a0, int a1, int ctop, int bbot, int p, double psize, const double pmid[3])
{
    double poff = psize / 4, nmid[3];
    if (tc_type(t, p) == TC_CELL) {
        for (int q = t->more[p]; q != t->next[p]; q = t->next[q]) {
            for (int k = 0; k < 3; k++) nmid[k] = pmid[k] + (t->pos[q][k] < pmid[k] ? -poff : poff);
            tc_walktree(t, a0, a1, ctop, bbot, q, psize / 2, nmid);
        }
    } else {
        for (int k = 0; k < 3; k++) nmid[k] = pmid[k] + (t->pos[p][k] < pmid[k] ? -poff : poff);
        tc_walktree(t, a0, a1, ctop, bbot, p, psize / 2, nmid);
    }
}

static void tc_free(tc_tree *t)
{
    free(t->mass);
    free(t->pos);
    free(t->next);
    free(t->more);
    free(t->sub);
    free(t->rcrit2);
    free(t->quad);
    free(t->level);
    free(t->active);
    free(t->imass);
    free(t->ipos);
    free(t->iquad);
    free(t);
}

/*
 * Force on every body (self skipped) and, when nt > 0, on nt test points.
 * crit: 0 treecode1 default, 1 bh86, 2 sw94.  Outputs per target (n bodies
 * then nt test points): phi, acc[3], n2b, nbc; the cells are dumped when
 * cell_* are not NULL (DFS order of the subnode tables, at most cell_cap):
 * level, mass, c.o.m., Rcrit2.  Returns 0, or 1 cell pool exhausted,
 * 2 identical positions, 3 c.o.m. outside its cell, 4 list overflow,
 * 5 walk ended on a cell, 6 allocation.
 */
static void tc_dump(const tc_tree *t, int p, int *k, int cap, int *lev, double *m, double *cm, double *rc)
{
    if (*k < cap) {
        lev[*k] = t->level[p];
        m[*k] = t->mass[p];
        for (int j = 0; j < 3; j++) cm[3 * *k + j] = t->pos[p][j];
        rc[*k] = t->rcrit2[p];
    }
    (*k)++;
    for (int i = 0; i < 8; i++)
        if (t->sub[p][i] != -1 && tc_type(t, t->sub[p][i]) == TC_CELL) tc_dump(t, t->sub[p][i], k, cap, lev, m, cm, rc);
}

int tc_force(int n, const double *mass, const double *pos, int nt, const double *tpos, const double *rmin,
             double rsize, double theta, double eps, int crit, int usequad, double *phi, double *acc,
             int64_t *n2b, int64_t *nbc, int64_t *ncell, int cell_cap, int *cell_level, double *cell_mass,
             double *cell_cm, double *cell_rcrit2)
{
    tc_tree *t = (tc_tree *)calloc(1, sizeof *t);
    if (!t) return 6;
    t->n = n;
    t->nt = nt;
    t->maxcell = 2 * n + 8;
    t->crit = crit;
    t->theta = theta;
    t->eps = eps;
    t->rsize = rsize;
    t->usequad = usequad;
    int nnode = n + t->maxcell + nt;
    t->mass = (double *)calloc(nnode, sizeof(double));
    t->pos = (double(*)[3])calloc(nnode, sizeof *t->pos);
    t->next = (int *)calloc(nnode, sizeof(int));
    t->more = (int *)calloc(nnode, sizeof(int));
    t->sub = (int(*)[8])calloc(nnode, sizeof *t->sub);
    t->rcrit2 = (double *)calloc(nnode, sizeof(double));
    t->quad = (double(*)[3][3])calloc(nnode, sizeof *t->quad);
    t->level = (int *)calloc(nnode, sizeof(int));
    t->actlen = nnode + 64;
    t->active = (int *)calloc(t->actlen, sizeof(int));
    t->imass = (double *)calloc(t->actlen, sizeof(double));
    t->ipos = (double(*)[3])calloc(t->actlen, sizeof *t->ipos);
    t->iquad = (double(*)[3][3])calloc(t->actlen, sizeof *t->iquad);
    if (!t->mass || !t->pos || !t->next || !t->more || !t->sub || !t->rcrit2 || !t->quad || !t->level ||
        !t->active || !t->imass || !t->ipos || !t->iquad) {
        tc_free(t);
        return 6;
    }
    t->phi = phi;
    t->acc = (double(*)[3])acc;
    t->n2b = n2b;
    t->nbc = nbc;
    for (int i = 0; i < n; i++) {
        t->mass[i] = mass[i];
        for (int k = 0; k < 3; k++) t->pos[i][k] = pos[3 * i + k];
    }
    for (int i = 0; i < nt; i++)
        for (int k = 0; k < 3; k++) t->pos[n + t->maxcell + i][k] = tpos[3 * i + k];
    int root = tc_makecell(t);
    double rmid[3];
    for (int k = 0; k < 3; k++) rmid[k] = t->pos[root][k] = rmin[k] + 0.5 * rsize;
    t->level[root] = 0;
    for (int i = 0; i < n && !t->err; i++) tc_loadbody(t, root, i);
    if (!t->err) tc_hackcofm(t, root, rsize);
    if (!t->err) tc_threadtree(t, root, -1);
    if (!t->err && usequad) tc_hackquad(t, root);
    if (!t->err) {
        t->active[0] = root;
        tc_walktree(t, 0, 1, 0, t->actlen, root, rsize, rmid);
        for (int i = 0; i < nt && !t->err; i++)
            tc_walktree(t, 0, 1, 0, t->actlen, n + t->maxcell + i, rsize, rmid);
    }
    if (ncell) *ncell = t->ncell;
    if (!t->err && cell_level) {
        int k = 0;
        tc_dump(t, root, &k, cell_cap, cell_level, cell_mass, cell_cm, cell_rcrit2);
    }
    int err = t->err;
    tc_free(t);
    return err;
}
