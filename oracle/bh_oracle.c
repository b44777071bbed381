/*
 * bh_oracle.c -- CPU restatement of NEMO's hackcode1 Barnes-Hut force path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker the tests, the smoke
 * check and the CPU-baseline leg of bench.py compare the HIP engine against.
 * It is never linked into, loaded by or called from the product library
 * (nemo_amd/libnemo_bh.so) or the hackforce CLI.
 *
 * Semantics restated (file:line under the reference tree):
 *   tree build      src/nbody/evolve/hackcode/hackcode1/load.c:32-62  (maketree)
 *   box expansion   load.c:68-95   (expandbox; order dependent, graft old root)
 *   insertion       load.c:101-128 (loadtree: one body per leaf, split on collision)
 *   integer coords  load.c:135-156 (intcoord: floor(2^30 (x-rmin)/rsize), half-open)
 *   child index     load.c:162-177 (subindex: 4x+2y+z)
 *   moments         load.c:183-226 (hackcofm: FP64 c.o.m., traceless quadrupole)
 *   walk            grav.c:82-115  (hackwalk/walksub: DFS, dsq/4 per level)
 *   opening test    grav.c:122-131 (subdivp: tol^2 |com-x|^2 < size^2, bodies never open)
 *   interaction     grav.c:42-73   (gravsub: softened monopole + quadrupole)
 *   force driver    hackforce.c:177-215 (force_calc)
 *   leapfrog        code.c:159-191 (stepsystem) with box init code.c:119-120
 *   diagnostics     code_io.c:351-388 (diagnostics)
 *   direct sum      src/nbody/evolve/directcode/grav.c:23-65 (eps >= 0 branch, gravc = 1)
 *
 * Arithmetic follows the operation order of inc/vectmath.h (DOTVP, MULMV,
 * OUTVP, ...) with real = double and no fused multiply-add: build with
 * -ffp-contract=off.  Node storage is index based (cells in a pool, bodies
 * by index) instead of the reference's pointer tree; the traversal order,
 * the rounding sequence and the counters are the same.
 *
 * Pinning: every function here is checked by tests/test_oracle.py against
 * the reference's own fixtures (data/snapshot/hack128.sun4 Energy and
 * PETensor, man/man1/hackcode1.1:137-156 interaction totals 3864 and 4899,
 * src/nbody/evolve/hackcode/hackcode1/Testfile:27 bsf line) and against
 * mathematical closed forms (tol=0 equals the direct sum, two-body closed
 * form, point-set quadrupole, Newton's third law at tol=0).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_IMAX (1 << 30)
#define ORC_NULL (-1)

/* node reference: >= 0 body index, <= -2 cell index c encoded as -(c+2) */
#define IS_BODY(r) ((r) >= 0)
#define IS_CELL(r) ((r) <= -2)
#define CELL_OF(r) (-(r)-2)
#define CELL_REF(c) (-(c)-2)

typedef struct {
    double mass;
    double pos[3];
    int sub[8];
    double quad[3][3];
    int bhead, btail, nbk;   /* depth-capped trees: bodies of a bucket cell, in insertion order */
} orc_cell;

typedef struct {
    int ncell, maxcell;
    orc_cell *cells;
    int root;            /* node ref or ORC_NULL */
    double rmin[3];
    double rsize;
    const double *bmass; /* borrowed body arrays */
    const double *bpos;  /* [n][3] */
    int usequad;
    int err;             /* 0 ok, 1 cell pool exhausted, 2 out of bits, 3 non-finite */
    int depth_cap;       /* 0: the reference's tree; > 0: cells stop at this level (buckets) */
    int *bnext;          /* bucket lists: next body, -1 at the end */
    int nbucket;         /* bucket cells (cells at level depth_cap) */
} orc_tree;

static int orc_intcoord(const orc_tree *t, int xp[3], const double rp[3])
{
    int inb = 1;
    for (int k = 0; k < 3; k++) {
        double xsc = (rp[k] - t->rmin[k]) / t->rsize;
        if (0.0 <= xsc && xsc < 1.0)
            xp[k] = (int)floor(ORC_IMAX * xsc);
        else
            inb = 0;
    }
    return inb;
}

static int orc_subindex(const int x[3], int l)
{
    int i = 0;
    for (int k = 0; k < 3; k++)
        if (x[k] & l) i += 8 >> (k + 1);
    return i;
}

static int orc_makecell(orc_tree *t)
{
    if (t->ncell >= t->maxcell) { t->err = 1; return -1; }
    int c = t->ncell++;
    for (int i = 0; i < 8; i++) t->cells[c].sub[i] = ORC_NULL;
    t->cells[c].bhead = t->cells[c].btail = -1;
    t->cells[c].nbk = 0;
    return c;
}

static void orc_expandbox(orc_tree *t, const double *pos)
{
    int xtmp[3], xmid[3];
    double rmid[3];
    int guard = 0;
    while (!orc_intcoord(t, xtmp, pos)) {
        if (!(isfinite(pos[0]) && isfinite(pos[1]) && isfinite(pos[2])) || ++guard > 2000) {
            t->err = 3;   /* the reference loops forever here; report instead */
            return;
        }
        for (int k = 0; k < 3; k++) rmid[k] = t->rmin[k] + 0.5 * t->rsize;
        for (int k = 0; k < 3; k++)
            if (pos[k] < rmid[k]) t->rmin[k] -= t->rsize;
        t->rsize = 2.0 * t->rsize;
        if (t->root != ORC_NULL) {
            int c = orc_makecell(t);
            if (c < 0) return;
            orc_intcoord(t, xmid, rmid);
            int k = orc_subindex(xmid, ORC_IMAX >> 1);
            t->cells[c].sub[k] = t->root;
            t->root = CELL_REF(c);
        }
    }
}

static void orc_loadtree(orc_tree *t, int b)
{
    int xp[3], xq[3];
    orc_intcoord(t, xp, t->bpos + 3 * (size_t)b);
    int l = ORC_IMAX >> 1;
    int *qptr = &t->root;
    while (*qptr != ORC_NULL) {
        if (l == 0) { t->err = 2; return; }   /* coincident bodies: reference asserts */
        if (IS_BODY(*qptr)) {
            int c = orc_makecell(t);
            if (c < 0) return;
            orc_intcoord(t, xq, t->bpos + 3 * (size_t)(*qptr));
            t->cells[c].sub[orc_subindex(xq, l)] = *qptr;
            *qptr = CELL_REF(c);
        }
        qptr = &t->cells[CELL_OF(*qptr)].sub[orc_subindex(xp, l)];
        l = l >> 1;
    }
    *qptr = b;
}

/* Depth-capped insertion: the device's tree (63-bit keys, 21 levels below
 * the root).  The same descent as loadtree, except that a collision in a slot
 * of level depth_cap makes that slot a bucket cell, and a body reaching a
 * bucket is appended to it, so the bodies of a bucket are its children in
 * insertion (input) order.  Levels count from the root cell (level 0), so
 * the caller inserts with the final box (no expandbox grafts afterwards). */
static void orc_loadtree_capped(orc_tree *t, int b)
{
    int xp[3], xq[3];
    orc_intcoord(t, xp, t->bpos + 3 * (size_t)b);
    int l = ORC_IMAX >> 1, level = 0;
    int *qptr = &t->root;
    while (*qptr != ORC_NULL) {
        if (IS_CELL(*qptr) && t->cells[CELL_OF(*qptr)].nbk > 0) {
            orc_cell *c = &t->cells[CELL_OF(*qptr)];
            t->bnext[c->btail] = b;
            t->bnext[b] = -1;
            c->btail = b;
            c->nbk++;
            return;
        }
        if (IS_BODY(*qptr)) {
            int c = orc_makecell(t);
            if (c < 0) return;
            if (level == t->depth_cap) {
                orc_cell *cc = &t->cells[c];
                cc->bhead = *qptr;
                t->bnext[*qptr] = b;
                t->bnext[b] = -1;
                cc->btail = b;
                cc->nbk = 2;
                t->nbucket++;
                *qptr = CELL_REF(c);
                return;
            }
            orc_intcoord(t, xq, t->bpos + 3 * (size_t)(*qptr));
            t->cells[c].sub[orc_subindex(xq, l)] = *qptr;
            *qptr = CELL_REF(c);
        }
        qptr = &t->cells[CELL_OF(*qptr)].sub[orc_subindex(xp, l)];
        l = l >> 1;
        level++;
    }
    *qptr = b;
}

static double node_mass(const orc_tree *t, int r)
{
    return IS_BODY(r) ? t->bmass[r] : t->cells[CELL_OF(r)].mass;
}

static const double *node_pos(const orc_tree *t, int r)
{
    return IS_BODY(r) ? t->bpos + 3 * (size_t)r : t->cells[CELL_OF(r)].pos;
}

/* hackcofm's sums over a bucket's bodies in list order (the device sums a
 * bucket's bodies in sorted order, which for equal keys is input order). */
static void bucket_cofm(orc_tree *t, orc_cell *c)
{
    c->mass = 0.0;
    c->pos[0] = c->pos[1] = c->pos[2] = 0.0;
    for (int r = c->bhead; r >= 0; r = t->bnext[r]) {
        double mr = t->bmass[r];
        const double *pr = t->bpos + 3 * (size_t)r;
        c->mass += mr;
        double tmpv[3] = {pr[0] * mr, pr[1] * mr, pr[2] * mr};
        c->pos[0] = c->pos[0] + tmpv[0];
        c->pos[1] = c->pos[1] + tmpv[1];
        c->pos[2] = c->pos[2] + tmpv[2];
    }
    for (int k = 0; k < 3; k++) c->pos[k] = c->pos[k] / c->mass;
    if (!t->usequad) return;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) c->quad[i][j] = 0.0;
    for (int r = c->bhead; r >= 0; r = t->bnext[r]) {
        const double *pr = t->bpos + 3 * (size_t)r;
        double dr[3] = {pr[0] - c->pos[0], pr[1] - c->pos[1], pr[2] - c->pos[2]};
        double drsq = dr[0] * dr[0];
        drsq += dr[1] * dr[1];
        drsq += dr[2] * dr[2];
        double mr = t->bmass[r];
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) {
                double idr = (a == b ? 1.0 : 0.0) * drsq;
                double tm = (dr[a] * dr[b]) * 3.0;
                tm = tm - idr;
                tm = tm * mr;
                c->quad[a][b] = c->quad[a][b] + tm;
            }
    }
}

static void orc_cofm(orc_tree *t, int q)
{
    if (!IS_CELL(q)) return;
    orc_cell *c = &t->cells[CELL_OF(q)];
    if (c->nbk > 0) {
        bucket_cofm(t, c);
        return;
    }
    c->mass = 0.0;
    c->pos[0] = c->pos[1] = c->pos[2] = 0.0;
    for (int i = 0; i < 8; i++) {
        int r = c->sub[i];
        if (r == ORC_NULL) continue;
        orc_cofm(t, r);
        double mr = node_mass(t, r);
        const double *pr = node_pos(t, r);
        c->mass += mr;
        double tmpv[3] = {pr[0] * mr, pr[1] * mr, pr[2] * mr};
        c->pos[0] = c->pos[0] + tmpv[0];
        c->pos[1] = c->pos[1] + tmpv[1];
        c->pos[2] = c->pos[2] + tmpv[2];
    }
    for (int k = 0; k < 3; k++) c->pos[k] = c->pos[k] / c->mass;
    if (!t->usequad) return;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) c->quad[i][j] = 0.0;
    for (int i = 0; i < 8; i++) {
        int r = c->sub[i];
        if (r == ORC_NULL) continue;
        const double *pr = node_pos(t, r);
        double dr[3] = {pr[0] - c->pos[0], pr[1] - c->pos[1], pr[2] - c->pos[2]};
        double drsq = dr[0] * dr[0];
        drsq += dr[1] * dr[1];
        drsq += dr[2] * dr[2];
        double mr = node_mass(t, r);
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) {
                double idr = (a == b ? 1.0 : 0.0) * drsq;
                double tm = (dr[a] * dr[b]) * 3.0;
                tm = tm - idr;
                tm = tm * mr;
                if (IS_CELL(r)) tm = tm + t->cells[CELL_OF(r)].quad[a][b];
                c->quad[a][b] = c->quad[a][b] + tm;
            }
    }
}

/* ---- public tree API ---------------------------------------------------- */

/* Build the tree over the massive bodies, in input order (load.c:32-62).
 * rmin[3]/rsize are in/out: the box only grows.  maxcell is the cell pool
 * size (fcells*nbody in the reference).  depth_cap 0 builds the reference's
 * tree; depth_cap > 0 first runs expandbox over every massive body (the box
 * sequence does not depend on the tree) and then inserts with the final box,
 * capping cells at that level (see orc_loadtree_capped).  Returns 0 or an
 * error code. */
static int tree_build(orc_tree **out, int n, const double *mass, const double *pos, double *rmin, double *rsize,
                      int maxcell, int usequad, int depth_cap)
{
    orc_tree *t = (orc_tree *)calloc(1, sizeof(orc_tree));
    t->maxcell = maxcell;
    t->cells = (orc_cell *)calloc((size_t)(maxcell > 0 ? maxcell : 1), sizeof(orc_cell));
    t->root = ORC_NULL;
    memcpy(t->rmin, rmin, 3 * sizeof(double));
    t->rsize = *rsize;
    t->bmass = mass;
    t->bpos = pos;
    t->usequad = usequad;
    t->depth_cap = depth_cap;
    if (depth_cap > 0) {
        t->bnext = (int *)malloc((size_t)(n > 0 ? n : 1) * sizeof(int));
        for (int b = 0; b < n && !t->err; b++)
            if (mass[b] != 0.0) orc_expandbox(t, pos + 3 * (size_t)b);
        for (int b = 0; b < n && !t->err; b++)
            if (mass[b] != 0.0) orc_loadtree_capped(t, b);
    } else {
        for (int b = 0; b < n && !t->err; b++) {
            if (mass[b] != 0.0) {
                orc_expandbox(t, pos + 3 * (size_t)b);
                if (t->err) break;
                orc_loadtree(t, b);
            }
        }
    }
    if (!t->err && t->root != ORC_NULL) orc_cofm(t, t->root);
    memcpy(rmin, t->rmin, 3 * sizeof(double));
    *rsize = t->rsize;
    *out = t;
    return t->err;
}

int orc_tree_build(orc_tree **out, int n, const double *mass, const double *pos,
                   double *rmin, double *rsize, int maxcell, int usequad)
{
    return tree_build(out, n, mass, pos, rmin, rsize, maxcell, usequad, 0);
}

int orc_tree_build_capped(orc_tree **out, int n, const double *mass, const double *pos,
                          double *rmin, double *rsize, int maxcell, int usequad, int depth_cap)
{
    return tree_build(out, n, mass, pos, rmin, rsize, maxcell, usequad, depth_cap);
}

void orc_tree_free(orc_tree *t)
{
    if (!t) return;
    free(t->cells);
    free(t->bnext);
    free(t);
}

int orc_tree_ncell(const orc_tree *t) { return t->ncell; }
int orc_tree_root(const orc_tree *t) { return t->root; }
int orc_tree_nbucket(const orc_tree *t) { return t->nbucket; }

/* Export cells in DFS pre-order (children 0..7): level, mass, com, quad[9],
 * and the body count under each cell.  Returns the number written. */
typedef struct { double *mass, *pos, *quad; int *level, *nbody; int k; } orc_dump;

static int dump_rec(const orc_tree *t, int r, int level, orc_dump *d)
{
    if (IS_BODY(r)) return 1;
    const orc_cell *c = &t->cells[CELL_OF(r)];
    int me = d->k++;
    d->mass[me] = c->mass;
    for (int k = 0; k < 3; k++) d->pos[3 * me + k] = c->pos[k];
    for (int a = 0; a < 9; a++) d->quad[9 * me + a] = t->usequad ? c->quad[a / 3][a % 3] : 0.0;
    d->level[me] = level;
    int nb = c->nbk;
    for (int i = 0; i < 8; i++)
        if (c->sub[i] != ORC_NULL) nb += dump_rec(t, c->sub[i], level + 1, d);
    d->nbody[me] = nb;
    return nb;
}

int orc_tree_dump(const orc_tree *t, double *mass, double *pos, double *quad, int *level, int *nbody)
{
    orc_dump d = {mass, pos, quad, level, nbody, 0};
    if (t->root != ORC_NULL) dump_rec(t, t->root, 0, &d);
    return d.k;
}

/* ---- walk (grav.c) ------------------------------------------------------- */

typedef struct {
    const orc_tree *t;
    int skip;           /* body index to skip, or -1 (test particle) */
    double pos0[3];
    double phi0, acc0[3];
    double tolsq, eps;
    int n2b, nbc;
} orc_walk;

static void orc_gravsub(orc_walk *w, int r, const double *drm, double drsqm, int memo)
{
    const orc_tree *t = w->t;
    double dr[3], drsq;
    const double *p = node_pos(t, r);
    if (memo) {
        dr[0] = drm[0]; dr[1] = drm[1]; dr[2] = drm[2]; drsq = drsqm;
    } else {
        dr[0] = p[0] - w->pos0[0];
        dr[1] = p[1] - w->pos0[1];
        dr[2] = p[2] - w->pos0[2];
        drsq = dr[0] * dr[0];
        drsq += dr[1] * dr[1];
        drsq += dr[2] * dr[2];
    }
    drsq += w->eps * w->eps;
    double drabs = sqrt(drsq);
    double phii = node_mass(t, r) / drabs;
    w->phi0 -= phii;
    double mor3 = phii / drsq;
    double ai[3] = {dr[0] * mor3, dr[1] * mor3, dr[2] * mor3};
    w->acc0[0] = w->acc0[0] + ai[0];
    w->acc0[1] = w->acc0[1] + ai[1];
    w->acc0[2] = w->acc0[2] + ai[2];
    if (t->usequad && IS_CELL(r)) {
        const orc_cell *c = &t->cells[CELL_OF(r)];
        double dr5inv = 1.0 / (drsq * drsq * drabs);
        double qdr[3];
        for (int i = 0; i < 3; i++) {
            qdr[i] = 0.0;
            for (int j = 0; j < 3; j++) qdr[i] += c->quad[i][j] * dr[j];
        }
        double drqdr = dr[0] * qdr[0];
        drqdr += dr[1] * qdr[1];
        drqdr += dr[2] * qdr[2];
        double phiq = -0.5 * dr5inv * drqdr;
        w->phi0 = w->phi0 + phiq;
        phiq = 5.0 * phiq / drsq;
        for (int k = 0; k < 3; k++) ai[k] = dr[k] * phiq;
        for (int k = 0; k < 3; k++) w->acc0[k] = w->acc0[k] - ai[k];
        for (int k = 0; k < 3; k++) qdr[k] = qdr[k] * dr5inv;
        for (int k = 0; k < 3; k++) w->acc0[k] = w->acc0[k] - qdr[k];
    }
}

static void orc_walksub(orc_walk *w, int r, double dsq)
{
    const orc_tree *t = w->t;
    if (IS_CELL(r)) {
        const double *p = node_pos(t, r);
        double dr[3] = {p[0] - w->pos0[0], p[1] - w->pos0[1], p[2] - w->pos0[2]};
        double drsq = dr[0] * dr[0];
        drsq += dr[1] * dr[1];
        drsq += dr[2] * dr[2];
        if (w->tolsq * drsq < dsq) {
            const orc_cell *c = &t->cells[CELL_OF(r)];
            for (int b = c->bhead; b >= 0; b = t->bnext[b]) orc_walksub(w, b, dsq / 4.0);
            for (int k = 0; k < 8; k++)
                if (c->sub[k] != ORC_NULL) orc_walksub(w, c->sub[k], dsq / 4.0);
            return;
        }
        orc_gravsub(w, r, dr, drsq, 1);
        w->nbc++;
        return;
    }
    if (r != w->skip) {
        orc_gravsub(w, r, NULL, 0.0, 0);
        w->n2b++;
    }
}

/* Field at one point (hackgrav, grav.c:22-32).  skip = body index excluded
 * (pointer identity in the reference), or -1. */
void orc_tree_grav(const orc_tree *t, const double *pos0, int skip, double tol, double eps,
                   double *phi, double *acc, int *n2b, int *nbc)
{
    orc_walk w;
    memset(&w, 0, sizeof w);
    w.t = t;
    w.skip = skip;
    w.pos0[0] = pos0[0]; w.pos0[1] = pos0[1]; w.pos0[2] = pos0[2];
    w.tolsq = tol * tol;
    w.eps = eps;
    if (t->root != ORC_NULL) orc_walksub(&w, t->root, t->rsize * t->rsize);
    *phi = w.phi0;
    acc[0] = w.acc0[0]; acc[1] = w.acc0[1]; acc[2] = w.acc0[2];
    *n2b = w.n2b;
    *nbc = w.nbc;
}

/* hackforce force_calc (hackforce.c:177-215).  ntest/tpos NULL => the mass
 * set is the test set and each body skips itself.  Counters are 64-bit
 * totals (the reference's are 32-bit ints).  depth_cap as tree_build; the
 * bucket count goes to *nbucket when it is not NULL.  Returns the tree
 * error. */
int orc_force_capped(int n, const double *mass, const double *pos, int ntest, const double *tpos,
                     double tol, double eps, int usequad, double *rmin, double *rsize, double fcells,
                     int depth_cap, double *phi, double *acc, int64_t *n2btot, int64_t *nbctot, int64_t *nbucket)
{
    orc_tree *t;
    int maxcell = (int)(fcells * n);
    int err = tree_build(&t, n, mass, pos, rmin, rsize, maxcell, usequad, depth_cap);
    *n2btot = *nbctot = 0;
    if (nbucket) *nbucket = t->nbucket;
    if (err) { orc_tree_free(t); return err; }
    int self = (tpos == NULL);
    if (self) { ntest = n; tpos = pos; }
    for (int i = 0; i < ntest; i++) {
        int a, b;
        orc_tree_grav(t, tpos + 3 * (size_t)i, self ? i : -1, tol, eps, phi + i, acc + 3 * (size_t)i, &a, &b);
        *n2btot += a;
        *nbctot += b;
    }
    orc_tree_free(t);
    return 0;
}

int orc_force(int n, const double *mass, const double *pos, int ntest, const double *tpos,
              double tol, double eps, int usequad, double *rmin, double *rsize, double fcells,
              double *phi, double *acc, int64_t *n2btot, int64_t *nbctot)
{
    return orc_force_capped(n, mass, pos, ntest, tpos, tol, eps, usequad, rmin, rsize, fcells, 0, phi, acc,
                            n2btot, nbctot, NULL);
}

/* FP64 direct sum (directcode/grav.c:23-65, eps >= 0 branch, gravc = 1).
 * Evaluates targets idx[0..nt) (indices into the same body set); each
 * skips itself.  idx == NULL means all n. */
void orc_direct(int n, const double *mass, const double *pos, int nt, const int64_t *idx,
                double eps, double *phi, double *acc)
{
    if (!idx) nt = n;
    for (int t = 0; t < nt; t++) {
        int64_t i = idx ? idx[t] : t;
        const double *p0 = pos + 3 * i;
        double phi0 = 0.0, a0[3] = {0.0, 0.0, 0.0};
        for (int j = 0; j < n; j++) {
            if (j == i) continue;
            const double *p = pos + 3 * (size_t)j;
            double dr[3] = {p[0] - p0[0], p[1] - p0[1], p[2] - p0[2]};
            double drsq = dr[0] * dr[0];
            drsq += dr[1] * dr[1];
            drsq += dr[2] * dr[2];
            drsq += eps * eps;
            double drabs = sqrt(drsq);
            double phii = mass[j] / drabs;
            phi0 -= phii;
            double mor3 = phii / drsq;
            a0[0] = a0[0] + dr[0] * mor3;
            a0[1] = a0[1] + dr[1] * mor3;
            a0[2] = a0[2] + dr[2] * mor3;
        }
        phi[t] = phi0;
        acc[3 * t + 0] = a0[0];
        acc[3 * t + 1] = a0[1];
        acc[3 * t + 2] = a0[2];
    }
}

/* ---- hackcode1 integrator (code.c:159-191) and diagnostics --------------- */

/* Diagnostics (code_io.c:351-388): out[0..36] = Energy[3], KETensor[9],
 * PETensor[9], AMTensor[9], CMPhaseSpace[6]. */
void orc_diagnostics(int n, const double *mass, const double *pos, const double *vel,
                     const double *acc, const double *phi, double *out)
{
    double mtot = 0.0, etot[3] = {0, 0, 0};
    double ke[3][3] = {{0}}, pe[3][3] = {{0}}, am[3][3] = {{0}}, cm[2][3] = {{0}};
    for (int b = 0; b < n; b++) {
        const double *x = pos + 3 * (size_t)b, *v = vel + 3 * (size_t)b, *a = acc + 3 * (size_t)b;
        double m = mass[b];
        mtot += m;
        double velsq = v[0] * v[0];
        velsq += v[1] * v[1];
        velsq += v[2] * v[2];
        etot[1] += 0.5 * m * velsq;
        etot[2] += 0.5 * m * phi[b];
        double tv[3];
        for (int k = 0; k < 3; k++) tv[k] = v[k] * (0.5 * m);
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) ke[i][j] = ke[i][j] + tv[i] * v[j];
        for (int k = 0; k < 3; k++) tv[k] = x[k] * m;
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) pe[i][j] = pe[i][j] + tv[i] * a[j];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) am[i][j] = am[i][j] + tv[i] * v[j];
        for (int k = 0; k < 3; k++) tv[k] = x[k] * m;
        for (int k = 0; k < 3; k++) cm[0][k] = cm[0][k] + tv[k];
        for (int k = 0; k < 3; k++) tv[k] = v[k] * m;
        for (int k = 0; k < 3; k++) cm[1][k] = cm[1][k] + tv[k];
    }
    etot[0] = etot[1] + etot[2];
    double tr[3][3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) tr[i][j] = am[j][i];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) am[i][j] = am[i][j] - tr[i][j];
    for (int k = 0; k < 3; k++) { cm[0][k] = cm[0][k] / mtot; cm[1][k] = cm[1][k] / mtot; }
    memcpy(out, etot, 3 * sizeof(double));
    memcpy(out + 3, ke, 9 * sizeof(double));
    memcpy(out + 12, pe, 9 * sizeof(double));
    memcpy(out + 21, am, 9 * sizeof(double));
    memcpy(out + 30, cm, 6 * sizeof(double));
}

/* stepsystem() (code.c:159-191) split at the point where output() runs:
 * orc_step_force builds the tree, evaluates every body and applies the
 * second-order velocity correction when nstep > 0; the caller may then read
 * the output-time state; orc_step_drift does the half-kick/drift/half-kick
 * and increments nstep.  Counters of the force pass go to n2b/nbc. */
int orc_step_force(int n, const double *mass, const double *pos, double *vel, double *acc, double *phi,
                   int nstep, double dt, double tol, double eps, int usequad,
                   double *rmin, double *rsize, double fcells, int64_t *n2b, int64_t *nbc)
{
    double dthf = 0.5 * dt;
    orc_tree *t;
    int err = orc_tree_build(&t, n, mass, pos, rmin, rsize, (int)(fcells * n), usequad);
    if (err) { orc_tree_free(t); return err; }
    *n2b = *nbc = 0;
    for (int b = 0; b < n; b++) {
        double a1[3] = {acc[3 * b], acc[3 * b + 1], acc[3 * b + 2]};
        int x, y;
        orc_tree_grav(t, pos + 3 * (size_t)b, b, tol, eps, phi + b, acc + 3 * (size_t)b, &x, &y);
        *n2b += x;
        *nbc += y;
        if (nstep > 0) {
            for (int k = 0; k < 3; k++) {
                double dacc = acc[3 * b + k] - a1[k];
                double dvel = dacc * dthf;
                vel[3 * b + k] = vel[3 * b + k] + dvel;
            }
        }
    }
    orc_tree_free(t);
    return 0;
}

void orc_step_drift(int n, double *pos, double *vel, const double *acc, double dt)
{
    double dthf = 0.5 * dt;
    for (int b = 0; b < n; b++) {
        double vel1[3];
        for (int k = 0; k < 3; k++) {
            double dvel = acc[3 * b + k] * dthf;
            vel1[k] = vel[3 * b + k] + dvel;
        }
        for (int k = 0; k < 3; k++) {
            double dpos = vel1[k] * dt;
            pos[3 * b + k] = pos[3 * b + k] + dpos;
        }
        for (int k = 0; k < 3; k++) {
            double dvel = acc[3 * b + k] * dthf;
            vel[3 * b + k] = vel1[k] + dvel;
        }
    }
}

/* ---- Morton keys (the CPU reference for the device key kernel) ----------- */

/* 63-bit key: 21 bits per dimension, x the most significant bit of each
 * triple, from the 30-bit integer coordinates of intcoord() shifted right
 * by 9 (floor(2^21 xsc) == floor(2^30 xsc) >> 9).  Coordinates outside the
 * box are clamped to the box (cannot happen for bodies loaded in the tree). */
void orc_morton(int n, const double *pos, const double *rmin, double rsize, uint64_t *keys)
{
    for (int b = 0; b < n; b++) {
        uint64_t c[3];
        for (int k = 0; k < 3; k++) {
            double xsc = (pos[3 * (size_t)b + k] - rmin[k]) / rsize;
            int64_t xi;
            if (!(xsc >= 0.0)) xi = 0;
            else if (!(xsc < 1.0)) xi = ORC_IMAX - 1;
            else xi = (int64_t)floor(ORC_IMAX * xsc);
            c[k] = (uint64_t)(xi >> 9);
        }
        uint64_t key = 0;
        for (int l = 20; l >= 0; l--) {
            key = (key << 3) | (((c[0] >> l) & 1) << 2) | (((c[1] >> l) & 1) << 1) | ((c[2] >> l) & 1);
        }
        keys[b] = key;
    }
}

/* ---- timed sample for the CPU baseline ----------------------------------- */
#include <time.h>

static double now_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

/* Build the full tree (timed), then evaluate the bodies idx[0..ns) of the
 * mass set, each skipping itself (timed), on nthreads host threads sharing
 * the read-only tree (each thread walks a contiguous slice of idx).  The
 * caller extrapolates per-body walk time to the whole set. */
#include <pthread.h>

typedef struct {
    const orc_tree *t;
    const double *pos;
    const int64_t *idx;
    int k0, k1;
    double tol, eps;
    double *phi, *acc;
    int64_t n2b, nbc;
} orc_slice;

static void *orc_walk_slice(void *arg)
{
    orc_slice *s = (orc_slice *)arg;
    s->n2b = s->nbc = 0;
    for (int k = s->k0; k < s->k1; k++) {
        int a, b;
        int64_t i = s->idx[k];
        orc_tree_grav(s->t, s->pos + 3 * i, (int)i, s->tol, s->eps, s->phi + k, s->acc + 3 * (size_t)k, &a, &b);
        s->n2b += a;
        s->nbc += b;
    }
    return NULL;
}

int orc_force_sample_mt(int n, const double *mass, const double *pos, int ns, const int64_t *idx,
                        double tol, double eps, int usequad, double *rmin, double *rsize, double fcells,
                        int nthreads, double *phi, double *acc, int64_t *n2btot, int64_t *nbctot,
                        double *t_tree, double *t_walk)
{
    enum { MAXT = 256 };
    if (nthreads < 1) nthreads = 1;
    if (nthreads > MAXT) nthreads = MAXT;
    orc_tree *t;
    double t0 = now_s();
    int err = orc_tree_build(&t, n, mass, pos, rmin, rsize, (int)(fcells * n), usequad);
    double t1 = now_s();
    *n2btot = *nbctot = 0;
    if (err) { orc_tree_free(t); return err; }
    orc_slice sl[MAXT];
    pthread_t th[MAXT];
    int started[MAXT] = {0};
    for (int j = 0; j < nthreads; j++) {
        sl[j] = (orc_slice){t, pos, idx, (int)((int64_t)ns * j / nthreads), (int)((int64_t)ns * (j + 1) / nthreads),
                            tol, eps, phi, acc, 0, 0};
        if (j > 0) started[j] = pthread_create(&th[j], NULL, orc_walk_slice, &sl[j]) == 0;
    }
    /* slice 0, and any slice whose thread could not start, runs on this thread */
    for (int j = 0; j < nthreads; j++)
        if (!started[j]) orc_walk_slice(&sl[j]);
    for (int j = 1; j < nthreads; j++)
        if (started[j]) pthread_join(th[j], NULL);
    double t2 = now_s();
    for (int j = 0; j < nthreads; j++) {
        *n2btot += sl[j].n2b;
        *nbctot += sl[j].nbc;
    }
    *t_tree = t1 - t0;
    *t_walk = t2 - t1;
    orc_tree_free(t);
    return 0;
}

/* Walk the bodies idx[0..ns) of the mass set over a built tree (each skips
 * itself) on nthreads host threads; seconds of the walk in *t_walk.  Lets a
 * caller try several tol on one tree. */
int orc_tree_walk_idx(const orc_tree *t, const double *pos, int ns, const int64_t *idx, double tol, double eps,
                      int nthreads, double *phi, double *acc, int64_t *n2btot, int64_t *nbctot, double *t_walk)
{
    enum { MAXT = 256 };
    if (nthreads < 1) nthreads = 1;
    if (nthreads > MAXT) nthreads = MAXT;
    orc_slice sl[MAXT];
    pthread_t th[MAXT];
    int started[MAXT] = {0};
    double t1 = now_s();
    for (int j = 0; j < nthreads; j++) {
        sl[j] = (orc_slice){t, pos, idx, (int)((int64_t)ns * j / nthreads), (int)((int64_t)ns * (j + 1) / nthreads),
                            tol, eps, phi, acc, 0, 0};
        if (j > 0) started[j] = pthread_create(&th[j], NULL, orc_walk_slice, &sl[j]) == 0;
    }
    for (int j = 0; j < nthreads; j++)
        if (!started[j]) orc_walk_slice(&sl[j]);
    for (int j = 1; j < nthreads; j++)
        if (started[j]) pthread_join(th[j], NULL);
    *t_walk = now_s() - t1;
    *n2btot = *nbctot = 0;
    for (int j = 0; j < nthreads; j++) {
        *n2btot += sl[j].n2b;
        *nbctot += sl[j].nbc;
    }
    return 0;
}

int orc_force_sample(int n, const double *mass, const double *pos, int ns, const int64_t *idx,
                     double tol, double eps, int usequad, double *rmin, double *rsize, double fcells,
                     double *phi, double *acc, int64_t *n2btot, int64_t *nbctot, double *t_tree, double *t_walk)
{
    return orc_force_sample_mt(n, mass, pos, ns, idx, tol, eps, usequad, rmin, rsize, fcells, 1, phi, acc,
                               n2btot, nbctot, t_tree, t_walk);
}
