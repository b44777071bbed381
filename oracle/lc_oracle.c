/*
 * lc_oracle.c -- CPU restatement of the learned far-field correction's data
 * path: the exact residual field of an accepted cell, the MLP's input
 * features, and the first-order group correction.
 *
 * TEST INFRASTRUCTURE ONLY.  Called by tests/ (through oracle/oracle.py) as
 * the checker of the device's sample generator (k_samples), feature records
 * and correction pass; never linked into the product library.
 *
 * What a cell's residual field is (FP64 throughout):
 *   exact part      the softened pair sum over the cell's bodies at point t,
 *                   directcode's form (src/nbody/evolve/directcode/grav.c:33-45,
 *                   gravc = 1): a = sum m (x - t) / (|x - t|^2 + eps^2)^(3/2)
 *   multipole part  gravsub's force of the cell seen from t
 *                   (src/nbody/evolve/hackcode/hackcode1/grav.c:42-73):
 *                   softened monopole, plus the quadrupole terms when asked
 *   residual        res = exact - multipole, and its Jacobian
 *                   J[i][k] = d res_i / d t_k.
 * The exact and monopole Jacobians are written out analytically
 * (m (3 d d^T / r^5 - I / r^3) with d = source - t); the quadrupole part's
 * Jacobian is a Richardson-extrapolated central difference of its force.
 *
 * Pinning (tests/test_lc_oracle.py): a one-body cell has zero residual and
 * Jacobian; the exact part equals the oracle's direct sum (orc_direct) at a
 * massless test point; the multipole part equals the oracle's hackforce walk
 * when that walk accepts the cell as a whole; the Jacobian equals central
 * differences of the residual; the first-order correction converges to the
 * exact residual at second order in the offset; the features' direction is
 * a unit vector and their quadrupole is the point-set formula.
 */
#include <math.h>
#include <string.h>

/* gravsub's force of a cell (mass M, c.o.m. com, traceless quadrupole Q
 * [3][3] or NULL) at point t, softening eps (grav.c:42-73). */
static void multipole_force(double M, const double com[3], const double *Q, double eps, const double t[3],
                            double a[3])
{
    double dr[3] = {com[0] - t[0], com[1] - t[1], com[2] - t[2]};
    double drsq = dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2] + eps * eps;
    double drabs = sqrt(drsq);
    double phii = M / drabs;
    double mor3 = phii / drsq;
    for (int k = 0; k < 3; k++) a[k] = dr[k] * mor3;
    if (Q) {
        double dr5inv = 1.0 / (drsq * drsq * drabs);
        double qdr[3];
        for (int i = 0; i < 3; i++) qdr[i] = Q[3 * i] * dr[0] + Q[3 * i + 1] * dr[1] + Q[3 * i + 2] * dr[2];
        double drqdr = dr[0] * qdr[0] + dr[1] * qdr[1] + dr[2] * qdr[2];
        double phiq = -0.5 * dr5inv * drqdr;
        phiq = 5.0 * phiq / drsq;
        for (int k = 0; k < 3; k++) a[k] = a[k] - dr[k] * phiq - qdr[k] * dr5inv;
    }
}

/* Jacobian (J[i][k] = d a_i / d t_k) of the field m d / (|d|^2 + eps^2)^(3/2),
 * d = src - t, accumulated into J. */
static void point_jacobian(double m, const double src[3], double eps, const double t[3], double *J)
{
    double d[3] = {src[0] - t[0], src[1] - t[1], src[2] - t[2]};
    double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2] + eps * eps;
    double ri = 1.0 / sqrt(r2);
    double w = m * ri * ri * ri;
    double w5 = 3.0 * w * ri * ri;
    for (int i = 0; i < 3; i++)
        for (int k = 0; k < 3; k++) J[3 * i + k] += w5 * d[i] * d[k] - (i == k ? w : 0.0);
}

void lc_residual_field(int nb, const double *m, const double *x, double M, const double *com, const double *Q,
                       double eps, const double *t, double *res, double *J)
{
    double ex[3] = {0.0, 0.0, 0.0};
    memset(J, 0, 9 * sizeof(double));
    for (int j = 0; j < nb; j++) {
        const double *p = x + 3 * j;
        double d[3] = {p[0] - t[0], p[1] - t[1], p[2] - t[2]};
        double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2] + eps * eps;
        double mor3 = m[j] / (r2 * sqrt(r2));
        for (int k = 0; k < 3; k++) ex[k] += d[k] * mor3;
        point_jacobian(m[j], p, eps, t, J);
    }
    double a[3];
    multipole_force(M, com, Q, eps, t, a);
    for (int k = 0; k < 3; k++) res[k] = ex[k] - a[k];
    /* minus the monopole's Jacobian (analytic) */
    double Jm[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    point_jacobian(M, com, eps, t, Jm);
    for (int i = 0; i < 9; i++) J[i] -= Jm[i];
    if (!Q) return;
    /* minus the quadrupole terms' Jacobian: central differences at steps h
     * and h/2, Richardson-combined (error O(h^4)) */
    double r = sqrt((com[0] - t[0]) * (com[0] - t[0]) + (com[1] - t[1]) * (com[1] - t[1]) +
                    (com[2] - t[2]) * (com[2] - t[2]) + eps * eps);
    double h = 1e-2 * r;
    for (int k = 0; k < 3; k++) {
        double D[2][3];
        for (int s = 0; s < 2; s++) {
            double hs = s ? 0.5 * h : h;
            double tp[3] = {t[0], t[1], t[2]}, tm[3] = {t[0], t[1], t[2]};
            tp[k] += hs;
            tm[k] -= hs;
            double ap[3], am[3], bp[3], bm[3];
            multipole_force(M, com, Q, eps, tp, ap);
            multipole_force(M, com, Q, eps, tm, am);
            multipole_force(M, com, NULL, eps, tp, bp);   /* the monopole part is taken analytically */
            multipole_force(M, com, NULL, eps, tm, bm);
            for (int i = 0; i < 3; i++) D[s][i] = ((ap[i] - bp[i]) - (am[i] - bm[i])) / (2.0 * hs);
        }
        for (int i = 0; i < 3; i++) J[3 * i + k] -= (4.0 * D[1][i] - D[0][i]) / 3.0;
    }
}

/* Input features of a cell (c.o.m. com, mass M, traceless quadrupole Q[3][3],
 * geometric centre cc, side s) seen from point t:
 *   f[0..2] = d / r with d = com - t, r = |d| (unsoftened)
 *   f[3]    = s / r
 *   f[4..6] = (com - cc) / s
 *   f[7..11] = (Qxx, Qyy, Qxy, Qxz, Qyz) / (M s^2)
 * Returns the residual's unit M / r^2 (the Jacobian's is M / r^3) and sets
 * *rinv = 1 / r; a point at the c.o.m. gives zero features and unit. */
double lc_features(const double *com, double M, const double *Q, const double *cc, double s, const double *t,
                   double *f, double *rinv)
{
    double d[3] = {com[0] - t[0], com[1] - t[1], com[2] - t[2]};
    double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    if (!(r2 > 0.0)) {
        memset(f, 0, 12 * sizeof(double));
        *rinv = 0.0;
        return 0.0;
    }
    double r = sqrt(r2);
    for (int k = 0; k < 3; k++) f[k] = d[k] / r;
    f[3] = s / r;
    for (int k = 0; k < 3; k++) f[4 + k] = (com[k] - cc[k]) / s;
    double qn = 1.0 / (M * s * s);
    f[7] = Q[0] * qn;
    f[8] = Q[4] * qn;
    f[9] = Q[1] * qn;
    f[10] = Q[2] * qn;
    f[11] = Q[5] * qn;
    *rinv = 1.0 / r;
    return M / r2;
}

/* First-order correction of n points x[n][3] of a group with centre c from
 * the group's summed residual R[3] and symmetric Jacobian j[6] (xx xy xz yy
 * yz zz): out[i] = R + J (x_i - c). */
void lc_apply(int n, const double *x, const double *c, const double *R, const double *j, double *out)
{
    for (int i = 0; i < n; i++) {
        double d[3] = {x[3 * i] - c[0], x[3 * i + 1] - c[1], x[3 * i + 2] - c[2]};
        out[3 * i] = R[0] + j[0] * d[0] + j[1] * d[1] + j[2] * d[2];
        out[3 * i + 1] = R[1] + j[1] * d[0] + j[3] * d[1] + j[4] * d[2];
        out[3 * i + 2] = R[2] + j[2] * d[0] + j[4] * d[1] + j[5] * d[2];
    }
}
