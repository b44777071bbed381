/*
 * le_oracle.c -- CPU restatement of the fast walk's group local expansion
 * (bh_params.local_alpha): the second-order expansion, about a group centre,
 * of the softened monopole fields of the far cells a group accepts.
 *
 * TEST INFRASTRUCTURE ONLY.  Called by tests/ (through oracle/oracle.py) as
 * the checker of the walk's local-expansion kernels; never linked into the
 * product library.
 *
 * A cell (mass m, c.o.m. s) seen from a point t, with d = s - t and
 * rho^2 = |d|^2 + eps^2, gives gravsub's softened monopole
 * (src/nbody/evolve/hackcode/hackcode1/grav.c:42-73, monopole part):
 *     phi(t) = -m / rho,    a(t) = m d / rho^3.
 * With g(t) = 1/rho, a = m grad g and phi = -m g.  About a centre c, for a
 * member t = c + delta:
 *     g(t)   = g + g_i delta_i + 1/2 g_ij delta_i delta_j
 *              + 1/6 g_ijk delta_i delta_j delta_k + O(delta^4)
 *     a_i(t) = m (g_i + g_ij delta_j + 1/2 g_ijk delta_j delta_k) + O(delta^3)
 * where, at t = c (d = s - c):
 *     g_i   = d_i / rho^3
 *     g_ij  = 3 d_i d_j / rho^5 - delta_ij / rho^3
 *     g_ijk = 15 d_i d_j d_k / rho^7
 *             - 3 (delta_ij d_k + delta_ik d_j + delta_jk d_i) / rho^5.
 * The coefficients are summed over the cells (times m) and stored scaled by
 * a length ell (any ell > 0 gives the same values of a and phi):
 *     G[0] = sum m g                      potential
 *     G[1..3]   = sum m g_i               x y z
 *     G[4..9]   = ell sum m g_ij          xx xy xz yy yz zz
 *     G[10..19] = ell^2 sum m g_ijk       xxx xxy xxz xyy xyz xzz yyy yyz yzz zzz
 * and evaluated at u = delta / ell.
 *
 * Pinning (tests/test_le_oracle.py): at the centre the expansion equals the
 * exact field and potential; its error falls at third order in delta for the
 * field and fourth for the potential; G[4..9] and G[10..19] equal central
 * differences of the exact field and of G[4..9]; with eps = 0 both tensors
 * are traceless (1/r is harmonic); coefficients are additive over cells.
 */
#include <math.h>
#include <string.h>

/* index of the symmetric rank-3 component (i <= j <= k) in G[10..19] */
static int sym3(int i, int j, int k)
{
    int a[3] = {i, j, k};
    for (int p = 0; p < 2; p++)
        for (int q = 0; q < 2 - p; q++)
            if (a[q] > a[q + 1]) {
                int t = a[q];
                a[q] = a[q + 1];
                a[q + 1] = t;
            }
    static const int tab[3][3][3] = {
        {{0, 1, 2}, {-1, 3, 4}, {-1, -1, 5}},
        {{-1, -1, -1}, {-1, 6, 7}, {-1, -1, 8}},
        {{-1, -1, -1}, {-1, -1, -1}, {-1, -1, 9}},
    };
    return 10 + tab[a[0]][a[1]][a[2]];
}

static int sym2(int i, int j)
{
    static const int tab[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
    return 4 + tab[i][j];
}

/* Summed, scaled coefficients G[20] of k cells (mass m[k], c.o.m. s[k][3])
 * about centre c with softening eps and length ell. */
void le_coeffs(int k, const double *m, const double *s, const double *c, double eps, double ell, double *G)
{
    memset(G, 0, 20 * sizeof(double));
    for (int n = 0; n < k; n++) {
        double d[3] = {s[3 * n] - c[0], s[3 * n + 1] - c[1], s[3 * n + 2] - c[2]};
        double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2] + eps * eps;
        double r = sqrt(r2);
        double r3 = r2 * r, r5 = r3 * r2, r7 = r5 * r2;
        G[0] += m[n] / r;
        for (int i = 0; i < 3; i++) G[1 + i] += m[n] * d[i] / r3;
        for (int i = 0; i < 3; i++)
            for (int j = i; j < 3; j++)
                G[sym2(i, j)] += ell * m[n] * (3.0 * d[i] * d[j] / r5 - (i == j ? 1.0 / r3 : 0.0));
        for (int i = 0; i < 3; i++)
            for (int j = i; j < 3; j++)
                for (int l = j; l < 3; l++) {
                    double t = 15.0 * d[i] * d[j] * d[l] / r7;
                    double dl = ((i == j) ? d[l] : 0.0) + ((i == l) ? d[j] : 0.0) + ((j == l) ? d[i] : 0.0);
                    G[sym3(i, j, l)] += ell * ell * m[n] * (t - 3.0 * dl / r5);
                }
    }
}

/* Potential phi[n] (negative, NEMO's sign) and field acc[n][3] of the
 * expansion G (length ell, centre c) at the points x[n][3]. */
void le_eval(const double *G, double ell, const double *c, int n, const double *x, double *phi, double *acc)
{
    for (int p = 0; p < n; p++) {
        double u[3];
        for (int i = 0; i < 3; i++) u[i] = (x[3 * p + i] - c[i]) / ell;
        double g2u[3] = {0, 0, 0}, g3uu[3] = {0, 0, 0};
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                g2u[i] += G[sym2(i, j)] * u[j];
                for (int l = 0; l < 3; l++) g3uu[i] += G[sym3(i, j, l)] * u[j] * u[l];
            }
        double g = G[0];
        for (int i = 0; i < 3; i++) {
            acc[3 * p + i] = G[1 + i] + g2u[i] + 0.5 * g3uu[i];
            g += ell * u[i] * (G[1 + i] + 0.5 * g2u[i] + g3uu[i] / 6.0);
        }
        phi[p] = -g;
    }
}
