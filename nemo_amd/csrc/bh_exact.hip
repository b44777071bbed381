// FP64 walks that reproduce the reference arithmetic: the strict
// per-particle Barnes-Hut walk (hackforce's decisions, counters and summation
// order) and the direct sum.  This translation unit is compiled with
// -ffp-contract=off so that no multiply-add is fused, matching the
// reference's x86-64 FP64 rounding sequence.
//
// Strict walk = walksub/subdivp/gravsub of src/nbody/evolve/hackcode/
// hackcode1/grav.c:42-131 run as a stackless DFS over the pre-order node
// array: a cell at pre-order position p opens to p+1 (its first child) or is
// accepted and continues at skip[c]; a body continues at p+1.  Children are
// visited in octant order 0..7, the same order as the recursive walk.
// Direct sum = src/nbody/evolve/directcode/grav.c:23-65 (eps >= 0 branch).
//
// treecode1 criteria (bh_params.criterion != BH_CRIT_HACK): k_rcrit gives
// every cell treecode1's critical radius (setrcrit, src/nbody/evolve/hackcode/
// treecode1/treeload.c:232-252) from its c.o.m. and its geometric centre,
// which is built like loadbody's (treeload.c:154-156: parent centre +- parent
// size / 4, level by level from the root centre).  The strict walk then
// replaces subdivp by treecode1's accept (treegrav.c:137-156) against the
// sink cell of the cell's level: the cube of that side holding the body,
// whose centre walksub builds by the same halving with body-position
// comparisons (treegrav.c:182-186).  A cell of level L is tested only after
// its parent was opened, as on treecode1's level-by-level active list, so
// the per-body DFS accepts the same cells.
#include "bh_ctx.h"

#include <algorithm>

namespace bh {
namespace {

constexpr int kTpb = 256;

struct Acc64 {
    double phi = 0.0, ax = 0.0, ay = 0.0, az = 0.0;
};

__device__ __forceinline__ void gravsub64(Acc64 &a, double dx, double dy, double dz, double drsq, double m,
                                          double eps, const double *q)
{
    drsq += eps * eps;
    double drabs = sqrt(drsq);
    double phii = m / drabs;
    a.phi -= phii;
    double mor3 = phii / drsq;
    a.ax = a.ax + dx * mor3;
    a.ay = a.ay + dy * mor3;
    a.az = a.az + dz * mor3;
    if (q) {
        // q: xx xy xz yy yz zz (exactly symmetric, so MULMV's row order is kept)
        double dr5inv = 1.0 / (drsq * drsq * drabs);
        double qx = 0.0, qy = 0.0, qz = 0.0;
        qx += q[0] * dx; qx += q[1] * dy; qx += q[2] * dz;
        qy += q[1] * dx; qy += q[3] * dy; qy += q[4] * dz;
        qz += q[2] * dx; qz += q[4] * dy; qz += q[5] * dz;
        double drqdr = dx * qx;
        drqdr += dy * qy;
        drqdr += dz * qz;
        double phiq = -0.5 * dr5inv * drqdr;
        a.phi = a.phi + phiq;
        phiq = 5.0 * phiq / drsq;
        a.ax = a.ax - dx * phiq;
        a.ay = a.ay - dy * phiq;
        a.az = a.az - dz * phiq;
        a.ax = a.ax - qx * dr5inv;
        a.ay = a.ay - qy * dr5inv;
        a.az = a.az - qz * dr5inv;
    }
}

// treecode1's Rcrit2 per cell (criterion 1 default, 2 bh86, 3 sw94), also
// into cB[c].w for the fast walk's group test.
__global__ __launch_bounds__(kTpb) void k_rcrit(const uint64_t *__restrict__ keys, const int *__restrict__ cfirst,
                                                const uint8_t *__restrict__ level, const double *__restrict__ mo,
                                                const unsigned long long *__restrict__ ncellp,
                                                const double4 *__restrict__ box, double theta, int crit, double *rc2,
                                                float4 *cB)
{
    const int ncell = (int)*ncellp;
    const double4 b = *box;
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < ncell; c += gridDim.x * blockDim.x) {
        const int L = level[c];
        const uint64_t key = keys[cfirst[c]];
        double m0 = b.x + 0.5 * b.w, m1 = b.y + 0.5 * b.w, m2 = b.z + 0.5 * b.w, s = b.w;
        for (int j = 1; j <= L; j++) {
            const int d = (int)((key >> (63 - 3 * j)) & 7);
            m0 = m0 + ((d & 4) ? s : -s) / 4;
            m1 = m1 + ((d & 2) ? s : -s) / 4;
            m2 = m2 + ((d & 1) ? s : -s) / 4;
            s = s / 2;
        }
        const double *mc = mo + 10 * (int64_t)c;
        const double cm[3] = {mc[1], mc[2], mc[3]}, mid[3] = {m0, m1, m2};
        double r;
        if (theta == 0.0) {
            r = (2 * b.w) * (2 * b.w);
        } else if (crit == BH_CRIT_SW94) {
            double bmax2 = 0.0;
            for (int k = 0; k < 3; k++) {
                const double d = cm[k] - mid[k] + s / 2;
                const double e = d > s - d ? d : s - d;
                bmax2 += e * e;
            }
            r = bmax2 / (theta * theta);
        } else if (crit == BH_CRIT_BH86) {
            const double e = s / theta;
            r = e * e;
        } else {
            double t = 0.0;
            for (int k = 0; k < 3; k++) t += (cm[k] - mid[k]) * (cm[k] - mid[k]);
            const double e = s / theta + sqrt(t);
            r = e * e;
        }
        rc2[c] = r;
        reinterpret_cast<float *>(cB + c)[3] = (float)r;
    }
}

// Centre of the level-L sink cube of a body whose walksub comparisons gave
// the bits bx, by, bz (bit kMaxLevel - j for level j); its side goes to *side.
__device__ __forceinline__ void sink_centre(const double4 b, int bx, int by, int bz, int L, double mid[3],
                                            double *side)
{
    double m0 = b.x + 0.5 * b.w, m1 = b.y + 0.5 * b.w, m2 = b.z + 0.5 * b.w, s = b.w;
    for (int j = 1; j <= L; j++) {
        const int sh = kMaxLevel - j;
        const double poff = s / 4;
        m0 = m0 + (((bx >> sh) & 1) ? poff : -poff);
        m1 = m1 + (((by >> sh) & 1) ? poff : -poff);
        m2 = m2 + (((bz >> sh) & 1) ? poff : -poff);
        s = s / 2;
    }
    mid[0] = m0;
    mid[1] = m1;
    mid[2] = m2;
    *side = s;
}

__global__ __launch_bounds__(kTpb) void k_walk_strict(
    const double4 *__restrict__ td, const int *__restrict__ tself, int selfmode, const uint32_t *__restrict__ tperm,
    int64_t nt,
    const int *__restrict__ pre, const int *__restrict__ skip, const uint8_t *__restrict__ level,
    const double *__restrict__ mo,
    const double4 *__restrict__ bd, int nb, const double4 *__restrict__ box, double tol, double eps, int usequad,
    int crit, const double *__restrict__ rc2, double *phi_out, double *acc_out, unsigned long long *counters)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int nnode = nb + (int)counters[S_NCELL];
    const double4 bb = *box;
    const double rsize = bb.w;
    unsigned long long n2b = 0, nbc = 0;
    if (i < nt) {
        const double4 x0 = td[i];
        const int self = selfmode == 1 ? (int)i : (selfmode == 2 ? tself[i] : -1);
        const double tolsq = tol * tol;
        const double dsq0 = rsize * rsize;
        // walksub's descent of the body's sink: upper half when centre <= x
        int bx = 0, by = 0, bz = 0;
        if (crit) {
            double m0 = bb.x + 0.5 * bb.w, m1 = bb.y + 0.5 * bb.w, m2 = bb.z + 0.5 * bb.w, s = bb.w;
            for (int j = 1; j <= kMaxLevel; j++) {
                const double poff = s / 4;
                const int ux = !(x0.x < m0), uy = !(x0.y < m1), uz = !(x0.z < m2);
                bx = (bx << 1) | ux;
                by = (by << 1) | uy;
                bz = (bz << 1) | uz;
                m0 = m0 + (ux ? poff : -poff);
                m1 = m1 + (uy ? poff : -poff);
                m2 = m2 + (uz ? poff : -poff);
                s = s / 2;
            }
        }
        Acc64 a;
        int p = 0;
        while (p < nnode) {
            int r = pre[p];
            if (r < 0) {
                int c = ~r;
                const double *mc = mo + 10 * (int64_t)c;
                double dx = mc[1] - x0.x;
                double dy = mc[2] - x0.y;
                double dz = mc[3] - x0.z;
                double drsq = dx * dx;
                drsq += dy * dy;
                drsq += dz * dz;
                bool open;
                if (crit) {
                    double pmid[3], psize;
                    sink_centre(bb, bx, by, bz, (int)level[c], pmid, &psize);
                    double dmax = psize, dsq = 0.0;
                    for (int k = 0; k < 3; k++) {
                        double dk = mc[1 + k] - pmid[k];
                        if (dk < 0) dk = -dk;
                        if (dk > dmax) dmax = dk;
                        dk -= 0.5 * psize;
                        if (dk > 0) dsq += dk * dk;
                    }
                    open = !(dsq > rc2[c] && dmax > 1.5 * psize);
                } else {
                    double dsq = dsq0 * ldexp(1.0, -2 * (int)level[c]);
                    open = tolsq * drsq < dsq;
                }
                if (open) {
                    p++;
                    continue;
                }
                gravsub64(a, dx, dy, dz, drsq, mc[0], eps, usequad ? mc + 4 : nullptr);
                nbc++;
                p = skip[c];
            } else {
                if (r != self) {
                    double4 b = bd[r];
                    double dx = b.x - x0.x, dy = b.y - x0.y, dz = b.z - x0.z;
                    double drsq = dx * dx;
                    drsq += dy * dy;
                    drsq += dz * dz;
                    gravsub64(a, dx, dy, dz, drsq, b.w, eps, nullptr);
                    n2b++;
                }
                p++;
            }
        }
        int64_t o = tperm[i];
        phi_out[o] = a.phi;
        acc_out[3 * o] = a.ax;
        acc_out[3 * o + 1] = a.ay;
        acc_out[3 * o + 2] = a.az;
    }
    for (int d = 32; d >= 1; d >>= 1) {
        n2b += __shfl_xor(n2b, d, 64);
        nbc += __shfl_xor(nbc, d, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        if (n2b) atomicAdd(&counters[0], n2b);
        if (nbc) atomicAdd(&counters[1], nbc);
    }
}

__global__ void k_zero_out(double *phi, int64_t n, double *acc)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) phi[i] = 0.0;
    if (i < 3 * n) acc[i] = 0.0;
}

// Direct sum in input order over all n mass-set bodies (self skipped by
// index), for the targets idx[0..nt) of the mass set.
constexpr int kDirTile = 256;
__global__ __launch_bounds__(kTpb) void k_direct(const double *__restrict__ pos, const double *__restrict__ mass,
                                                 int64_t n, const int64_t *__restrict__ idx, int64_t nt, double eps,
                                                 double *phi_out, double *acc_out)
{
    __shared__ double4 tile[kDirTile];
    int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const bool act = t < nt;
    int64_t i = act ? idx[t] : 0;
    double x = pos[3 * i], y = pos[3 * i + 1], z = pos[3 * i + 2];
    Acc64 a;
    for (int64_t base = 0; base < n; base += kDirTile) {
        __syncthreads();
        int64_t j = base + threadIdx.x;
        if (j < n) tile[threadIdx.x] = make_double4(pos[3 * j], pos[3 * j + 1], pos[3 * j + 2], mass[j]);
        __syncthreads();
        int lim = (int)std::min<int64_t>(kDirTile, n - base);
        for (int k = 0; k < lim; k++) {
            if (base + k == i) continue;
            double4 b = tile[k];
            double dx = b.x - x, dy = b.y - y, dz = b.z - z;
            double drsq = dx * dx;
            drsq += dy * dy;
            drsq += dz * dz;
            gravsub64(a, dx, dy, dz, drsq, b.w, eps, nullptr);
        }
    }
    if (act) {
        phi_out[t] = a.phi;
        acc_out[3 * t] = a.ax;
        acc_out[3 * t + 1] = a.ay;
        acc_out[3 * t + 2] = a.az;
    }
}

}  // namespace

void walk_strict(bh_ctx *ctx, const bh_params &p)
{
    const Tree &T = ctx->tree;
    const int64_t nt = ctx->ntest;
    double *phi = ptr<double>(ctx->phi), *acc = ptr<double>(ctx->acc);
    unsigned long long *cnt = ptr<unsigned long long>(ctx->counters);
    if (T.nb == 0) {
        k_zero_out<<<std::max<int64_t>(1, (3 * nt + kTpb - 1) / kTpb), kTpb, 0, ctx->stream>>>(phi, nt, acc);
        return;
    }
    int grid = (int)std::max<int64_t>(1, (nt + kTpb - 1) / kTpb);
    k_walk_strict<<<grid, kTpb, 0, ctx->stream>>>(T.td, T.tself, T.selfmode, T.tperm, nt, T.pre, T.skip, T.level, T.mo, T.bd, T.nb, T.dbox,
                                                  p.tol, p.eps, p.usequad, p.criterion, T.rc2, phi, acc, cnt);
    BH_HIP(hipGetLastError());
}

double *cell_rcrit(bh_ctx *ctx, const bh_params &p, const uint64_t *keys, const int *cfirst, const uint8_t *level,
                   const double *mo, const unsigned long long *ncellp, int cap, const double4 *box, float4 *cB)
{
    double *rc2 = ensure<double>(ctx->crc2, cap);
    k_rcrit<<<ctx->ncu * 4, kTpb, 0, ctx->stream>>>(keys, cfirst, level, mo, ncellp, box, p.tol, p.criterion, rc2, cB);
    BH_HIP(hipGetLastError());
    return rc2;
}

void direct_sum(bh_ctx *ctx, double eps, int64_t nidx, const int64_t *d_idx, double *d_phi, double *d_acc)
{
    int grid = (int)std::max<int64_t>(1, (nidx + kTpb - 1) / kTpb);
    k_direct<<<grid, kTpb, 0, ctx->stream>>>(ptr<double>(ctx->pos), ptr<double>(ctx->mass), ctx->nmass, d_idx, nidx, eps,
                                             d_phi, d_acc);
    BH_HIP(hipGetLastError());
}

}  // namespace bh
