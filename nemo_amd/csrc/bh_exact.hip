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

__global__ __launch_bounds__(kTpb) void k_walk_strict(
    const double4 *__restrict__ td, const int *__restrict__ tself, int selfmode, const uint32_t *__restrict__ tperm,
    int64_t nt,
    const int *__restrict__ pre, const int *__restrict__ skip, const uint8_t *__restrict__ level,
    const double *__restrict__ mo,
    const double4 *__restrict__ bd, int nb, const double4 *__restrict__ box, double tol, double eps, int usequad,
    double *phi_out, double *acc_out, unsigned long long *counters)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int nnode = nb + (int)counters[S_NCELL];
    const double rsize = box->w;
    unsigned long long n2b = 0, nbc = 0;
    if (i < nt) {
        const double4 x0 = td[i];
        const int self = selfmode == 1 ? (int)i : (selfmode == 2 ? tself[i] : -1);
        const double tolsq = tol * tol;
        const double dsq0 = rsize * rsize;
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
                double dsq = dsq0 * ldexp(1.0, -2 * (int)level[c]);
                if (tolsq * drsq < dsq) {
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
                                                  p.tol, p.eps, p.usequad, phi, acc, cnt);
    BH_HIP(hipGetLastError());
}

void direct_sum(bh_ctx *ctx, double eps, int64_t nidx, const int64_t *d_idx, double *d_phi, double *d_acc)
{
    int grid = (int)std::max<int64_t>(1, (nidx + kTpb - 1) / kTpb);
    k_direct<<<grid, kTpb, 0, ctx->stream>>>(ptr<double>(ctx->pos), ptr<double>(ctx->mass), ctx->nmass, d_idx, nidx, eps,
                                             d_phi, d_acc);
    BH_HIP(hipGetLastError());
}

}  // namespace bh
