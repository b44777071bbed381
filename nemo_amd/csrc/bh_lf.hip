// hackcode1's leapfrog on the device: the kick correction, the FP64
// diagnostics and the drift of stepsystem/diagnostics
// (src/nbody/evolve/hackcode/hackcode1/code.c:159-191, code_io.c:351-388),
// with the force pass in between done by the tree build and walk.
//
// Compiled with -ffp-contract=off: the per-body updates keep the reference's
// rounding sequence exactly (dvel = (a - a_old) * dt/2; v = v + dvel; v1 =
// v + a dt/2; x = x + v1 dt; v = v1 + a dt/2), so with the strict walk the
// trajectory is bit-identical to the reference.  The diagnostics are FP64
// sums reduced in a fixed order (deterministic, not the reference's
// sequential order).
#include "bh_ctx.h"

#include <algorithm>

namespace bh {
namespace {

constexpr int kTpb = 256;
constexpr int kDiagBlocks = 512;

// Second-order velocity correction after the force pass (nstep > 0), then
// remember the new accelerations.
__global__ void k_lf_kick(int64_t n, double *vel, const double *acc, double *acc_old, const int *nstep, double dthf)
{
    const bool corr = *nstep > 0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < 3 * n; i += (int64_t)gridDim.x * blockDim.x) {
        double a = acc[i];
        if (corr) {
            double dacc = a - acc_old[i];
            double dvel = dacc * dthf;
            vel[i] = vel[i] + dvel;
        }
        acc_old[i] = a;
    }
}

__global__ void k_lf_drift(int64_t n, double *pos, double *vel, const double *acc, double dt, double dthf, int *nstep)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < 3 * n; i += (int64_t)gridDim.x * blockDim.x) {
        double dvel = acc[i] * dthf;
        double vel1 = vel[i] + dvel;
        double dpos = vel1 * dt;
        pos[i] = pos[i] + dpos;
        vel[i] = vel1 + dvel;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) *nstep += 1;
}

// Per-body terms of diagnostics(): slots 0 M, 1 T, 2 U, 3..11 KE tensor,
// 12..20 PE tensor, 21..29 m x v^T (antisymmetrised at the end), 30..35
// m x and m v.
__device__ __forceinline__ void body_terms(const double *x, const double *v, const double *a, double m, double phi,
                                           double *t)
{
    double velsq = v[0] * v[0];
    velsq += v[1] * v[1];
    velsq += v[2] * v[2];
    t[0] = m;
    t[1] = 0.5 * m * velsq;
    t[2] = 0.5 * m * phi;
    double hv[3] = {v[0] * (0.5 * m), v[1] * (0.5 * m), v[2] * (0.5 * m)};
    double mx[3] = {x[0] * m, x[1] * m, x[2] * m};
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            t[3 + 3 * i + j] = hv[i] * v[j];
            t[12 + 3 * i + j] = mx[i] * a[j];
            t[21 + 3 * i + j] = mx[i] * v[j];
        }
    for (int k = 0; k < 3; k++) {
        t[30 + k] = mx[k];
        t[33 + k] = v[k] * m;
    }
}

// Block b sums bodies [b*chunk, (b+1)*chunk) in index order per thread, then
// the 256 thread sums in a fixed tree order.
__global__ __launch_bounds__(kTpb) void k_diag_partial(int64_t n, int64_t chunk, const double *mass, const double *pos,
                                                       const double *vel, const double *acc, const double *phi,
                                                       double *partial)
{
    __shared__ double red[kTpb];
    double acc36[36];
    for (int k = 0; k < 36; k++) acc36[k] = 0.0;
    const int64_t lo = blockIdx.x * chunk, hi = min(n, lo + chunk);
    for (int64_t i = lo + threadIdx.x; i < hi; i += kTpb) {
        double t[36];
        body_terms(pos + 3 * i, vel + 3 * i, acc + 3 * i, mass[i], phi[i], t);
        for (int k = 0; k < 36; k++) acc36[k] += t[k];
    }
    for (int k = 0; k < 36; k++) {
        red[threadIdx.x] = acc36[k];
        __syncthreads();
        for (int s = kTpb / 2; s > 0; s >>= 1) {
            if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
            __syncthreads();
        }
        if (threadIdx.x == 0) partial[blockIdx.x * 36 + k] = red[0];
        __syncthreads();
    }
}

// Final sums in block order and the closing steps of diagnostics(): Energy =
// (T+U, T, U), AM tensor antisymmetrised, c.o.m. phase divided by M.  Writes
// the 36 values to out[*rec * 36] and bumps *rec.
__global__ void k_diag_final(const double *partial, int nblk, double *out, int *rec)
{
    __shared__ double s[36];
    int k = threadIdx.x;
    if (k < 36) {
        double a = 0.0;
        for (int b = 0; b < nblk; b++) a += partial[b * 36 + k];
        s[k] = a;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double *o = out + 36 * (int64_t)(*rec);
        const double mtot = s[0];
        o[0] = s[1] + s[2];
        o[1] = s[1];
        o[2] = s[2];
        for (int j = 0; j < 9; j++) {
            o[3 + j] = s[3 + j];
            o[12 + j] = s[12 + j];
        }
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) o[21 + 3 * i + j] = s[21 + 3 * i + j] - s[21 + 3 * j + i];
        for (int c = 0; c < 6; c++) o[30 + c] = s[30 + c] / mtot;
    }
}

// Batches: the 36 diagnostics of every system (one block per system, the
// same per-thread order and fixed tree reduction as above), written to
// out[(*rec * nsys + s) * 36].
__global__ __launch_bounds__(kTpb) void k_diag_sys(const int64_t *__restrict__ off, int nsys, const double *mass,
                                                   const double *pos, const double *vel, const double *acc,
                                                   const double *phi, double *out, const int *rec)
{
    __shared__ double red[kTpb];
    __shared__ double s[36];
    const int sy = blockIdx.x;
    double acc36[36];
    for (int k = 0; k < 36; k++) acc36[k] = 0.0;
    for (int64_t i = off[sy] + threadIdx.x; i < off[sy + 1]; i += kTpb) {
        double t[36];
        body_terms(pos + 3 * i, vel + 3 * i, acc + 3 * i, mass[i], phi[i], t);
        for (int k = 0; k < 36; k++) acc36[k] += t[k];
    }
    for (int k = 0; k < 36; k++) {
        red[threadIdx.x] = acc36[k];
        __syncthreads();
        for (int h = kTpb / 2; h > 0; h >>= 1) {
            if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
            __syncthreads();
        }
        if (threadIdx.x == 0) s[k] = red[0];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        double *o = out + 36 * ((int64_t)(*rec) * nsys + sy);
        const double mtot = s[0];
        o[0] = s[1] + s[2];
        o[1] = s[1];
        o[2] = s[2];
        for (int j = 0; j < 9; j++) {
            o[3 + j] = s[3 + j];
            o[12 + j] = s[12 + j];
        }
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) o[21 + 3 * i + j] = s[21 + 3 * i + j] - s[21 + 3 * j + i];
        for (int c = 0; c < 6; c++) o[30 + c] = s[30 + c] / mtot;
    }
}

// Per-step interaction counts (the walk's counters of this step).
__global__ void k_lf_record(unsigned long long *stat, const int *rec, long long *n2b, long long *nbc)
{
    int r = *rec;
    n2b[r] = (long long)stat[S_N2B];
    nbc[r] = (long long)stat[S_NBC];
    stat[S_STICKY] |= stat[S_FLAGS];
}

__global__ void k_lf_next(int *rec) { *rec += 1; }
__global__ void k_lf_begin(int *rec, unsigned long long *stat)
{
    *rec = 0;
    stat[S_STICKY] = 0;
}
__global__ void k_set(int *p, int v) { *p = v; }

inline int grid_for(int64_t n) { return (int)std::min<int64_t>(4096, std::max<int64_t>(1, (3 * n + kTpb - 1) / kTpb)); }

}  // namespace

void lf_reset(bh_ctx *ctx)
{
    k_set<<<1, 1, 0, ctx->stream>>>(ensure<int>(ctx->lfstep, 2), 0);
    BH_HIP(hipGetLastError());
}

// Restart the per-call record index and clear the flags or-ed over the
// steps of the call (the step counter is kept).
void lf_reset_record(bh_ctx *ctx)
{
    k_lf_begin<<<1, 1, 0, ctx->stream>>>(ptr<int>(ctx->lfstep) + 1,
                                          ensure<unsigned long long>(ctx->counters, S_NSLOTS));
    BH_HIP(hipGetLastError());
}

// Force pass of one step: tree + walk (done by the caller) then the kick
// correction and the diagnostics of the output-time state.
void lf_after_force(bh_ctx *ctx, double *diag_out, long long *n2b_out, long long *nbc_out)
{
    hipStream_t st = ctx->stream;
    const int64_t n = ctx->nmass;
    int *nstep = ptr<int>(ctx->lfstep), *rec = nstep + 1;
    k_lf_kick<<<grid_for(n), kTpb, 0, st>>>(n, ptr<double>(ctx->vel), ptr<double>(ctx->acc), ptr<double>(ctx->acc_old),
                                            nstep, 0.5 * ctx->lf_dt);
    if (ctx->nsys) {
        k_diag_sys<<<ctx->nsys, kTpb, 0, st>>>(ptr<int64_t>(ctx->boff), ctx->nsys, ptr<double>(ctx->mass),
                                               ptr<double>(ctx->pos), ptr<double>(ctx->vel), ptr<double>(ctx->acc),
                                               ptr<double>(ctx->phi), diag_out, rec);
    } else {
        const int nblk = (int)std::min<int64_t>(kDiagBlocks, std::max<int64_t>(1, (n + kTpb - 1) / kTpb));
        const int64_t chunk = (n + nblk - 1) / nblk;
        double *partial = ensure<double>(ctx->diag_partial, 36 * (size_t)kDiagBlocks);
        k_diag_partial<<<nblk, kTpb, 0, st>>>(n, chunk, ptr<double>(ctx->mass), ptr<double>(ctx->pos),
                                              ptr<double>(ctx->vel), ptr<double>(ctx->acc), ptr<double>(ctx->phi),
                                              partial);
        k_diag_final<<<1, 64, 0, st>>>(partial, nblk, diag_out, rec);
    }
    k_lf_record<<<1, 1, 0, st>>>(ptr<unsigned long long>(ctx->counters), rec, n2b_out, nbc_out);
    BH_HIP(hipGetLastError());
}

void lf_drift(bh_ctx *ctx)
{
    const int64_t n = ctx->nmass;
    int *nstep = ptr<int>(ctx->lfstep);
    k_lf_drift<<<grid_for(n), kTpb, 0, ctx->stream>>>(n, ptr<double>(ctx->pos), ptr<double>(ctx->vel),
                                                      ptr<double>(ctx->acc), ctx->lf_dt, 0.5 * ctx->lf_dt, nstep);
    k_lf_next<<<1, 1, 0, ctx->stream>>>(nstep + 1);
    BH_HIP(hipGetLastError());
}

}  // namespace bh
