// Group-wise Barnes-Hut walk with FP32 split-precision interactions.
//
// One wavefront owns one group of 64*NPL test particles that are contiguous
// in Morton order.  The wave walks the tree breadth first: it takes 64
// frontier nodes at a time (one per lane), applies hackforce's opening rule
// (grav.c:122-131: open iff tol^2 |com - x|^2 < size^2) conservatively for
// the whole group by using the distance from the cell's c.o.m. to the
// group's bounding box, and sorts every node into one of three places:
//   opened cell   -> its children are appended to the next frontier,
//   accepted cell -> the P2C list (monopole + optional quadrupole),
//   body          -> the P2P list.
// Both lists live in LDS; whenever one holds 64 entries the wave consumes it
// as a 64 x (64*NPL) tile (every lane reads the same LDS entry, a broadcast)
// and continues.  The frontier lives in a per-wave global scratch area.
//
// Interaction arithmetic follows gravsub (grav.c:42-73): r^2 = |d|^2 + eps^2,
// phi -= m/r, a += m d/r^3, and for quadrupole cells phi -= 1/2 d.Q.d / r^5,
// a -= 5 (-1/2 d.Q.d / r^5) d / r^2 + Q d / r^5, softened r throughout.
// Positions are split in FP32 hi+lo parts; every source is converted to
// coordinates relative to the group centre when it enters a list, so the
// inner loops subtract small numbers.  Each lane sums a tile in plain FP32
// and folds the tile sum into a (hi, lo) pair with TwoSum.
//
// Local expansions (local_alpha > 0, monopole cells): an accepted cell whose
// c.o.m. lies farther than R/alpha from the group centre (R: the group's
// radius about the centre) is not put on the P2C list.  Its softened
// monopole potential, expanded about the centre to third order (field to
// second order), is summed into 20 coefficients per group: in units of the
// group radius l, with n = d/rho, F = m/rho^2 and s = l/rho (< alpha),
//   G0  += m/rho                       (potential)
//   G1  += F n_i                       (field at the centre)
//   G2  += F s (3 n_i n_j - delta_ij)  (its gradient times l)
//   G3  += F s^2 (15 n_i n_j n_k - 3 (delta_ij n_k + delta_ik n_j + delta_jk n_i))
// and each member at u = (x - centre)/l receives
//   a   = G1 + G2 u + 1/2 G3[u, u]
//   phi = -(G0 + l u.(G1 + 1/2 G2 u + 1/6 G3[u, u])).
// Far cells wait in an LDS list; every 64 of them the wave computes one
// cell's terms per lane and sums them across the wave into the group's row
// of 20 coefficients (LDS).
//
// A body's own contribution is removed exactly: with eps > 0 the self pair
// adds m*rsq(eps^2) to the potential sum and nothing to the acceleration, so
// that term is subtracted at the end; with eps == 0 the kernel masks r = 0.
#include "bh_ctx.h"

#include <algorithm>

namespace bh {
namespace {

constexpr int kWaves = 4;          // waves per block
constexpr int kTpb = 64 * kWaves;
constexpr int kListCap = 128;      // entries per list (a tile is 64)

__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }

__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ float wave_min(float v)
{
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v = fminf(v, __shfl_xor(v, d, 64));
    return v;
}
__device__ __forceinline__ float wave_max(float v)
{
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v = fmaxf(v, __shfl_xor(v, d, 64));
    return v;
}
__device__ __forceinline__ int wave_min_i(int v)
{
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v = min(v, __shfl_xor(v, d, 64));
    return v;
}
__device__ __forceinline__ int wave_max_i(int v)
{
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v = max(v, __shfl_xor(v, d, 64));
    return v;
}
__device__ __forceinline__ int wave_excl_scan(int v, int *total)
{
    const int lane = lane_id();
    int x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        int t = __shfl_up(x, d, 64);
        if (lane >= d) x += t;
    }
    *total = __shfl(x, 63, 64);
    return x - v;
}
__device__ __forceinline__ int mbcnt(unsigned long long m)
{
    return __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0));
}

// TwoSum fold of a tile sum into a compensated accumulator.
__device__ __forceinline__ void fold(float &hi, float &lo, float v)
{
    float s = hi + v;
    float bb = s - hi;
    float err = (hi - (s - bb)) + (v - bb);
    hi = s;
    lo += err;
}

__device__ __forceinline__ float rsq(float x) { return __builtin_amdgcn_rsqf(x); }

template <int NPL>
struct Lane {
    float px[NPL], py[NPL], pz[NPL];
    float axh[NPL], axl[NPL], ayh[NPL], ayl[NPL], azh[NPL], azl[NPL], ph[NPL], pl[NPL];
};

// NQ <= NPL: the particle slots in use (a group of at most 64 points uses
// one slot per lane).
template <int NPL, int NQ, bool SOFT0>
__device__ __forceinline__ void eval_p2p(Lane<NPL> &L, const float4 *src, int cnt, float eps2)
{
    float sx[NPL], sy[NPL], sz[NPL], sp[NPL];
#pragma unroll
    for (int q = 0; q < NQ; q++) sx[q] = sy[q] = sz[q] = sp[q] = 0.f;
#pragma unroll 4
    for (int j = 0; j < cnt; j++) {
        const float4 s = src[j];
#pragma unroll
        for (int q = 0; q < NQ; q++) {
            float dx = s.x - L.px[q], dy = s.y - L.py[q], dz = s.z - L.pz[q];
            float r2 = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, eps2)));
            float ri = rsq(r2);
            if (SOFT0) ri = r2 > 0.f ? ri : 0.f;
            float mr = s.w * ri;
            sp[q] += mr;
            float mr3 = mr * ri * ri;
            sx[q] = fmaf(dx, mr3, sx[q]);
            sy[q] = fmaf(dy, mr3, sy[q]);
            sz[q] = fmaf(dz, mr3, sz[q]);
        }
    }
#pragma unroll
    for (int q = 0; q < NQ; q++) {
        fold(L.axh[q], L.axl[q], sx[q]);
        fold(L.ayh[q], L.ayl[q], sy[q]);
        fold(L.azh[q], L.azl[q], sz[q]);
        fold(L.ph[q], L.pl[q], sp[q]);
    }
}

// P2C entry: e[0] = (dx0, dy0, dz0, m) relative to the group centre,
// e[1] = (qxx, qxy, qxz, qyy), e[2] = (qyz, qzz, -, -).
template <int NPL, int NQ, bool QUAD, bool SOFT0>
__device__ __forceinline__ void eval_p2c(Lane<NPL> &L, const float4 *src, int cnt, float eps2)
{
    constexpr int S = QUAD ? 3 : 1;
    float sx[NPL], sy[NPL], sz[NPL], sp[NPL];
#pragma unroll
    for (int q = 0; q < NQ; q++) sx[q] = sy[q] = sz[q] = sp[q] = 0.f;
#pragma unroll 2
    for (int j = 0; j < cnt; j++) {
        const float4 s = src[S * j];
        float4 q0, q1;
        if (QUAD) {
            q0 = src[S * j + 1];
            q1 = src[S * j + 2];
        }
#pragma unroll
        for (int q = 0; q < NQ; q++) {
            float dx = s.x - L.px[q], dy = s.y - L.py[q], dz = s.z - L.pz[q];
            float r2 = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, eps2)));
            float ri = rsq(r2);
            if (SOFT0) ri = r2 > 0.f ? ri : 0.f;
            float ri2 = ri * ri;
            float mr = s.w * ri;
            if (QUAD) {
                float qx = fmaf(q0.x, dx, fmaf(q0.y, dy, q0.z * dz));
                float qy = fmaf(q0.y, dx, fmaf(q0.w, dy, q1.x * dz));
                float qz = fmaf(q0.z, dx, fmaf(q1.x, dy, q1.y * dz));
                float dqd = fmaf(dx, qx, fmaf(dy, qy, dz * qz));
                float r5 = ri2 * ri2 * ri;
                float phq = -0.5f * r5 * dqd;          // quadrupole potential term
                float coef = (mr - 5.f * phq) * ri2;   // radial factor of a
                sp[q] += mr - phq;
                sx[q] = fmaf(dx, coef, fmaf(-qx, r5, sx[q]));
                sy[q] = fmaf(dy, coef, fmaf(-qy, r5, sy[q]));
                sz[q] = fmaf(dz, coef, fmaf(-qz, r5, sz[q]));
            } else {
                float mr3 = mr * ri2;
                sp[q] += mr;
                sx[q] = fmaf(dx, mr3, sx[q]);
                sy[q] = fmaf(dy, mr3, sy[q]);
                sz[q] = fmaf(dz, mr3, sz[q]);
            }
        }
    }
#pragma unroll
    for (int q = 0; q < NQ; q++) {
        fold(L.axh[q], L.axl[q], sx[q]);
        fold(L.ayh[q], L.ayl[q], sy[q]);
        fold(L.azh[q], L.azl[q], sz[q]);
        fold(L.ph[q], L.pl[q], sp[q]);
    }
}

// Tiles for the first nq particle slots of each lane (groups of at most
// 64 nq points leave the other slots idle).
template <int NPL, bool SOFT0>
__device__ __forceinline__ void p2p_slots(int nq, Lane<NPL> &L, const float4 *src, int cnt, float eps2)
{
    if (nq <= 1) eval_p2p<NPL, 1, SOFT0>(L, src, cnt, eps2);
    else if (NPL > 2 && nq == 2) eval_p2p<NPL, (NPL > 2 ? 2 : NPL), SOFT0>(L, src, cnt, eps2);
    else if (NPL > 3 && nq == 3) eval_p2p<NPL, (NPL > 3 ? 3 : NPL), SOFT0>(L, src, cnt, eps2);
    else eval_p2p<NPL, NPL, SOFT0>(L, src, cnt, eps2);
}

template <int NPL, bool QUAD, bool SOFT0>
__device__ __forceinline__ void p2c_slots(int nq, Lane<NPL> &L, const float4 *src, int cnt, float eps2)
{
    if (nq <= 1) eval_p2c<NPL, 1, QUAD, SOFT0>(L, src, cnt, eps2);
    else if (NPL > 2 && nq == 2) eval_p2c<NPL, (NPL > 2 ? 2 : NPL), QUAD, SOFT0>(L, src, cnt, eps2);
    else if (NPL > 3 && nq == 3) eval_p2c<NPL, (NPL > 3 ? 3 : NPL), QUAD, SOFT0>(L, src, cnt, eps2);
    else eval_p2c<NPL, NPL, QUAD, SOFT0>(L, src, cnt, eps2);
}

// Local expansion coefficients, in the order G0, G1 (x y z), G2 (xx xy xz yy
// yz zz), G3 (xxx xxy xxz xyy xyz xzz yyy yyz yzz zzz).
constexpr int kLoc = 20;

// v + (v of another lane chosen by the DPP control CTRL); rows outside RM add 0.
template <int CTRL, int RM = 0xF>
__device__ __forceinline__ float dpp_add(float v)
{
    return v + __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, RM, 0xF, true));
}

// Sum over the wave, valid in lane 63: quad swaps (xor 1, xor 2), half-row
// and row mirrors (sums of 8 and 16 lanes), then row 15 -> next row and
// lane 31 -> rows 2-3 broadcasts.  All VALU, no LDS round trips.
__device__ __forceinline__ float wave_sum63(float v)
{
    v = dpp_add<0xB1>(v);          // quad_perm [1,0,3,2]
    v = dpp_add<0x4E>(v);          // quad_perm [2,3,0,1]
    v = dpp_add<0x141>(v);         // row_half_mirror
    v = dpp_add<0x140>(v);         // row_mirror
    v = dpp_add<0x142, 0xA>(v);    // row_bcast:15 into rows 1 and 3
    v = dpp_add<0x143, 0xC>(v);    // row_bcast:31 into rows 2 and 3
    return v;
}

// Adds the first cnt far cells of F (c.o.m. relative to the centre, mass) to
// the group's coefficients row[kLoc]: lane k takes cell k, then the wave sums.
__device__ __forceinline__ void local_flush(const float4 *F, int cnt, float *row, float eps2, float ell)
{
    const int lane = lane_id();
    const float4 e = lane < cnt ? F[lane] : make_float4(1.f, 0.f, 0.f, 0.f);   // massless: no terms
    const float rho2 = fmaf(e.x, e.x, fmaf(e.y, e.y, fmaf(e.z, e.z, eps2)));
    const float ri = rsq(rho2);
    const float nx = e.x * ri, ny = e.y * ri, nz = e.z * ri;
    const float F0 = e.w * ri * ri;
    const float sc = ell * ri;
    const float Fs = F0 * sc, Fs2 = Fs * sc;
    const float e3 = 3.f * Fs, t15 = 15.f * Fs2, f3 = 3.f * Fs2;
    // one coefficient at a time: its wave sum, added by lane 63
    auto put = [&](int k, float v) {
        v = wave_sum63(v);
        if (lane == 63) row[k] += v;
    };
    put(0, e.w * ri);
    put(1, F0 * nx);
    put(2, F0 * ny);
    put(3, F0 * nz);
    put(4, fmaf(e3 * nx, nx, -Fs));
    put(5, e3 * nx * ny);
    put(6, e3 * nx * nz);
    put(7, fmaf(e3 * ny, ny, -Fs));
    put(8, e3 * ny * nz);
    put(9, fmaf(e3 * nz, nz, -Fs));
    put(10, fmaf(t15 * nx, nx * nx, -3.f * f3 * nx));
    put(11, fmaf(t15 * nx, nx * ny, -f3 * ny));
    put(12, fmaf(t15 * nx, nx * nz, -f3 * nz));
    put(13, fmaf(t15 * nx, ny * ny, -f3 * nx));
    put(14, t15 * nx * (ny * nz));
    put(15, fmaf(t15 * nx, nz * nz, -f3 * nx));
    put(16, fmaf(t15 * ny, ny * ny, -3.f * f3 * ny));
    put(17, fmaf(t15 * ny, ny * nz, -f3 * nz));
    put(18, fmaf(t15 * ny, nz * nz, -f3 * ny));
    put(19, fmaf(t15 * nz, nz * nz, -3.f * f3 * nz));
}

// Field and potential sum (+m/r convention of the tiles) of the summed
// coefficients G at the member u = (x - centre)/l.
__device__ __forceinline__ void local_eval(const float *G, float ux, float uy, float uz, float ell, float &ax,
                                           float &ay, float &az, float &ph)
{
    const float gx = fmaf(G[4], ux, fmaf(G[5], uy, G[6] * uz));
    const float gy = fmaf(G[5], ux, fmaf(G[7], uy, G[8] * uz));
    const float gz = fmaf(G[6], ux, fmaf(G[8], uy, G[9] * uz));
    const float xx = ux * ux, yy = uy * uy, zz = uz * uz;
    const float xy2 = 2.f * ux * uy, xz2 = 2.f * ux * uz, yz2 = 2.f * uy * uz;
    const float hx = fmaf(G[10], xx, fmaf(G[11], xy2, fmaf(G[12], xz2, fmaf(G[13], yy, fmaf(G[14], yz2, G[15] * zz)))));
    const float hy = fmaf(G[11], xx, fmaf(G[13], xy2, fmaf(G[14], xz2, fmaf(G[16], yy, fmaf(G[17], yz2, G[18] * zz)))));
    const float hz = fmaf(G[12], xx, fmaf(G[14], xy2, fmaf(G[15], xz2, fmaf(G[17], yy, fmaf(G[18], yz2, G[19] * zz)))));
    ax = fmaf(0.5f, hx, G[1] + gx);
    ay = fmaf(0.5f, hy, G[2] + gy);
    az = fmaf(0.5f, hz, G[3] + gz);
    const float px = fmaf(1.f / 6.f, hx, fmaf(0.5f, gx, G[1]));
    const float py = fmaf(1.f / 6.f, hy, fmaf(0.5f, gy, G[2]));
    const float pz = fmaf(1.f / 6.f, hz, fmaf(0.5f, gz, G[3]));
    ph = fmaf(ell, fmaf(ux, px, fmaf(uy, py, uz * pz)), G[0]);
}

struct WalkArgs {
    const float4 *thi, *tlo;
    const int *tself, *tinv;
    int selfmode;
    const uint32_t *tperm;
    int64_t ntest;
    const unsigned long long *stat;
    const int *gfirst, *gcount;
    const float4 *bhi, *blo;
    const float4 *cA, *cB, *cQ0, *cQ1;
    const int *sub, *nkids;
    const int2 *fnb;
    int nleaf;
    const uint8_t *level;
    int nb;
    const unsigned long long *ncellp;
    float tolsq, eps2;
    int *frontier;
    int fcap;
    int *gctr;
    unsigned long long *counters;
    double *phi, *acc;
    double4 *sacc;        // lmode 1: (a, phi) per test point in sorted order (the correction pass writes phi, acc)
    // learned correction: record, per group, the accepted cells whose
    // size/distance lies in [lr_lo, lr_hi] (squared bounds) into a per-wave
    // region: [used][g, n, cx, cy, cz, cell ids ...] ...
    int lmode;            // 0 off, 1 every group, 2 every lstride-th group
    int lstride;
    float lr_lo2, lr_hi2;
    int *lrec;
    int lcap;
    // batches: system offsets, per-system (tol^2, eps^2) and walk roots
    int nsys;
    const int64_t *boff;
    const float2 *sysp;
    const int *sysroot;
    // treecode1 criteria: root box; cB.w holds each cell's Rcrit2
    const double4 *box;
    // super-groups: K consecutive groups share one traversal; per-wave
    // scratch for the shared lists (shcap entries each)
    int K;
    float lalpha2;        // local expansions: local_alpha^2 (0 = off)
    float4 *shlist;       // [wave][shcap P entries][shcap * S C entries]
    int *shid;            // [wave][shcap body ids of the P entries][shcap deferred nodes]
    int shcap;
};

// CRIT: treecode1's accept (treegrav.c:137-156) for the whole group.  With
// U the union of the side-s cubes (s = the cell's side, cubes of the cell's
// level) that hold the group's points, a cell is accepted when the squared
// distance from its c.o.m. to U exceeds its Rcrit2 and, along some axis, the
// c.o.m. lies more than s beyond U.  Every body's own sink cube lies in U, so
// each body's treecode1 test passes whenever the group's does.  U is formed
// in integer coordinates of the finest level (2^21 per side).
// The monopole variants with three points per lane ask for 4 waves per SIMD
// (128 VGPRs): left to itself the compiler gives them ~160 and 3 waves.
template <int NPL, bool QUAD, bool SOFT0, bool LREC, bool CRIT, bool LOCAL>
__global__ __launch_bounds__(kTpb) __attribute__((amdgpu_waves_per_eu(!QUAD && !LREC && NPL == 3 ? 4 : 1))) void k_walk(WalkArgs a)
{
    static_assert(!(LOCAL && QUAD), "local expansions take monopole cells");
    constexpr int G = 64 * NPL;
    constexpr int S = QUAD ? 3 : 1;
    __shared__ float4 sP[kWaves][kListCap];
    __shared__ float4 sC[kWaves][kListCap * S];
    __shared__ unsigned char sSelf[kWaves][G];
    __shared__ float4 sF[LOCAL ? kWaves : 1][LOCAL ? kListCap : 1];    // far cells awaiting the expansion
    __shared__ float sG[LOCAL ? kWaves : 1][kLoc];                     // the group's coefficients
    // the first kLFr entries of each frontier level live in LDS, the rest in
    // the wave's global frontier (index - kLFr): most levels never leave LDS
    constexpr int kLFr = QUAD ? 192 : 384;
    __shared__ int sFr[kWaves][2][kLFr];
    const int w = threadIdx.x >> 6, lane = lane_id();
    float4 *P = sP[w];
    float4 *Cl = sC[w];
    int *fr0 = a.frontier + (int64_t)(blockIdx.x * kWaves + w) * 2 * a.fcap;
    int *fr1 = fr0 + a.fcap;
    int *lf0 = sFr[w][0], *lf1 = sFr[w][1];

    int lused = 1;                       // used ints of this wave's record region
    for (;;) {
        int g = 0;
        if (lane == 0) g = atomicAdd(a.gctr, 1);
        g = __shfl(g, 0, 64);
        if (g >= (int)a.stat[S_NGROUPS]) break;
        const int64_t gbase = a.gfirst[g];
        const int gcount = a.gcount[g];
        const int nq = (gcount + 63) >> 6;             // particle slots in use (wave-uniform)
        // opening and softening of this group's system
        float tolsq = a.tolsq, eps2 = a.eps2;
        int sys = -1;
        if (a.nsys) {
            int lo = 0, hi = a.nsys;
            while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (a.boff[mid] <= gbase) lo = mid;
                else hi = mid;
            }
            sys = lo;
            const float2 pp = a.sysp[sys];
            tolsq = pp.x;
            eps2 = pp.y;
        }

        // particles of this lane; inactive slots copy the first particle
        float4 h[NPL], l[NPL];
        bool act[NPL];
#pragma unroll
        for (int q = 0; q < NPL; q++) {
            int k = q * 64 + lane;
            act[q] = k < gcount;
            int64_t i = gbase + (act[q] ? k : 0);
            h[q] = a.thi[i];
            l[q] = a.tlo[i];
        }
        float mnx = h[0].x, mny = h[0].y, mnz = h[0].z, mxx = mnx, mxy = mny, mxz = mnz;
#pragma unroll
        for (int q = 1; q < NPL; q++) {
            mnx = fminf(mnx, h[q].x); mny = fminf(mny, h[q].y); mnz = fminf(mnz, h[q].z);
            mxx = fmaxf(mxx, h[q].x); mxy = fmaxf(mxy, h[q].y); mxz = fmaxf(mxz, h[q].z);
        }
        const float cx = 0.5f * (wave_min(mnx) + wave_max(mxx));
        const float cy = 0.5f * (wave_min(mny) + wave_max(mxy));
        const float cz = 0.5f * (wave_min(mnz) + wave_max(mxz));
        Lane<NPL> L;
        float bminx = 3.0e38f, bminy = 3.0e38f, bminz = 3.0e38f, bmaxx = -3.0e38f, bmaxy = -3.0e38f, bmaxz = -3.0e38f;
#pragma unroll
        for (int q = 0; q < NPL; q++) {
            L.px[q] = (h[q].x - cx) + l[q].x;
            L.py[q] = (h[q].y - cy) + l[q].y;
            L.pz[q] = (h[q].z - cz) + l[q].z;
            L.axh[q] = L.axl[q] = L.ayh[q] = L.ayl[q] = L.azh[q] = L.azl[q] = L.ph[q] = L.pl[q] = 0.f;
            bminx = fminf(bminx, L.px[q]); bminy = fminf(bminy, L.py[q]); bminz = fminf(bminz, L.pz[q]);
            bmaxx = fmaxf(bmaxx, L.px[q]); bmaxy = fmaxf(bmaxy, L.py[q]); bmaxz = fmaxf(bmaxz, L.pz[q]);
        }
        bminx = wave_min(bminx); bminy = wave_min(bminy); bminz = wave_min(bminz);
        bmaxx = wave_max(bmaxx); bmaxy = wave_max(bmaxy); bmaxz = wave_max(bmaxz);
        // local expansions: group radius^2 about the centre (inactive slots
        // copy the first particle), its length unit, this lane's column
        float lrg2 = 0.f, lell = 0.f;
        float4 *Fl = sF[LOCAL ? w : 0];
        float *lrow = sG[LOCAL ? w : 0];
        int nfl = 0;
        if (LOCAL) {
#pragma unroll
            for (int q = 0; q < NPL; q++)
                lrg2 = fmaxf(lrg2, fmaf(L.px[q], L.px[q], fmaf(L.py[q], L.py[q], L.pz[q] * L.pz[q])));
            lrg2 = wave_max(lrg2);
            lell = sqrtf(lrg2);
            if (lane < kLoc) lrow[lane] = 0.f;
        }
        int ulo[3] = {0, 0, 0}, uhi[3] = {0, 0, 0};
        float uoff[3] = {0.f, 0.f, 0.f}, uunit = 0.f, urs = 0.f;
        if (CRIT) {
            const double4 bb = *a.box;
            const double org[3] = {bb.x, bb.y, bb.z}, cen[3] = {cx, cy, cz};
            const double sc = 2097152.0 / bb.w;
#pragma unroll
            for (int k = 0; k < 3; k++) {
                int lo = INT32_MAX, hi = INT32_MIN;
#pragma unroll
                for (int q = 0; q < NPL; q++) {
                    const float hk = k == 0 ? h[q].x : (k == 1 ? h[q].y : h[q].z);
                    const float lk = k == 0 ? l[q].x : (k == 1 ? l[q].y : l[q].z);
                    double f = floor((((double)hk + (double)lk) - org[k]) * sc);
                    f = fmin(fmax(f, -536870912.0), 536870912.0);
                    lo = min(lo, (int)f);
                    hi = max(hi, (int)f);
                }
                ulo[k] = wave_min_i(lo);
                uhi[k] = wave_max_i(hi);
                uoff[k] = (float)((org[k] + (double)ulo[k] * (bb.w / 2097152.0)) - cen[k]);
            }
            uunit = (float)(bb.w / 2097152.0);
            urs = (float)bb.w;
        }
#pragma unroll
        for (int q = 0; q < NPL; q++) sSelf[w][q * 64 + lane] = 0;

        int np = 0, nc = 0, nf = 0;
        long long tot_pp = 0, tot_pc = 0, tot_lc = 0;
        const bool lrecord = LREC && (a.lmode == 1 || (a.lmode == 2 && g % a.lstride == 0));
        int *lreg = a.lrec + (int64_t)(blockIdx.x * kWaves + w) * a.lcap;
        int lhead = 0, lcnt = 0;
        if (lrecord) {
            lhead = lused;
            if (lhead + 5 + 64 > a.lcap) {
                if (lane == 0) atomicOr(&a.counters[S_FLAGS], (unsigned long long)F_LREC_OVF);
            } else if (lane == 0) {
                lreg[lhead] = g;
                lreg[lhead + 2] = __float_as_int(cx);
                lreg[lhead + 3] = __float_as_int(cy);
                lreg[lhead + 4] = __float_as_int(cz);
            }
        }
        if (sys >= 0) {
            if (lane == 0) lf0[0] = a.sysroot[sys];
            nf = 1;
        } else if (*a.ncellp == 0) {
            if (a.nb == 1 && lane == 0) lf0[0] = 0;
            nf = a.nb == 1 ? 1 : 0;
        } else {
            if (lane == 0) lf0[0] = ~0;
            nf = 1;
        }
        bool overflow = false;
        while (nf > 0 && !overflow) {
            int nn = 0;
            for (int base = 0; base < nf; base += 64) {
                wave_lds_sync();
                const int k = base + lane;
                const bool v = k < nf;
                const int node = v ? (k < kLFr ? lf0[k] : fr0[k - kLFr]) : 0;
                const bool isBody = v && node >= 0;
                const bool isCell = v && node < 0;
                bool open = false, accept = false;
                int c = 0, nk = 0, bfirst = -1;
                float4 A = make_float4(0.f, 0.f, 0.f, 0.f);
                float rx = 0.f, ry = 0.f, rz = 0.f;
                int4 s0 = make_int4(kNoChild, kNoChild, kNoChild, kNoChild), s1 = s0;
                if (isCell) {
                    c = ~node;
                    s0 = reinterpret_cast<const int4 *>(a.sub)[2 * (int64_t)c];
                    s1 = reinterpret_cast<const int4 *>(a.sub)[2 * (int64_t)c + 1];
                    A = a.cA[c];
                    const float4 B = a.cB[c];
                    rx = (A.x - cx) + B.x;
                    ry = (A.y - cy) + B.y;
                    rz = (A.z - cz) + B.z;
                    float dx = fmaxf(fmaxf(bminx - rx, rx - bmaxx), 0.f);
                    float dy = fmaxf(fmaxf(bminy - ry, ry - bmaxy), 0.f);
                    float dz = fmaxf(fmaxf(bminz - rz, rz - bmaxz), 0.f);
                    float d2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
                    const int2 fn = a.fnb[c];
                    const int lv = a.level[c];
                    const bool small = fn.y <= a.nleaf || lv == kMaxLevel;
                    if (CRIT) {
                        const int sh = kMaxLevel - lv, step = 1 << sh;
                        const float side = ldexpf(urs, -lv), rr[3] = {rx, ry, rz};
                        float dd[3];
#pragma unroll
                        for (int k = 0; k < 3; k++) {
                            const float flo = (float)((ulo[k] >> sh) * step - ulo[k]) * uunit + uoff[k];
                            const float fhi = (float)(((uhi[k] >> sh) + 1) * step - ulo[k]) * uunit + uoff[k];
                            dd[k] = fmaxf(fmaxf(flo - rr[k], rr[k] - fhi), 0.f);
                        }
                        const float usq = fmaf(dd[0], dd[0], fmaf(dd[1], dd[1], dd[2] * dd[2]));
                        const bool adj_ok = fmaxf(dd[0], fmaxf(dd[1], dd[2])) > side;
                        open = !(usq > B.w && adj_ok) || fn.y <= a.nleaf;
                    } else {
                        open = tolsq * d2 < B.w || fn.y <= a.nleaf;
                    }
                    accept = !open;
                    if (open) {
                        int cnt = 0;
                        const int so[8] = {s0.x, s0.y, s0.z, s0.w, s1.x, s1.y, s1.z, s1.w};
#pragma unroll
                        for (int o = 0; o < 8; o++) cnt += so[o] != kNoChild;
                        nk = small ? fn.y : cnt;
                    }
                    bfirst = small ? fn.x : -1;
                }
                // the learned record keeps every accepted cell in its range,
                // also those the expansion takes (their residual is added on
                // top of whichever evaluates their monopole)
                const bool accepted = accept;
                // far accepted cells -> the local expansion's list
                if (LOCAL) {
                    const bool far = accept && lrg2 < a.lalpha2 * fmaf(rx, rx, fmaf(ry, ry, rz * rz));
                    const unsigned long long m = __ballot(far);
                    if (far) {
                        Fl[nfl + mbcnt(m)] = make_float4(rx, ry, rz, A.w);
                        accept = false;
                    }
                    nfl += __popcll(m);
                    tot_lc += __popcll(m);
                }
                // bodies -> P2P list
                {
                    const unsigned long long m = __ballot(isBody);
                    if (isBody) {
                        const float4 bh = a.bhi[node], bl = a.blo[node];
                        P[np + mbcnt(m)] = make_float4((bh.x - cx) + bl.x, (bh.y - cy) + bl.y, (bh.z - cz) + bl.z, bh.w);
                        // the group member that is this body, if any
                        const int64_t rel = a.selfmode == 1   ? (int64_t)node - gbase
                                            : a.selfmode == 2 ? (int64_t)a.tinv[node] - gbase
                                                              : -1;
                        if (rel >= 0 && rel < gcount) sSelf[w][rel] = 1;
                    }
                    np += __popcll(m);
                }
                // accepted cells in the learned range -> this group's record
                if (LREC && lrecord) {
                    const float r2 = fmaf(rx, rx, fmaf(ry, ry, rz * rz));
                    const float q = accepted ? a.cB[c].w : 0.f;
                    const bool take = accepted && q >= a.lr_lo2 * r2 && q <= a.lr_hi2 * r2;
                    const unsigned long long m = __ballot(take);
                    const int at = lhead + 5 + lcnt;
                    if (at + 64 <= a.lcap) {
                        if (take) lreg[at + mbcnt(m)] = c;
                        lcnt += __popcll(m);
                    } else if (m && lane == 0) {
                        atomicOr(&a.counters[S_FLAGS], (unsigned long long)F_LREC_OVF);
                    }
                }
                // accepted cells -> P2C list
                {
                    const unsigned long long m = __ballot(accept);
                    if (accept) {
                        const int slot = nc + mbcnt(m);
                        Cl[S * slot] = make_float4(rx, ry, rz, A.w);
                        if (QUAD) {
                            Cl[S * slot + 1] = a.cQ0[c];
                            Cl[S * slot + 2] = a.cQ1[c];
                        }
                    }
                    nc += __popcll(m);
                }
                // opened cells -> next frontier
                {
                    int tot;
                    const int off = wave_excl_scan(nk, &tot);
                    if (nn + tot > kLFr + a.fcap) {
                        overflow = true;
                        break;
                    }
                    if (open) {
                        const int at = nn + off;
                        auto put = [&](int j, int r) {
                            if (at + j < kLFr) lf1[at + j] = r;
                            else fr1[at + j - kLFr] = r;
                        };
                        if (bfirst >= 0) {
                            for (int j = 0; j < nk; j++) put(j, bfirst + j);
                        } else {
                            const int so[8] = {s0.x, s0.y, s0.z, s0.w, s1.x, s1.y, s1.z, s1.w};
                            int j = 0;
#pragma unroll
                            for (int o = 0; o < 8; o++) {
                                int r = so[o];
                                if (r != kNoChild) put(j++, r);
                            }
                        }
                    }
                    nn += tot;
                }
                wave_lds_sync();
                while (np >= 64) {
                    p2p_slots<NPL, SOFT0>(nq, L, P, 64, eps2);
                    tot_pp += 64;
                    np -= 64;
                    wave_lds_sync();
                    float4 t = P[64 + (lane < np ? lane : 0)];
                    wave_lds_sync();
                    if (lane < np) P[lane] = t;
                    wave_lds_sync();
                }
                while (nc >= 64) {
                    p2c_slots<NPL, QUAD, SOFT0>(nq, L, Cl, 64, eps2);
                    tot_pc += 64;
                    nc -= 64;
                    wave_lds_sync();
                    float4 t0 = Cl[S * (64 + (lane < nc ? lane : 0))];
                    float4 t1, t2;
                    if (QUAD) {
                        t1 = Cl[S * (64 + (lane < nc ? lane : 0)) + 1];
                        t2 = Cl[S * (64 + (lane < nc ? lane : 0)) + 2];
                    }
                    wave_lds_sync();
                    if (lane < nc) {
                        Cl[S * lane] = t0;
                        if (QUAD) {
                            Cl[S * lane + 1] = t1;
                            Cl[S * lane + 2] = t2;
                        }
                    }
                    wave_lds_sync();
                }
                if (LOCAL && nfl >= 64) {
                    local_flush(Fl, 64, lrow, eps2, lell);
                    nfl -= 64;
                    wave_lds_sync();
                    const float4 t = Fl[64 + (lane < nfl ? lane : 0)];
                    wave_lds_sync();
                    if (lane < nfl) Fl[lane] = t;
                    wave_lds_sync();
                }
            }
            int *tmp = fr0;
            fr0 = fr1;
            fr1 = tmp;
            tmp = lf0;
            lf0 = lf1;
            lf1 = tmp;
            nf = nn;
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        }
        if (overflow) {
            if (lane == 0) atomicOr(&a.counters[S_FLAGS], (unsigned long long)F_FRONTIER);
            continue;
        }
        wave_lds_sync();
        if (np > 0) p2p_slots<NPL, SOFT0>(nq, L, P, np, eps2);
        if (nc > 0) p2c_slots<NPL, QUAD, SOFT0>(nq, L, Cl, nc, eps2);
        tot_pp += np;
        tot_pc += nc;
        wave_lds_sync();
        if (LOCAL && tot_lc > 0) {
            // the last far cells, then the expansion at every member
            if (nfl > 0) local_flush(Fl, nfl, lrow, eps2, lell);
            wave_lds_sync();
            float Gs[kLoc];
#pragma unroll
            for (int k = 0; k < kLoc; k++) Gs[k] = lrow[k];
            const float linv = lell > 0.f ? 1.f / lell : 0.f;
#pragma unroll
            for (int q = 0; q < NPL; q++) {
                float ex, ey, ez, ep;
                local_eval(Gs, L.px[q] * linv, L.py[q] * linv, L.pz[q] * linv, lell, ex, ey, ez, ep);
                fold(L.axh[q], L.axl[q], ex);
                fold(L.ayh[q], L.ayl[q], ey);
                fold(L.azh[q], L.azl[q], ez);
                fold(L.ph[q], L.pl[q], ep);
            }
        }

        if (lrecord && lhead + 5 + 64 <= a.lcap) {
            lused = lhead + 5 + lcnt;
            if (lane == 0) {
                lreg[lhead + 1] = lcnt;
                lreg[0] = lused;
            }
        }
        int nself = 0;
#pragma unroll
        for (int q = 0; q < NPL; q++) {
            if (!act[q]) continue;
            const int k = q * 64 + lane;
            double ph = (double)L.ph[q] + (double)L.pl[q];
            if (sSelf[w][k]) {
                nself++;
                // the self pair added m rsq(eps^2) unless eps = 0 masked it
                // (a batch with an unsoftened system runs the masking variant
                // for every system)
                if (eps2 > 0.f) {
                    const int64_t tb = a.selfmode == 1 ? gbase + k : a.tself[gbase + k];
                    ph -= (double)(a.bhi[tb].w * rsq(eps2));
                }
            }
            const double ax = (double)L.axh[q] + (double)L.axl[q];
            const double ay = (double)L.ayh[q] + (double)L.ayl[q];
            const double az = (double)L.azh[q] + (double)L.azl[q];
            if (LREC && a.lmode == 1) {
                // coalesced; the correction pass adds the group's residual and
                // writes the input-order arrays once (sample walks, lmode 2,
                // have no such pass)
                a.sacc[gbase + k] = make_double4(ax, ay, az, -ph);
            } else {
                const int64_t o = a.tperm[gbase + k];
                a.phi[o] = -ph;
                a.acc[3 * o] = ax;
                a.acc[3 * o + 1] = ay;
                a.acc[3 * o + 2] = az;
            }
        }
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) nself += __shfl_xor(nself, d, 64);
        if (lane == 0) {
            atomicAdd(&a.counters[0], (unsigned long long)(tot_pp * gcount - nself));
            atomicAdd(&a.counters[1], (unsigned long long)((tot_pc + tot_lc) * gcount));
            if (LOCAL) atomicAdd(&a.counters[S_NLOC], (unsigned long long)(tot_lc * gcount));
        }
    }
}


// ---- super-groups: one traversal shared by K consecutive groups --------------
//
// A wave takes K consecutive groups (Morton-adjacent) and first walks the
// tree for all of them at once, testing every node against each group's
// bounding box (hackforce's rule, grav.c:122-131, applied per group as in
// k_walk): a cell that every group would open is opened, a cell that every
// group would accept goes to a shared P2C list, a body to a shared P2P list,
// and a cell on which the groups disagree to a deferred list.  Then each
// group in turn consumes the shared lists and walks on from the deferred
// nodes alone.  Every group ends with exactly the interactions k_walk gives
// it (the counters are equal); only the order of the sums differs.  The
// shared lists live in a per-wave global scratch region, coordinates
// relative to the super-group's centre.
constexpr int kMaxK = 4;

template <int NPL, bool QUAD, bool SOFT0>
__global__ __launch_bounds__(kTpb) void k_walk_sg(WalkArgs a)
{
    constexpr int G = 64 * NPL;
    constexpr int S = QUAD ? 3 : 1;
    __shared__ float4 sP[kWaves][kListCap];
    __shared__ float4 sC[kWaves][kListCap * S];
    __shared__ unsigned char sSelf[kWaves][G];
    __shared__ float4 sBox[kWaves][kMaxK][2];   // each group's bounding box (lo, hi), relative to the centre
    const int w = threadIdx.x >> 6, lane = lane_id();
    const int64_t wid = (int64_t)blockIdx.x * kWaves + w;
    float4 *P = sP[w];
    float4 *Cl = sC[w];
    int *frA = a.frontier + wid * 2 * a.fcap;
    int *frB = frA + a.fcap;
    float4 *shP = a.shlist + wid * (int64_t)a.shcap * (1 + S);
    float4 *shC = shP + a.shcap;
    int *shId = a.shid + wid * 2 * (int64_t)a.shcap;
    int *shDef = shId + a.shcap;
    const int K = a.K;
    const float tolsq = a.tolsq, eps2 = a.eps2;
    for (;;) {
        int sg = 0;
        if (lane == 0) sg = atomicAdd(a.gctr, 1);
        sg = __shfl(sg, 0, 64);
        const int ng = (int)a.stat[S_NGROUPS];
        if ((int64_t)sg * K >= ng) break;
        const int g0 = sg * K, gn = min(K, ng - g0);

        // centre of the super-group (midpoint of the hi parts' extent)
        float mnx = 3.0e38f, mny = 3.0e38f, mnz = 3.0e38f, mxx = -3.0e38f, mxy = -3.0e38f, mxz = -3.0e38f;
        for (int i = 0; i < gn; i++) {
            const int64_t gb = a.gfirst[g0 + i];
            const int gc = a.gcount[g0 + i];
#pragma unroll
            for (int q = 0; q < NPL; q++) {
                const int k = q * 64 + lane;
                const float4 h = a.thi[gb + (k < gc ? k : 0)];
                mnx = fminf(mnx, h.x); mny = fminf(mny, h.y); mnz = fminf(mnz, h.z);
                mxx = fmaxf(mxx, h.x); mxy = fmaxf(mxy, h.y); mxz = fmaxf(mxz, h.z);
            }
        }
        const float cx = 0.5f * (wave_min(mnx) + wave_max(mxx));
        const float cy = 0.5f * (wave_min(mny) + wave_max(mxy));
        const float cz = 0.5f * (wave_min(mnz) + wave_max(mxz));
        // each group's bounding box relative to that centre (in LDS: the
        // values are wave-uniform and would cost 6 VGPRs per group)
        wave_lds_sync();
        for (int i = 0; i < gn; i++) {
            float lo0 = 3.0e38f, lo1 = 3.0e38f, lo2 = 3.0e38f, hi0 = -3.0e38f, hi1 = -3.0e38f, hi2 = -3.0e38f;
            const int64_t gb = a.gfirst[g0 + i];
            const int gc = a.gcount[g0 + i];
#pragma unroll
            for (int q = 0; q < NPL; q++) {
                const int k = q * 64 + lane;
                const int64_t t = gb + (k < gc ? k : 0);
                const float4 h = a.thi[t], l = a.tlo[t];
                const float px = (h.x - cx) + l.x, py = (h.y - cy) + l.y, pz = (h.z - cz) + l.z;
                lo0 = fminf(lo0, px); lo1 = fminf(lo1, py); lo2 = fminf(lo2, pz);
                hi0 = fmaxf(hi0, px); hi1 = fmaxf(hi1, py); hi2 = fmaxf(hi2, pz);
            }
            lo0 = wave_min(lo0); lo1 = wave_min(lo1); lo2 = wave_min(lo2);
            hi0 = wave_max(hi0); hi1 = wave_max(hi1); hi2 = wave_max(hi2);
            if (lane == 0) {
                sBox[w][i][0] = make_float4(lo0, lo1, lo2, 0.f);
                sBox[w][i][1] = make_float4(hi0, hi1, hi2, 0.f);
            }
        }
        wave_lds_sync();

        // ---- phase A: the shared traversal
        int np = 0, nc = 0, nd = 0, nf = 0;
        bool overflow = false;
        if (*a.ncellp == 0) {
            if (a.nb == 1 && lane == 0) frA[0] = 0;
            nf = a.nb == 1 ? 1 : 0;
        } else {
            if (lane == 0) frA[0] = ~0;
            nf = 1;
        }
        int *fr0 = frA, *fr1 = frB;
        while (nf > 0 && !overflow) {
            int nn = 0;
            for (int base = 0; base < nf; base += 64) {
                const int k = base + lane;
                const bool v = k < nf;
                const int node = v ? fr0[k] : 0;
                const bool isBody = v && node >= 0;
                const bool isCell = v && node < 0;
                bool open = false, accept = false, defer = false;
                int c = 0, nk = 0, bfirst = -1;
                float4 A = make_float4(0.f, 0.f, 0.f, 0.f);
                float rx = 0.f, ry = 0.f, rz = 0.f;
                if (isCell) {
                    c = ~node;
                    A = a.cA[c];
                    const float4 B = a.cB[c];
                    rx = (A.x - cx) + B.x;
                    ry = (A.y - cy) + B.y;
                    rz = (A.z - cz) + B.z;
                    const int2 fn = a.fnb[c];
                    const int lv = a.level[c];
                    const bool small = fn.y <= a.nleaf || lv == kMaxLevel;
                    bool any_open = fn.y <= a.nleaf, all_open = true;
                    for (int i = 0; i < gn; i++) {
                        const float4 lo = sBox[w][i][0], hi = sBox[w][i][1];
                        const float dx = fmaxf(fmaxf(lo.x - rx, rx - hi.x), 0.f);
                        const float dy = fmaxf(fmaxf(lo.y - ry, ry - hi.y), 0.f);
                        const float dz = fmaxf(fmaxf(lo.z - rz, rz - hi.z), 0.f);
                        const float d2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
                        const bool o = tolsq * d2 < B.w || fn.y <= a.nleaf;
                        any_open = any_open || o;
                        all_open = all_open && o;
                    }
                    open = all_open;
                    accept = !any_open;
                    defer = any_open && !all_open;
                    if (open) nk = small ? fn.y : a.nkids[c];
                    bfirst = small ? fn.x : -1;
                }
                // bodies -> shared P2P list (with their ids for the self pairs)
                {
                    const unsigned long long m = __ballot(isBody);
                    const int cnt = __popcll(m);
                    if (np + cnt > a.shcap) {
                        overflow = true;
                        break;
                    }
                    if (isBody) {
                        const float4 bh = a.bhi[node], bl = a.blo[node];
                        const int at = np + mbcnt(m);
                        shP[at] = make_float4((bh.x - cx) + bl.x, (bh.y - cy) + bl.y, (bh.z - cz) + bl.z, bh.w);
                        shId[at] = node;
                    }
                    np += cnt;
                }
                // cells every group accepts -> shared P2C list
                {
                    const unsigned long long m = __ballot(accept);
                    const int cnt = __popcll(m);
                    if (nc + cnt > a.shcap) {
                        overflow = true;
                        break;
                    }
                    if (accept) {
                        const int at = nc + mbcnt(m);
                        shC[S * at] = make_float4(rx, ry, rz, A.w);
                        if (QUAD) {
                            shC[S * at + 1] = a.cQ0[c];
                            shC[S * at + 2] = a.cQ1[c];
                        }
                    }
                    nc += cnt;
                }
                // cells the groups disagree on -> deferred
                {
                    const unsigned long long m = __ballot(defer);
                    const int cnt = __popcll(m);
                    if (nd + cnt > a.shcap) {
                        overflow = true;
                        break;
                    }
                    if (defer) shDef[nd + mbcnt(m)] = node;
                    nd += cnt;
                }
                // cells every group opens -> next frontier
                {
                    int tot;
                    const int off = wave_excl_scan(nk, &tot);
                    if (nn + tot > a.fcap) {
                        overflow = true;
                        break;
                    }
                    if (open) {
                        int *dst = fr1 + nn + off;
                        if (bfirst >= 0) {
                            for (int j = 0; j < nk; j++) dst[j] = bfirst + j;
                        } else {
                            const int *sub = a.sub + 8 * (int64_t)c;
                            int j = 0;
#pragma unroll
                            for (int o = 0; o < 8; o++) {
                                const int r = sub[o];
                                if (r != kNoChild) dst[j++] = r;
                            }
                        }
                    }
                    nn += tot;
                }
            }
            int *tmp = fr0;
            fr0 = fr1;
            fr1 = tmp;
            nf = nn;
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        }
        if (overflow) {
            if (lane == 0) atomicOr(&a.counters[S_FLAGS], (unsigned long long)F_FRONTIER);
            continue;
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");

        // ---- phase B: each group consumes the shared lists, then walks on alone
        for (int i = 0; i < gn; i++) {
            const int g = g0 + i;
            const int64_t gbase = a.gfirst[g];
            const int gcount = a.gcount[g];
            const int nq = (gcount + 63) >> 6;
            bool act[NPL];
            Lane<NPL> L;
#pragma unroll
            for (int q = 0; q < NPL; q++) {
                const int k = q * 64 + lane;
                act[q] = k < gcount;
                const int64_t t = gbase + (act[q] ? k : 0);
                const float4 h = a.thi[t], l = a.tlo[t];
                L.px[q] = (h.x - cx) + l.x;
                L.py[q] = (h.y - cy) + l.y;
                L.pz[q] = (h.z - cz) + l.z;
                L.axh[q] = L.axl[q] = L.ayh[q] = L.ayl[q] = L.azh[q] = L.azl[q] = L.ph[q] = L.pl[q] = 0.f;
                sSelf[w][k] = 0;
            }
            const float4 blo_i = sBox[w][i][0], bhi_i = sBox[w][i][1];
            const float bminx = blo_i.x, bminy = blo_i.y, bminz = blo_i.z;
            const float bmaxx = bhi_i.x, bmaxy = bhi_i.y, bmaxz = bhi_i.z;
            wave_lds_sync();
            // shared bodies, 64 at a time through LDS; members of this group
            // among them are its self pairs
            for (int t0 = 0; t0 < np; t0 += 64) {
                const int cnt = min(64, np - t0);
                if (lane < cnt) {
                    P[lane] = shP[t0 + lane];
                    const int node = shId[t0 + lane];
                    const int64_t rel = a.selfmode == 1   ? (int64_t)node - gbase
                                        : a.selfmode == 2 ? (int64_t)a.tinv[node] - gbase
                                                          : -1;
                    if (rel >= 0 && rel < gcount) sSelf[w][rel] = 1;
                }
                wave_lds_sync();
                p2p_slots<NPL, SOFT0>(nq, L, P, cnt, eps2);
                wave_lds_sync();
            }
            for (int t0 = 0; t0 < nc; t0 += 64) {
                const int cnt = min(64, nc - t0);
                if (lane < cnt) {
#pragma unroll
                    for (int s = 0; s < S; s++) Cl[S * lane + s] = shC[S * (t0 + lane) + s];
                }
                wave_lds_sync();
                p2c_slots<NPL, QUAD, SOFT0>(nq, L, Cl, cnt, eps2);
                wave_lds_sync();
            }
            long long tot_pp = np, tot_pc = nc;
            // own walk from the deferred nodes (read from the shared region in
            // the first step, then between the two frontier buffers)
            int mp = 0, mc = 0;
            int nf2 = nd;
            const int *src = shDef;
            int *dstb = frA, *other = frB;
            bool ovf = false;
            while (nf2 > 0 && !ovf) {
                int nn = 0;
                for (int base = 0; base < nf2; base += 64) {
                    wave_lds_sync();
                    const int k = base + lane;
                    const bool v = k < nf2;
                    const int node = v ? src[k] : 0;
                    const bool isBody = v && node >= 0;
                    const bool isCell = v && node < 0;
                    bool open = false, accept = false;
                    int c = 0, nk = 0, bfirst = -1;
                    float4 A = make_float4(0.f, 0.f, 0.f, 0.f);
                    float rx = 0.f, ry = 0.f, rz = 0.f;
                    if (isCell) {
                        c = ~node;
                        A = a.cA[c];
                        const float4 B = a.cB[c];
                        rx = (A.x - cx) + B.x;
                        ry = (A.y - cy) + B.y;
                        rz = (A.z - cz) + B.z;
                        const float dx = fmaxf(fmaxf(bminx - rx, rx - bmaxx), 0.f);
                        const float dy = fmaxf(fmaxf(bminy - ry, ry - bmaxy), 0.f);
                        const float dz = fmaxf(fmaxf(bminz - rz, rz - bmaxz), 0.f);
                        const float d2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
                        const int2 fn = a.fnb[c];
                        const int lv = a.level[c];
                        const bool small = fn.y <= a.nleaf || lv == kMaxLevel;
                        open = tolsq * d2 < B.w || fn.y <= a.nleaf;
                        accept = !open;
                        if (open) nk = small ? fn.y : a.nkids[c];
                        bfirst = small ? fn.x : -1;
                    }
                    {
                        const unsigned long long m = __ballot(isBody);
                        if (isBody) {
                            const float4 bh = a.bhi[node], bl = a.blo[node];
                            P[mp + mbcnt(m)] =
                                make_float4((bh.x - cx) + bl.x, (bh.y - cy) + bl.y, (bh.z - cz) + bl.z, bh.w);
                            const int64_t rel = a.selfmode == 1   ? (int64_t)node - gbase
                                                : a.selfmode == 2 ? (int64_t)a.tinv[node] - gbase
                                                                  : -1;
                            if (rel >= 0 && rel < gcount) sSelf[w][rel] = 1;
                        }
                        mp += __popcll(m);
                    }
                    {
                        const unsigned long long m = __ballot(accept);
                        if (accept) {
                            const int slot = mc + mbcnt(m);
                            Cl[S * slot] = make_float4(rx, ry, rz, A.w);
                            if (QUAD) {
                                Cl[S * slot + 1] = a.cQ0[c];
                                Cl[S * slot + 2] = a.cQ1[c];
                            }
                        }
                        mc += __popcll(m);
                    }
                    {
                        int tot;
                        const int off = wave_excl_scan(nk, &tot);
                        if (nn + tot > a.fcap) {
                            ovf = true;
                            break;
                        }
                        if (open) {
                            int *dst = dstb + nn + off;
                            if (bfirst >= 0) {
                                for (int j = 0; j < nk; j++) dst[j] = bfirst + j;
                            } else {
                                const int *sub = a.sub + 8 * (int64_t)c;
                                int j = 0;
#pragma unroll
                                for (int o = 0; o < 8; o++) {
                                    const int r = sub[o];
                                    if (r != kNoChild) dst[j++] = r;
                                }
                            }
                        }
                        nn += tot;
                    }
                    wave_lds_sync();
                    while (mp >= 64) {
                        p2p_slots<NPL, SOFT0>(nq, L, P, 64, eps2);
                        tot_pp += 64;
                        mp -= 64;
                        wave_lds_sync();
                        float4 t = P[64 + (lane < mp ? lane : 0)];
                        wave_lds_sync();
                        if (lane < mp) P[lane] = t;
                        wave_lds_sync();
                    }
                    while (mc >= 64) {
                        p2c_slots<NPL, QUAD, SOFT0>(nq, L, Cl, 64, eps2);
                        tot_pc += 64;
                        mc -= 64;
                        wave_lds_sync();
                        float4 t0 = Cl[S * (64 + (lane < mc ? lane : 0))];
                        float4 t1, t2;
                        if (QUAD) {
                            t1 = Cl[S * (64 + (lane < mc ? lane : 0)) + 1];
                            t2 = Cl[S * (64 + (lane < mc ? lane : 0)) + 2];
                        }
                        wave_lds_sync();
                        if (lane < mc) {
                            Cl[S * lane] = t0;
                            if (QUAD) {
                                Cl[S * lane + 1] = t1;
                                Cl[S * lane + 2] = t2;
                            }
                        }
                        wave_lds_sync();
                    }
                }
                src = dstb;
                int *tmp = dstb;
                dstb = other;
                other = tmp;
                nf2 = nn;
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
            }
            if (ovf) {
                if (lane == 0) atomicOr(&a.counters[S_FLAGS], (unsigned long long)F_FRONTIER);
                continue;
            }
            wave_lds_sync();
            if (mp > 0) p2p_slots<NPL, SOFT0>(nq, L, P, mp, eps2);
            if (mc > 0) p2c_slots<NPL, QUAD, SOFT0>(nq, L, Cl, mc, eps2);
            tot_pp += mp;
            tot_pc += mc;
            wave_lds_sync();
            int nself = 0;
#pragma unroll
            for (int q = 0; q < NPL; q++) {
                if (!act[q]) continue;
                const int k = q * 64 + lane;
                double ph = (double)L.ph[q] + (double)L.pl[q];
                if (sSelf[w][k]) {
                    nself++;
                    if (eps2 > 0.f) {
                        const int64_t tb = a.selfmode == 1 ? gbase + k : a.tself[gbase + k];
                        ph -= (double)(a.bhi[tb].w * rsq(eps2));
                    }
                }
                const int64_t o = a.tperm[gbase + k];
                a.phi[o] = -ph;
                a.acc[3 * o] = (double)L.axh[q] + (double)L.axl[q];
                a.acc[3 * o + 1] = (double)L.ayh[q] + (double)L.ayl[q];
                a.acc[3 * o + 2] = (double)L.azh[q] + (double)L.azl[q];
            }
#pragma unroll
            for (int d = 32; d >= 1; d >>= 1) nself += __shfl_xor(nself, d, 64);
            if (lane == 0) {
                atomicAdd(&a.counters[0], (unsigned long long)(tot_pp * gcount - nself));
                atomicAdd(&a.counters[1], (unsigned long long)(tot_pc * gcount));
            }
            wave_lds_sync();
        }
    }
}

__global__ void k_fill_zero(double *phi, int64_t n, double *acc)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) phi[i] = 0.0;
    if (i < 3 * n) acc[i] = 0.0;
}

__global__ void k_clear_regions(int *lrec, int cap)
{
    if (threadIdx.x == 0) lrec[(int64_t)blockIdx.x * cap] = 1;
}

template <int NPL, bool QUAD, bool SOFT0>
void launch(bh_ctx *ctx, WalkArgs &a, int grid, bool crit)
{
    if (a.lalpha2 > 0.f && !QUAD) {
        if (a.lmode) k_walk<NPL, false, SOFT0, true, false, true><<<grid, kTpb, 0, ctx->stream>>>(a);
        else if (crit) k_walk<NPL, false, SOFT0, false, true, true><<<grid, kTpb, 0, ctx->stream>>>(a);
        else k_walk<NPL, false, SOFT0, false, false, true><<<grid, kTpb, 0, ctx->stream>>>(a);
    } else if (crit) k_walk<NPL, QUAD, SOFT0, false, true, false><<<grid, kTpb, 0, ctx->stream>>>(a);
    else if (a.lmode) k_walk<NPL, QUAD, SOFT0, true, false, false><<<grid, kTpb, 0, ctx->stream>>>(a);
    else if (a.K > 1) k_walk_sg<NPL, QUAD, SOFT0><<<grid, kTpb, 0, ctx->stream>>>(a);
    else k_walk<NPL, QUAD, SOFT0, false, false, false><<<grid, kTpb, 0, ctx->stream>>>(a);
}

}  // namespace

void walk_fast(bh_ctx *ctx, const bh_params &p, int lmode, int lstride)
{
    const Tree &T = ctx->tree;
    const int64_t nt = ctx->ntest;
    double *phi = ptr<double>(ctx->phi), *acc = ptr<double>(ctx->acc);
    if (T.nb == 0) {
        k_fill_zero<<<std::max<int64_t>(1, (4 * nt + 255) / 256), 256, 0, ctx->stream>>>(phi, nt, acc);
        return;
    }
    const int npl = p.group_size / 64;
    WalkArgs a{};
    a.thi = T.thi;
    a.tlo = T.tlo;
    a.tself = T.tself;
    a.tinv = T.tinv;
    a.selfmode = T.selfmode;
    a.tperm = T.tperm;
    a.ntest = nt;
    a.stat = ptr<unsigned long long>(ctx->counters);
    a.gfirst = ptr<int>(ctx->gfirst);
    a.gcount = ptr<int>(ctx->gcount);
    a.bhi = T.bhi;
    a.blo = T.blo;
    a.cA = T.cA;
    a.cB = T.cB;
    a.cQ0 = T.cQ0;
    a.cQ1 = T.cQ1;
    a.sub = T.sub;
    a.nkids = T.nkids;
    a.fnb = T.fnb;
    a.nleaf = p.nleaf;
    a.level = T.level;
    a.nb = T.nb;
    a.ncellp = T.ncellp;
    a.tolsq = (float)(p.tol * p.tol);
    a.eps2 = (float)(p.eps * p.eps);
    a.counters = ptr<unsigned long long>(ctx->counters);
    a.phi = phi;
    a.acc = acc;
    a.sacc = lmode == 1 ? ensure<double4>(ctx->sacc, std::max<int64_t>(nt, 1)) : nullptr;
    a.lmode = lmode;
    a.lstride = std::max(1, lstride);
    a.lr_lo2 = (float)(ctx->mlp.norm[1] * ctx->mlp.norm[1]);
    a.lr_hi2 = (float)(ctx->mlp.norm[2] * ctx->mlp.norm[2]);
    a.nsys = ctx->nsys;
    if (ctx->nsys) {
        a.boff = ptr<int64_t>(ctx->boff);
        a.sysp = ptr<float2>(ctx->sysp);
        a.sysroot = ptr<int>(ctx->sysroot);
    }
    // persistent grid sized for the device; each wave loops over groups
    const int64_t max_groups = (nt + 1) / 1;
    int grid = (int)std::min<int64_t>((max_groups + kWaves - 1) / kWaves, (int64_t)ctx->ncu * 8);
    grid = std::max(grid, 1);
    if (ctx->fcap == 0) ctx->fcap = 1 << 15;
    a.fcap = (int)ctx->fcap;
    a.frontier = ensure<int>(ctx->frontier, (size_t)grid * kWaves * 2 * ctx->fcap);
    // the group counter is a status word cleared at the start of the tree
    // build (every tree build is followed by exactly one fast walk)
    a.gctr = reinterpret_cast<int *>(ptr<unsigned long long>(ctx->counters) + S_GCTR);
    const int nwave = grid * kWaves;
    // super-groups: the plain walk of a single system with the hackcode1
    // criterion (learned recordings, treecode1 criteria, batches and local
    // expansions walk group by group)
    // local expansions: the plain walk and the learned walk (lmode 1), not
    // the sample walks
    a.lalpha2 = (lmode <= 1 && !p.usequad) ? (float)(p.local_alpha * p.local_alpha) : 0.f;
    a.K = (lmode == 0 && p.criterion == BH_CRIT_HACK && ctx->nsys == 0 && a.lalpha2 == 0.f)
              ? std::max(1, std::min(p.supergroup, kMaxK))
              : 1;
    if (a.K > 1) {
        const int S = p.usequad ? 3 : 1;
        a.shcap = (int)std::max<int64_t>(ctx->fcap / 4, 1024);
        a.shlist = ensure<float4>(ctx->shlist, (size_t)nwave * a.shcap * (1 + S));
        a.shid = ensure<int>(ctx->shid, (size_t)nwave * 2 * a.shcap);
    }
    if (lmode) {
        if (ctx->lcap == 0) ctx->lcap = 1 << 16;
        a.lcap = (int)ctx->lcap;
        a.lrec = ensure<int>(ctx->lrec, (size_t)nwave * ctx->lcap);
        k_clear_regions<<<nwave, 64, 0, ctx->stream>>>(a.lrec, a.lcap);
    } else {
        a.lcap = 0;
        a.lrec = a.gctr;   // not written without records
    }
    ctx->lnwave = nwave;
    const bool soft0 = ctx->nsys ? ctx->sys_eps0 : p.eps == 0.0;
    // walk-kernel start: without records nothing follows the caller's
    // tree-end event, which then stands for it (an event record costs a few us)
    if (!ctx->capturing && lmode) BH_HIP(hipEventRecord(ctx->ev[4], ctx->stream));
    const bool crit = p.criterion != BH_CRIT_HACK;
    a.box = T.dbox;
    if (npl == 3) {
        if (p.usequad) soft0 ? launch<3, true, true>(ctx, a, grid, crit) : launch<3, true, false>(ctx, a, grid, crit);
        else soft0 ? launch<3, false, true>(ctx, a, grid, crit) : launch<3, false, false>(ctx, a, grid, crit);
    } else if (npl == 4) {
        if (p.usequad) soft0 ? launch<4, true, true>(ctx, a, grid, crit) : launch<4, true, false>(ctx, a, grid, crit);
        else soft0 ? launch<4, false, true>(ctx, a, grid, crit) : launch<4, false, false>(ctx, a, grid, crit);
    } else if (npl == 2) {
        if (p.usequad) soft0 ? launch<2, true, true>(ctx, a, grid, crit) : launch<2, true, false>(ctx, a, grid, crit);
        else soft0 ? launch<2, false, true>(ctx, a, grid, crit) : launch<2, false, false>(ctx, a, grid, crit);
    } else {
        if (p.usequad) soft0 ? launch<1, true, true>(ctx, a, grid, crit) : launch<1, true, false>(ctx, a, grid, crit);
        else soft0 ? launch<1, false, true>(ctx, a, grid, crit) : launch<1, false, false>(ctx, a, grid, crit);
    }
    BH_HIP(hipGetLastError());
    if (!ctx->capturing) BH_HIP(hipEventRecord(ctx->ev[5], ctx->stream));
}

}  // namespace bh
