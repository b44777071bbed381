// Tree construction on the device: root box, 63-bit Morton keys, LSD radix
// sort, octree topology from the sorted keys, and the bottom-up FP64
// monopole/quadrupole pass.
//
// Semantics follow hackcode1's maketree (src/nbody/evolve/hackcode/hackcode1/
// load.c:32-62): only massive bodies are loaded (56-60); the box starts from
// the caller's rmin/rsize and doubles towards the first out-of-box body in
// input order (expandbox, 68-95); a body's integer coordinate is
// floor(2^30 (x - rmin)/rsize) on the half-open unit interval (intcoord,
// 135-156); the child index is 4x+2y+z (subindex, 162-177); a cell exists for
// every key prefix shared by two or more bodies (loadtree splits a leaf on
// collision, 101-128); c.o.m. and the traceless quadrupole are accumulated
// over the children in octant order (hackcofm, 183-226).
//
// Layout: the tree is described by the sorted key array.  cpl[s] is the
// number of leading octal digits shared by keys s and s+1; the cells that
// start at body s are the levels cpl[s-1]+1 .. cpl[s]; numbering them with an
// exclusive scan gives cells in DFS pre-order.  Depth is capped at 21 levels
// (63-bit keys); a level-21 cell holding several bodies is a bucket whose
// bodies are its children.
#include "bh_ctx.h"

#include <algorithm>

namespace bh {
namespace {

constexpr int kTpb = 256;

__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v)
{
    const int lane = lane_id();
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t t = __shfl_up(v, d, 64);
        if (lane >= d) v += t;
    }
    return v;
}

// Exclusive scan across a 256-thread block; *total receives the block sum.
__device__ uint32_t block_excl_scan(uint32_t v, uint32_t *total)
{
    __shared__ uint32_t ws[kTpb / 64];
    const int w = threadIdx.x >> 6;
    uint32_t incl = wave_incl_scan(v);
    if (lane_id() == 63) ws[w] = incl;
    __syncthreads();
    uint32_t off = 0, tot = 0;
#pragma unroll
    for (int k = 0; k < kTpb / 64; k++) {
        uint32_t x = ws[k];
        if (k < w) off += x;
        tot += x;
    }
    __syncthreads();
    *total = tot;
    return off + incl - v;
}

// ---- box -----------------------------------------------------------------

__device__ __forceinline__ bool fits(const double *x, const double4 b)
{
    double r0 = (x[0] - b.x) / b.w;
    double r1 = (x[1] - b.y) / b.w;
    double r2 = (x[2] - b.z) / b.w;
    return (0.0 <= r0 && r0 < 1.0) && (0.0 <= r1 && r1 < 1.0) && (0.0 <= r2 && r2 < 1.0);
}

// The same test for many bodies against one box.  When the side is a power
// of two, d * (1/side) is the correctly rounded d / side (both round the same
// real number), so the three divisions become multiplications with identical
// results; any other side divides.
struct BoxTest {
    double4 b;
    double inv;
    bool p2;
};

__device__ __forceinline__ BoxTest box_test(const double4 b)
{
    const unsigned long long bits = (unsigned long long)__double_as_longlong(b.w);
    const unsigned e = (unsigned)((bits >> 52) & 0x7ff);
    const bool p2 = (bits & 0x000FFFFFFFFFFFFFULL) == 0 && e - 1u < 0x7feu;
    return {b, p2 ? 1.0 / b.w : 0.0, p2};
}

__device__ __forceinline__ bool fits(const double *x, const BoxTest &t)
{
    if (!t.p2) return fits(x, t.b);
    const double r0 = (x[0] - t.b.x) * t.inv, r1 = (x[1] - t.b.y) * t.inv, r2 = (x[2] - t.b.z) * t.inv;
    return (0.0 <= r0 && r0 < 1.0) && (0.0 <= r1 && r1 < 1.0) && (0.0 <= r2 && r2 < 1.0);
}

__device__ __forceinline__ bool fits(const double4 c, const BoxTest &t)
{
    const double x[3] = {c.x, c.y, c.z};
    return fits(x, t);
}

// Massive bodies outside the starting box (the only ones that can ever make
// expandbox grow it, since the box only grows), compacted in input order:
// per-block counts, a scan, then per-block writes.  Block b owns the input
// range [b*chunk, (b+1)*chunk).
__device__ __forceinline__ bool outside_massive(const double *__restrict__ pos, const double *__restrict__ mass,
                                                int64_t i, int64_t n, const BoxTest &b, int *bad)
{
    if (i >= n) return false;
    // mass and position loads issued together (one memory round trip)
    const double m = mass[i];
    const double x[3] = {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
    if (m == 0.0 || fits(x, b)) return false;
    if (!(isfinite(x[0]) && isfinite(x[1]) && isfinite(x[2]))) *bad = 1;
    return true;
}

// Compaction in any order (the rounds take the lowest index).  Block b owns
// bodies [b*kBoxChunk, (b+1)*kBoxChunk); in step k thread t tests body
// b*kBoxChunk + k*kTpb + t (coalesced) into bit k of a mask, a block scan of
// the counts gives the slots and one atomic per block (boxi[4]) reserves the
// block's range.  A candidate is stored with its position (x, y, z, index),
// so the rounds read it contiguously.
constexpr int kBoxPer = 16, kBoxChunk = kTpb * kBoxPer;

__global__ __launch_bounds__(kTpb) void k_box_compact(const double *__restrict__ pos, const double *__restrict__ mass,
                                                      int64_t n, const double4 *__restrict__ box, double4 *cand,
                                                      int *boxi, unsigned long long *stat)
{
    __shared__ uint32_t base;
    const BoxTest b = box_test(*box);
    const int64_t i0 = (int64_t)blockIdx.x * kBoxChunk + threadIdx.x;
    int bad = 0;
    uint32_t mask = 0;
#pragma unroll
    for (int k = 0; k < kBoxPer; k++)
        if (outside_massive(pos, mass, i0 + k * kTpb, n, b, &bad)) mask |= 1u << k;
    uint32_t tot;
    uint32_t off = block_excl_scan((uint32_t)__popc(mask), &tot);
    if (threadIdx.x == 0 && tot) base = atomicAdd(reinterpret_cast<uint32_t *>(&boxi[4]), tot);
    __syncthreads();
    while (mask) {
        const int k = __ffs(mask) - 1;
        mask &= mask - 1;
        const int64_t i = i0 + k * kTpb;
        const double *x = pos + 3 * i;
        cand[base + off++] = make_double4(x[0], x[1], x[2], (double)i);
    }
    if (bad) atomicOr(&stat[S_FLAGS], (unsigned long long)F_NONFINITE);
}

// Doubling of expandbox (load.c:68-95) until body x fits.
__device__ __forceinline__ double4 expand_for(const double *x, double4 b)
{
    for (int guard = 0; guard < 2000 && !fits(x, b); guard++) {
        double rmid0 = b.x + 0.5 * b.w, rmid1 = b.y + 0.5 * b.w, rmid2 = b.z + 0.5 * b.w;
        if (x[0] < rmid0) b.x -= b.w;
        if (x[1] < rmid1) b.y -= b.w;
        if (x[2] < rmid2) b.z -= b.w;
        b.w = 2.0 * b.w;
    }
    return b;
}

// One expansion round: every block reduces the lowest candidate index >=
// start that lies outside the box (atomicMin into boxi[0]); the last block
// to finish (ticket in boxi[3]) expands the box toward that body and moves
// start past it, or, when no candidate is outside, sets boxi[5] (done), after
// which the remaining rounds return at once.
__global__ __launch_bounds__(kTpb) void k_box_round(const double *__restrict__ pos, const double4 *__restrict__ cand,
                                                    double4 *box, int *boxi)
{
    __shared__ int wbest[kTpb / 64];
    if (boxi[5]) return;
    const double4 b0 = *box;
    const BoxTest b = box_test(b0);
    const int start = boxi[1], nc = boxi[4];
    int best = INT32_MAX;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nc; k += gridDim.x * blockDim.x) {
        const double4 c = cand[k];
        const int i = (int)c.w;
        if (i >= start && i < best && !fits(c, b)) best = i;
    }
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) best = min(best, __shfl_xor(best, d, 64));
    if (lane_id() == 0) wbest[threadIdx.x >> 6] = best;
    __syncthreads();
    if (threadIdx.x == 0) {
        int m = wbest[0];
#pragma unroll
        for (int k = 1; k < kTpb / 64; k++) m = min(m, wbest[k]);
        if (m != INT32_MAX) atomicMin(&boxi[0], m);
        __threadfence();
        const unsigned t = atomicAdd(reinterpret_cast<unsigned *>(&boxi[3]), 1u);
        if (t == gridDim.x - 1) {
            __threadfence();
            const int first = atomicMin(&boxi[0], INT32_MAX);   // atomic read
            boxi[3] = 0;
            if (first != INT32_MAX) {
                *box = expand_for(pos + 3 * (int64_t)first, b0);
                boxi[2] += 1;
                boxi[0] = INT32_MAX;
                boxi[1] = first + 1;
            } else {
                boxi[5] = 1;
            }
        }
    }
}

// Starting box: the caller's (rmin, rsize), or the previous box when keep.
// boxi: [0] first outside candidate (min), [1] start index, [2] expansions,
// [3] round ticket, [4] candidate count, [5] done.
// Start of a build: clear the per-evaluation status words and, when box is
// given, set the start box and the expansion bookkeeping (one launch).
__global__ void k_build_start(unsigned long long *stat, int nzero, uint32_t *lwork, int nlwork, double4 *box,
                              int *boxi, double rx, double ry, double rz, double rsize, int keep)
{
    if ((int)threadIdx.x < nzero) stat[threadIdx.x] = 0;
    if (threadIdx.x == 0) stat[S_GCTR] = 0;
    if ((int)threadIdx.x < nlwork) lwork[threadIdx.x] = 0;
    if (threadIdx.x != 0 || !box) return;
    if (!keep) *box = make_double4(rx, ry, rz, rsize);
    boxi[0] = INT32_MAX;
    boxi[1] = 0;
    boxi[2] = 0;
    boxi[3] = 0;
    boxi[4] = 0;
    boxi[5] = 0;
}

// ---- batched ensembles: one box per system -----------------------------------

// expandbox for every system of a batch (one block each): from the caller's
// box, or the system's previous box, take the first body in input order that
// lies outside, double the box towards it (load.c:68-95), and go on from the
// next body; every body of a batch is massive.  All threads compute the same
// box.  A round scans forward from its start in chunks of kTpb * kSysBoxPer
// bodies (all loads of a chunk in flight) and stops at the first chunk that
// holds a body outside; the next round starts after that body, so all rounds
// together read each body about once.  At most `rounds` expansions; a box
// still open is flagged.
constexpr int kSysBoxPer = 8;

__global__ __launch_bounds__(kTpb) void k_sys_box(const double *__restrict__ pos, const int64_t *__restrict__ off,
                                                  double4 *sysbox, double4 *box0, double rx, double ry, double rz,
                                                  double rsize, int keep, int rounds, unsigned long long *stat)
{
    __shared__ int red[kTpb / 64];
    const int s = blockIdx.x;
    const int lo = (int)off[s], hi = (int)off[s + 1];
    double4 b = keep ? sysbox[s] : make_double4(rx, ry, rz, rsize);
    int start = lo;
    bool closed = false;
    for (int round = 0; round <= rounds; round++) {
        const BoxTest bt = box_test(b);
        int f = hi;
        for (int c0 = start; c0 < hi && f == hi; c0 += kTpb * kSysBoxPer) {
            int first = hi;
#pragma unroll
            for (int k = 0; k < kSysBoxPer; k++) {
                const int i = c0 + k * kTpb + (int)threadIdx.x;
                if (i < hi && !fits(pos + 3 * (int64_t)i, bt)) first = min(first, i);
            }
#pragma unroll
            for (int d = 32; d >= 1; d >>= 1) first = min(first, __shfl_xor(first, d, 64));
            if (lane_id() == 0) red[threadIdx.x >> 6] = first;
            __syncthreads();
            f = red[0];
#pragma unroll
            for (int k = 1; k < kTpb / 64; k++) f = min(f, red[k]);
            __syncthreads();
        }
        if (f == hi) {
            closed = true;
            break;
        }
        if (round == rounds) break;
        const double *x = pos + 3 * (int64_t)f;
        if (!(isfinite(x[0]) && isfinite(x[1]) && isfinite(x[2]))) {
            if (threadIdx.x == 0) atomicOr(&stat[S_FLAGS], (unsigned long long)F_NONFINITE);
            closed = true;
            break;
        }
        for (int guard = 0; guard < 2000 && !fits(x, b); guard++) {
            double rmid0 = b.x + 0.5 * b.w, rmid1 = b.y + 0.5 * b.w, rmid2 = b.z + 0.5 * b.w;
            if (x[0] < rmid0) b.x -= b.w;
            if (x[1] < rmid1) b.y -= b.w;
            if (x[2] < rmid2) b.z -= b.w;
            b.w = 2.0 * b.w;
        }
        start = f + 1;
    }
    if (threadIdx.x == 0) {
        sysbox[s] = b;
        if (s == 0) *box0 = b;
        if (!closed) atomicOr(&stat[S_FLAGS], (unsigned long long)F_BOX_OPEN);
    }
}

// ---- keys ------------------------------------------------------------------

__device__ __forceinline__ uint64_t spread3(uint64_t v)
{
    v &= 0x1fffffULL;
    v = (v | (v << 32)) & 0x1f00000000ffffULL;
    v = (v | (v << 16)) & 0x1f0000ff0000ffULL;
    v = (v | (v << 8)) & 0x100f00f00f00f00fULL;
    v = (v | (v << 4)) & 0x10c30c30c30c30c3ULL;
    v = (v | (v << 2)) & 0x1249249249249249ULL;
    return v;
}

// Interleaved top `levels` bits of the three 30-bit integer coordinates.
__device__ __forceinline__ uint64_t morton_of(const double *x, const double4 b, int levels = kMaxLevel)
{
    uint64_t c[3];
    const double rm[3] = {b.x, b.y, b.z};
#pragma unroll
    for (int k = 0; k < 3; k++) {
        double xsc = (x[k] - rm[k]) / b.w;
        int64_t xi;
        if (!(xsc >= 0.0)) xi = 0;
        else if (!(xsc < 1.0)) xi = (1 << 30) - 1;
        else xi = (int64_t)floor(1073741824.0 * xsc);
        c[k] = (uint64_t)(xi >> (30 - levels));
    }
    return (spread3(c[0]) << 2) | (spread3(c[1]) << 1) | spread3(c[2]);
}

// Batch keys: [system id : sysbits][Morton code of kMaxLevel - sysbits/3
// levels in the system's own box]; the system digits form the top levels.
__global__ void k_keys_batch(const double *__restrict__ pos, int64_t n, const int64_t *__restrict__ off, int nsys,
                             const double4 *__restrict__ sysbox, int sysbits, uint64_t *keys, uint32_t *vals)
{
    const int levels = kMaxLevel - sysbits / 3;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        int lo = 0, hi = nsys;          // last system with off[s] <= i
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (off[mid] <= i) lo = mid;
            else hi = mid;
        }
        const uint64_t m = morton_of(pos + 3 * i, sysbox[lo], levels);
        keys[i] = ((uint64_t)lo << (3 * levels)) | m;
        vals[i] = (uint32_t)i;
    }
}

// Mass set: massless bodies get bit 63 so they sort after every massive body.
// pk (optional): the body packed as (x, y, z, m) in input order, so that the
// gather into sorted order reads one aligned 32-byte record per body.
// boxi (mass-set keys after the expansion rounds, else nullptr): the rounds'
// done flag; a box that no round found closed (every candidate inside) is
// flagged, and the host grows the number of rounds
__global__ void k_keys(const double *__restrict__ pos, const double *__restrict__ mass, int64_t n,
                       const double4 *__restrict__ box, uint64_t *keys, uint32_t *vals, double4 *pk,
                       const int *__restrict__ boxi = nullptr, unsigned long long *stat = nullptr)
{
    if (boxi && blockIdx.x == 0 && threadIdx.x == 0 && !boxi[5]) atomicOr(&stat[S_FLAGS], (unsigned long long)F_BOX_OPEN);
    const double4 b = *box;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double *x = pos + 3 * i;
        uint64_t k = morton_of(x, b);
        const double m = mass != nullptr ? mass[i] : 0.0;
        if (mass != nullptr && m == 0.0) k |= (1ULL << 63);
        keys[i] = k;
        vals[i] = (uint32_t)i;
        if (pk) pk[i] = make_double4(x[0], x[1], x[2], m);
    }
}

// ---- exclusive scan (uint32) -------------------------------------------------

constexpr int kScanItems = 16;
constexpr int kScanChunk = kTpb * kScanItems;

// n: entries known to the host; dn (optional): count on the device, which
// limits the scan to min(n, *dn) entries (the rest of out is not written).
__device__ __forceinline__ int64_t scan_len(int64_t n, const unsigned long long *dn)
{
    return dn ? min(n, (int64_t)*dn) : n;
}

// Sum of each chunk; a thread adds its kScanItems consecutive entries (16-byte
// vectors when aligned and in range), as k_scan_down reads them.
__global__ void k_scan_reduce(const uint32_t *__restrict__ in, int64_t n, const unsigned long long *dn,
                              uint32_t *partial)
{
    n = scan_len(n, dn);
    const int64_t mine = (int64_t)blockIdx.x * kScanChunk + (int64_t)threadIdx.x * kScanItems;
    uint32_t s = 0;
    if ((reinterpret_cast<uintptr_t>(in) & 15) == 0 && mine + kScanItems <= n) {
#pragma unroll
        for (int r = 0; r < kScanItems / 4; r++) {
            const uint4 q = reinterpret_cast<const uint4 *>(in + mine)[r];
            s += (q.x + q.y) + (q.z + q.w);
        }
    } else {
#pragma unroll
        for (int r = 0; r < kScanItems; r++)
            if (mine + r < n) s += in[mine + r];
    }
    uint32_t tot;
    block_excl_scan(s, &tot);
    if (threadIdx.x == 0) partial[blockIdx.x] = tot;
}

// Scans the block sums; the total goes to partial[nparts] and out[n].
__global__ void k_scan_top(uint32_t *partial, int64_t nparts, uint32_t *out, int64_t n)
{
    uint32_t carry = 0;
    for (int64_t base = 0; base < nparts; base += kTpb) {
        int64_t i = base + threadIdx.x;
        uint32_t v = i < nparts ? partial[i] : 0;
        uint32_t tot;
        uint32_t ex = block_excl_scan(v, &tot);
        if (i < nparts) partial[i] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) {
        partial[nparts] = carry;
        out[n] = carry;
    }
}

// Each thread owns kScanItems consecutive entries; when in and out are
// 16-byte aligned, a thread's full run moves as 16-byte vectors.
__global__ void k_scan_down(const uint32_t *in, uint32_t *out, int64_t n, const unsigned long long *dn,
                            const uint32_t *__restrict__ partial)
{
    n = scan_len(n, dn);
    const int64_t base = (int64_t)blockIdx.x * kScanChunk;
    if (base >= n) return;
    const int64_t mine = base + (int64_t)threadIdx.x * kScanItems;
    const bool vec = ((reinterpret_cast<uintptr_t>(in) | reinterpret_cast<uintptr_t>(out)) & 15) == 0;
    const bool full = vec && mine + kScanItems <= n;
    uint32_t v[kScanItems];
    uint32_t s = 0;
    if (full) {
#pragma unroll
        for (int r = 0; r < kScanItems / 4; r++) {
            const uint4 q = reinterpret_cast<const uint4 *>(in + mine)[r];
            v[4 * r] = q.x;
            v[4 * r + 1] = q.y;
            v[4 * r + 2] = q.z;
            v[4 * r + 3] = q.w;
        }
    } else {
#pragma unroll
        for (int r = 0; r < kScanItems; r++) v[r] = mine + r < n ? in[mine + r] : 0;
    }
#pragma unroll
    for (int r = 0; r < kScanItems; r++) s += v[r];
    uint32_t tot;
    uint32_t ex = block_excl_scan(s, &tot) + partial[blockIdx.x];
    uint32_t o[kScanItems];
#pragma unroll
    for (int r = 0; r < kScanItems; r++) {
        o[r] = ex;
        ex += v[r];
    }
    if (full) {
#pragma unroll
        for (int r = 0; r < kScanItems / 4; r++)
            reinterpret_cast<uint4 *>(out + mine)[r] = make_uint4(o[4 * r], o[4 * r + 1], o[4 * r + 2], o[4 * r + 3]);
    } else {
#pragma unroll
        for (int r = 0; r < kScanItems; r++)
            if (mine + r < n) out[mine + r] = o[r];
    }
}

// ---- radix sort (LSD, 8-bit digits, stable) -----------------------------------

constexpr int kSortItems = 8;
static_assert(kTpb == 256, "the radix kernels use one thread per 8-bit digit");
constexpr int kSortTile = kTpb * kSortItems;

// One sub-histogram per wave (skewed digits, e.g. the top byte of clustered
// keys, would otherwise serialise the block on a few LDS counters).
__global__ __launch_bounds__(kTpb) void k_radix_hist(const uint64_t *__restrict__ keys, int64_t n, int shift,
                                                     uint32_t *hist, int nblk)
{
    __shared__ uint32_t h[kTpb / 64][256];
#pragma unroll
    for (int w = 0; w < kTpb / 64; w++) h[w][threadIdx.x] = 0;
    __syncthreads();
    uint32_t *hw = h[threadIdx.x >> 6];
    const int64_t base = (int64_t)blockIdx.x * kSortTile;
#pragma unroll
    for (int r = 0; r < kSortItems; r++) {
        int64_t i = base + (int64_t)r * kTpb + threadIdx.x;
        if (i < n) atomicAdd(&hw[(keys[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    uint32_t c = 0;
#pragma unroll
    for (int w = 0; w < kTpb / 64; w++) c += h[w][threadIdx.x];
    hist[(int64_t)threadIdx.x * nblk + blockIdx.x] = c;
}

// Exclusive scan of each digit's row of the digit-major block histogram
// (block d scans hist[d*nblk .. d*nblk+nblk)); the row total goes to dtot[d].
// The scatter adds the exclusive scan of dtot (the digit's first output slot).
__global__ __launch_bounds__(kTpb) void k_hist_rows(const uint32_t *__restrict__ hist, int nblk, uint32_t *hscan,
                                                    uint32_t *dtot)
{
    const int64_t row = (int64_t)blockIdx.x * nblk;
    uint32_t run = 0;
    for (int b0 = 0; b0 < nblk; b0 += kTpb * 4) {
        const int i0 = b0 + threadIdx.x * 4;
        uint32_t v[4];
#pragma unroll
        for (int j = 0; j < 4; j++) v[j] = i0 + j < nblk ? hist[row + i0 + j] : 0u;
        uint32_t tot;
        uint32_t acc = run + block_excl_scan(v[0] + v[1] + v[2] + v[3], &tot);
#pragma unroll
        for (int j = 0; j < 4; j++) {
            if (i0 + j < nblk) hscan[row + i0 + j] = acc;
            acc += v[j];
        }
        run += tot;
    }
    if (threadIdx.x == 0) dtot[blockIdx.x] = run;
}

// Scatter one tile of kSortTile keys.  The tile is first ranked into LDS in
// digit order: within a round of 256 consecutive keys, ranks among equal
// digits come from a wave-level match (8 ballots) and a per-digit prefix over
// the block's 4 waves, so equal digits keep their input order (round, wave,
// lane).  Then consecutive threads copy consecutive LDS slots to the output,
// so the writes of one digit's run are contiguous.
__global__ __launch_bounds__(kTpb) void k_radix_scatter(const uint64_t *__restrict__ kin, const uint32_t *__restrict__ vin,
                                                        uint64_t *__restrict__ kout, uint32_t *__restrict__ vout,
                                                        int64_t n, int shift, const uint32_t *__restrict__ hist,
                                                        const uint32_t *__restrict__ hist_scan,
                                                        const uint32_t *__restrict__ dtot, int nblk)
{
    __shared__ uint64_t skey[kSortTile];
    __shared__ uint32_t sval[kSortTile];
    __shared__ uint32_t toff[256];         // first LDS slot of each digit in this tile
    __shared__ uint32_t gofs[256];         // output position of that slot
    __shared__ uint32_t run[256];          // next free LDS slot per digit
    __shared__ uint32_t wcnt[kTpb / 64][256];
    const int t = threadIdx.x, w = t >> 6, lane = lane_id();
    {
        uint32_t tot;
        const uint32_t c = hist[(int64_t)t * nblk + blockIdx.x];
        const uint32_t o = block_excl_scan(c, &tot);
        const uint32_t dbase = block_excl_scan(dtot[t], &tot);
        toff[t] = o;
        run[t] = o;
        gofs[t] = dbase + hist_scan[(int64_t)t * nblk + blockIdx.x] - o;
    }
    const int64_t base = (int64_t)blockIdx.x * kSortTile;
    const int cnt = (int)min<int64_t>(kSortTile, n - base);
    const uint64_t lt = (lane == 0) ? 0ULL : (~0ULL >> (64 - lane));
    for (int r = 0; r < kSortItems; r++) {
#pragma unroll
        for (int k = 0; k < kTpb / 64; k++) wcnt[k][t] = 0;
        __syncthreads();
        const int li = r * kTpb + t;
        const bool ok = li < cnt;
        const uint64_t key = ok ? kin[base + li] : 0;
        const uint32_t val = ok ? vin[base + li] : 0;
        const uint32_t d = ok ? (uint32_t)((key >> shift) & 255u) : 0u;
        uint64_t peers = __ballot(ok);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const uint64_t bal = __ballot((d >> b) & 1u);
            peers &= ((d >> b) & 1u) ? bal : ~bal;
        }
        const uint32_t rank = __popcll(peers & lt);
        if (ok && rank == 0) wcnt[w][d] = __popcll(peers);
        __syncthreads();
        {
            uint32_t acc = run[t];
#pragma unroll
            for (int k = 0; k < kTpb / 64; k++) {
                const uint32_t c = wcnt[k][t];
                wcnt[k][t] = acc;
                acc += c;
            }
            run[t] = acc;
        }
        __syncthreads();
        if (ok) {
            const uint32_t slot = wcnt[w][d] + rank;
            skey[slot] = key;
            sval[slot] = val;
        }
        __syncthreads();   // wcnt is cleared at the top of the next round
    }
    for (int j = t; j < cnt; j += kTpb) {
        const uint64_t key = skey[j];
        const uint32_t dst = gofs[(key >> shift) & 255u] + (uint32_t)j;
        kout[dst] = key;
        vout[dst] = sval[j];
    }
}

// ---- gather sorted bodies ----------------------------------------------------

__device__ __forceinline__ void split2(double x, float &hi, float &lo)
{
    hi = (float)x;
    lo = (float)(x - (double)hi);
}

// Test set in its sorted order.  tmap (input index -> tree body or -1, NULL
// when no test point is a tree body) gives tself[s], the tree body of test
// point s, and tinv, the test index of each tree body.
__global__ void k_gather_test(const uint32_t *__restrict__ perm, int64_t nt, const double *__restrict__ pos,
                              float4 *thi, float4 *tlo, double4 *td, int *tself, const int *__restrict__ tmap,
                              int *tinv)
{
    int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (s >= nt) return;
    int64_t o = perm[s];
    double x = pos[3 * o], y = pos[3 * o + 1], z = pos[3 * o + 2];
    float4 h, l;
    split2(x, h.x, l.x);
    split2(y, h.y, l.y);
    split2(z, h.z, l.z);
    h.w = 0.f;
    l.w = 0.f;
    thi[s] = h;
    tlo[s] = l;
    td[s] = make_double4(x, y, z, 0.0);
    const int b = tmap ? tmap[o] : -1;
    tself[s] = b;
    if (b >= 0) tinv[b] = (int)s;
}

// Input index -> tree body (sorted index among the massive bodies) or -1 for
// a massless body (which sorts after every massive one).
__global__ void k_tree_index(const uint32_t *__restrict__ vs, int64_t n, int nb, int *tmap)
{
    int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (s < n) tmap[vs[s]] = s < nb ? (int)s : -1;
}

// ---- topology ----------------------------------------------------------------

// cpl and its copy cq = cpl + 1 (0..22, for the 8-entry range scans).
__global__ void k_cpl(const uint64_t *__restrict__ keys, int nb, int8_t *cpl, uint8_t *cq, uint32_t *ccnt)
{
    for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < nb; s += gridDim.x * blockDim.x) {
        int c = -1, cp = -1;
        if (s + 1 < nb) c = min(kMaxLevel, (__clzll(keys[s] ^ keys[s + 1]) - 1) / 3);
        if (s > 0) cp = min(kMaxLevel, (__clzll(keys[s - 1] ^ keys[s]) - 1) / 3);
        cpl[s] = (int8_t)c;
        cq[s] = (uint8_t)(c + 1);
        ccnt[s] = (uint32_t)max(0, c - cp);
    }
}

// Cells needed = cbase[n]; the count used is clamped to the capacity and an
// overflow is flagged (the host then grows the arrays and evaluates again).
struct CellClamp {
    const uint32_t *cbase;
    int n;
    unsigned long long cap;
    unsigned long long *stat;
    int slot, slot_raw;
    unsigned long long flag;
};

__device__ __forceinline__ void clamp_cells(const CellClamp &c)
{
    const unsigned long long raw = c.n >= 2 ? c.cbase[c.n] : 0;
    c.stat[c.slot_raw] = raw;
    c.stat[c.slot] = raw < c.cap ? raw : c.cap;
    if (raw > c.cap) atomicOr(&c.stat[S_FLAGS], c.flag);
}

// Block minima over 64 entries of a byte array in the cq domain (cpl + 1,
// values 0..22): level 1 over cq, levels 2 and 3 over the level below, for
// the range-end search.
// out[b] = min of in[64 b .. 64 b + 63] (entries past nin ignored): eight
// lanes per output, one 8-byte load each (in is 8-byte aligned), then a min
// across the eight.  Launch with 8 * nout threads.  cc (first pyramid level
// only): thread 0 also sets the cell count.
__global__ void k_cmin(const uint8_t *__restrict__ in, int nin, uint8_t *out, int nout, CellClamp cc = CellClamp{})
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (cc.stat && t == 0) clamp_cells(cc);
    const int b = t >> 3, lo = b * 64 + 8 * (t & 7);
    int m = 255;
    if (b < nout) {
        if (lo + 8 <= nin) {
            const uint64_t w = *reinterpret_cast<const uint64_t *>(in + lo);
#pragma unroll
            for (int j = 0; j < 8; j++) m = min(m, (int)((w >> (8 * j)) & 255));
        } else {
            for (int j = lo; j < nin; j++) m = min(m, (int)in[j]);
        }
    }
    m = min(m, __shfl_xor(m, 1, 64));
    m = min(m, __shfl_xor(m, 2, 64));
    m = min(m, __shfl_xor(m, 4, 64));
    if (b < nout && (t & 7) == 0) out[b] = (uint8_t)m;
}

// Flags (the high bit of a byte) of the entries k, k+1, .. of the aligned
// 8-entry word of a that holds entry k whose value is below n (per byte:
// n = L + 1 <= 22).  A byte below n is flagged by (x - n*0x01..01) & ~x &
// 0x80..80 exactly for the lowest such byte (borrows only reach higher
// bytes; bytes of 128 and above never flag); the entries before k are set
// to 0x7f so that they never match.
__device__ __forceinline__ uint64_t below(const uint8_t *__restrict__ a, int k, uint64_t nl)
{
    const int w = k & ~7;
    uint64_t x = *reinterpret_cast<const uint64_t *>(a + w);
    const int sh = (k - w) * 8;
    if (sh) x |= 0x7f7f7f7f7f7f7f7fULL & ((1ULL << sh) - 1);
    return (x - nl) & ~x & 0x8080808080808080ULL;
}

// First j >= s with cpl[j] < L (exists: cpl[nb-1] = -1), i.e. cq[j] < L + 1.
// Each level of the minima pyramid (m1: 64 bodies, m2: 4096, m3: 262144 per
// entry; all in the cq domain, 8-byte aligned, one word of padding) is read
// eight entries at a time: at an aligned position the word either lets the
// search skip up to eight blocks whose minimum is >= L, or names the first
// block holding an entry below L, which the next finer level then searches.
// At most ~8 words per level going up and down.
__device__ int range_end(const uint8_t *__restrict__ cq, const uint8_t *__restrict__ m1,
                         const uint8_t *__restrict__ m2, const uint8_t *__restrict__ m3, int s, int L)
{
    const uint64_t nl = 0x0101010101010101ULL * (uint64_t)(L + 1);
    int j = s;
    for (;;) {
        if ((j & 262143) == 0) {
            const int k = j >> 18;
            const uint64_t t = below(m3, k, nl);
            if (!t) { j = ((k | 7) + 1) << 18; continue; }
            j = ((k & ~7) + (__builtin_ctzll(t) >> 3)) << 18;
        }
        if ((j & 4095) == 0) {
            const int k = j >> 12;
            const uint64_t t = below(m2, k, nl);
            if (!t) { j = ((k | 7) + 1) << 12; continue; }
            j = ((k & ~7) + (__builtin_ctzll(t) >> 3)) << 12;
        }
        if ((j & 63) == 0) {
            const int k = j >> 6;
            const uint64_t t = below(m1, k, nl);
            if (!t) { j = ((k | 7) + 1) << 6; continue; }
            j = ((k & ~7) + (__builtin_ctzll(t) >> 3)) << 6;
        }
        const uint64_t t = below(cq, j, nl);
        if (t) return (j & ~7) + (__builtin_ctzll(t) >> 3);
        j = (j | 7) + 1;
    }
}

// Owner body of every cell: body s starts the cells cbase[s] .. cbase[s+1]-1.
__global__ void k_owner(const uint32_t *__restrict__ cbase, int nb, const unsigned long long *ncellp, int *owner)
{
    const int ncell = (int)*ncellp;
    for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < nb; s += gridDim.x * blockDim.x) {
        const int id1 = min((int)cbase[s + 1], ncell);
        for (int id = (int)cbase[s]; id < id1; id++) owner[id] = s;
    }
}

// One thread per cell (grid-stride over the device cell count).  The cells
// that start at body s are the levels cpl[s-1]+1 .. cpl[s] with ids from
// cbase[s]; the range of a cell of level L ends before the first j >= s with
// cpl[j] < L.
__global__ void k_cells(const int8_t *__restrict__ cpl, const uint8_t *__restrict__ cq, const uint32_t *__restrict__ cbase,
                        const int *__restrict__ owner,
                        const uint8_t *__restrict__ m1, const uint8_t *__restrict__ m2, const uint8_t *__restrict__ m3,
                        const unsigned long long *ncellp, int *cfirst, int *clast, uint8_t *clevel, int *cskip,
                        int *pre, int2 *fnb, unsigned long long *deep)
{
    const int ncell = (int)*ncellp;
    const int stride = gridDim.x * blockDim.x;
    for (int base = blockIdx.x * blockDim.x; base < ncell; base += stride) {
        const int id = base + threadIdx.x;
        bool isdeep = false;
        if (id < ncell) {
            const int s = owner[id];
            const int l0 = (s > 0 ? cpl[s - 1] : -1) + 1;
            const int L = l0 + (id - (int)cbase[s]);
            const int e = range_end(cq, m1, m2, m3, s, L);
            cfirst[id] = s;
            clast[id] = e;
            clevel[id] = (uint8_t)L;
            if (cskip) cskip[id] = (e + 1) + (int)cbase[e + 1];
            if (fnb) fnb[id] = make_int2(s, e - s + 1);
            if (pre) pre[s + id] = ~id;
            isdeep = L == kMaxLevel;
        }
        const unsigned long long m = __ballot(isdeep);
        if (deep && m && lane_id() == 0) atomicAdd(deep, (unsigned long long)__popcll(m));
    }
}

// Pre-order position of body s: after the cells that start at s.
__global__ void k_body_pre(const uint32_t *__restrict__ cbase, int nb, const unsigned long long *ncellp, int *pre)
{
    const uint32_t ncell = (uint32_t)*ncellp;
    for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < nb; s += gridDim.x * blockDim.x)
        if (cbase[s + 1] <= ncell) pre[s + cbase[s + 1]] = s;
}

__global__ void k_set_int(int *p, int v) { *p = v; }

__global__ void k_copy_count(const uint32_t *src, unsigned long long *dst) { *dst = *src; }

// Fixed-size groups of consecutive points (sets too small to need a topology).
__global__ void k_groups_trivial(int64_t n, int G, int *gfirst, int *gcount, unsigned long long *stat)
{
    int64_t ng = (n + G - 1) / G;
    for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < ng; g += (int64_t)gridDim.x * blockDim.x) {
        gfirst[g] = (int)(g * G);
        gcount[g] = (int)min<int64_t>(G, n - g * G);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) stat[S_NGROUPS] = (unsigned long long)ng;
}

// ---- groups of test points -------------------------------------------------
// A cell holding more than G points emits groups made of runs of consecutive
// children that hold at most G points each (siblings are adjacent in space, so
// a group's extent is at most its parent cell); children with more than G
// points emit their own groups.  A level-21 bucket with more than G points is
// cut into chunks of G.  Every point lands in exactly one group.
// Octants inside the axis-aligned box spanned by the octants of mask m.
__device__ __forceinline__ unsigned octant_box(unsigned m)
{
    unsigned lo[3] = {1, 1, 1}, hi[3] = {0, 0, 0};
    for (int q = 0; q < 8; q++)
        if (m & (1u << q))
            for (int a = 0; a < 3; a++) {
                unsigned b = (q >> (2 - a)) & 1u;
                lo[a] = min(lo[a], b);
                hi[a] = max(hi[a], b);
            }
    unsigned box = 0;
    for (int q = 0; q < 8; q++) {
        bool in = true;
        for (int a = 0; a < 3; a++) {
            unsigned b = (q >> (2 - a)) & 1u;
            in = in && b >= lo[a] && b <= hi[a];
        }
        if (in) box |= 1u << q;
    }
    return box;
}

// End (exclusive) of the child of cell c (bodies s..e, level L) that starts
// at body t: a child cell's last body + 1, or t + 1 for a body; *cc gets the
// child cell or -1.  A child cell past the built cells (a capacity overflow,
// which is evaluated again) or an inconsistent range ends the run at e + 1,
// so every loop over children advances and stays in bounds.
__device__ __forceinline__ int child_end(int c, int s, int e, int t, int L, const int8_t *__restrict__ cpl,
                                         const uint32_t *__restrict__ cbase, const int *__restrict__ clast, int ncell,
                                         int *cc)
{
    *cc = -1;
    if (cpl[t] < L + 1) return t + 1;
    const int k = (t == s) ? c + 1 : (int)cbase[t];
    if (k >= ncell) return e + 1;
    *cc = k;
    const int tn = clast[k] + 1;
    return (tn > t && tn <= e + 1) ? tn : e + 1;
}

template <bool WRITE>
__device__ uint32_t pack_groups(int c, const uint64_t *__restrict__ keys, const int8_t *__restrict__ cpl,
                                const uint32_t *__restrict__ cbase, const int *__restrict__ cfirst,
                                const int *__restrict__ clast, const uint8_t *__restrict__ clevel, int G, int lsys,
                                int ncell, const int *__restrict__ csub, const int2 *__restrict__ fnb, int *gfirst,
                                int *gcount, uint32_t off)
{
    int s, e;
    if (fnb) {
        const int2 f = fnb[c];
        if (f.y <= G && lsys == 0) return 0;
        s = f.x;
        e = f.x + f.y - 1;
    } else {
        s = cfirst[c];
        e = clast[c];
    }
    const int L = clevel[c];
    uint32_t k = 0;
    auto emit = [&](int start, int count) {
        if (WRITE) {
            gfirst[off + k] = start;
            gcount[off + k] = count;
        }
        k++;
    };
    if (L < lsys) {
        // a level of system digits (batches): every child that is a whole
        // system (a system-root cell or a one-body system) with at most G
        // points is a group of its own, never merged with another system
        for (int t = s; t <= e;) {
            int cc;
            const int tn = child_end(c, s, e, t, L, cpl, cbase, clast, ncell, &cc);
            if (!(cc >= 0 && L + 1 < lsys) && tn - t <= G) emit(t, tn - t);
            t = tn;
        }
        return k;
    }
    if (e - s + 1 <= G) return 0;
    if (L == kMaxLevel) {
        for (int t = s; t <= e; t += G) emit(t, min(G, e - t + 1));
        return k;
    }
    const int shift = 60 - 3 * L;
    // runs of consecutive small children; a run also ends when the box spanned
    // by its octants would cover a big child's octant (a child holding more
    // than G points).  A child that does not fit is split: its first bodies
    // fill the run to G and the rest start the next run (every group stays
    // inside this cell, and groups are full).
    unsigned big = 0;
    int run = 0, rs = s;
    unsigned rmask = 0;
    auto visit = [&](int t, int size, int oct) {
        const unsigned o = 1u << oct;
        if (size > G) {
            if (run) emit(rs, run);
            run = 0;
            rmask = 0;
        } else if (run && (octant_box(rmask | o) & big)) {
            emit(rs, run);
            rs = t;
            run = size;
            rmask = o;
        } else {
            if (run == 0) rs = t;
            if (run + size > G) {
                emit(rs, G);
                rs += G;
                run = run + size - G;
                rmask = o;
            } else {
                run += size;
                rmask |= o;
            }
        }
    };
    if (csub) {
        // children from the tree's child slots: (first point, count) per octant
        const int4 s0 = reinterpret_cast<const int4 *>(csub)[2 * (int64_t)c];
        const int4 s1 = reinterpret_cast<const int4 *>(csub)[2 * (int64_t)c + 1];
        const int r[8] = {s0.x, s0.y, s0.z, s0.w, s1.x, s1.y, s1.z, s1.w};
        int2 f[8];
#pragma unroll
        for (int o = 0; o < 8; o++)
            f[o] = r[o] == kNoChild ? make_int2(0, 0) : (r[o] >= 0 ? make_int2(r[o], 1) : fnb[~r[o]]);
#pragma unroll
        for (int o = 0; o < 8; o++)
            if (f[o].y > G) big |= 1u << o;
#pragma unroll
        for (int o = 0; o < 8; o++)
            if (f[o].y > 0) visit(f[o].x, f[o].y, o);
    } else {
        for (int t = s; t <= e;) {
            int cc;
            const int tn = child_end(c, s, e, t, L, cpl, cbase, clast, ncell, &cc);
            if (tn - t > G) big |= 1u << ((keys[t] >> shift) & 7);
            t = tn;
        }
        for (int t = s; t <= e;) {
            int cc;
            const int tn = child_end(c, s, e, t, L, cpl, cbase, clast, ncell, &cc);
            visit(t, tn - t, (int)((keys[t] >> shift) & 7));
            t = tn;
        }
    }
    if (run) emit(rs, run);
    return k;
}

__global__ void k_group_count(const uint64_t *__restrict__ keys, const int8_t *__restrict__ cpl,
                              const uint32_t *__restrict__ cbase,
                              const int *__restrict__ cfirst, const int *__restrict__ clast,
                              const uint8_t *__restrict__ clevel, const unsigned long long *ncellp, int G,
                              int lsys, const int *__restrict__ csub, const int2 *__restrict__ fnb, uint32_t *gcnt)
{
    const int ncell = (int)*ncellp;
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < ncell; c += gridDim.x * blockDim.x)
        gcnt[c] = pack_groups<false>(c, keys, cpl, cbase, cfirst, clast, clevel, G, lsys, ncell, csub, fnb, nullptr,
                                     nullptr, 0);
}

__global__ void k_group_write(const uint64_t *__restrict__ keys, const int8_t *__restrict__ cpl,
                              const uint32_t *__restrict__ cbase,
                              const int *__restrict__ cfirst, const int *__restrict__ clast,
                              const uint8_t *__restrict__ clevel, const unsigned long long *ncellp, int G, int lsys,
                              const int *__restrict__ csub, const int2 *__restrict__ fnb,
                              const uint32_t *__restrict__ goff, int *gfirst, int *gcount)
{
    const int ncell = (int)*ncellp;
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < ncell; c += gridDim.x * blockDim.x)
        pack_groups<true>(c, keys, cpl, cbase, cfirst, clast, clevel, G, lsys, ncell, csub, fnb, gfirst, gcount,
                          goff[c]);
}

// Walk roots of a batch: the system's root cell (the cell at the first
// spatial level, lsys) or, for a one-body system, that body.
__global__ void k_sysroot_init(const int64_t *__restrict__ off, int nsys, int *root)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < nsys) root[s] = (int)off[s];
}
__global__ void k_sysroot_cells(const uint8_t *__restrict__ clevel, const int *__restrict__ cfirst,
                                const uint64_t *__restrict__ keys, int lsys, int sysbits,
                                const unsigned long long *ncellp, int *root)
{
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= (int)*ncellp || clevel[c] != lsys) return;
    const int s = sysbits ? (int)(keys[cfirst[c]] >> (63 - sysbits)) : 0;
    root[s] = ~c;
}

// Children of each cell in octant order; parent links; child-cell counts.
__device__ void children_of(int c, const uint64_t *__restrict__ keys, const int8_t *__restrict__ cpl,
                            const uint32_t *__restrict__ cbase, const int *__restrict__ cfirst,
                            const int *__restrict__ clast, const uint8_t *__restrict__ clevel, int ncell, int *csub,
                            int *cnkids);

__global__ void k_children(const uint64_t *__restrict__ keys, const int8_t *__restrict__ cpl,
                           const uint32_t *__restrict__ cbase, const int *__restrict__ cfirst,
                           const int *__restrict__ clast, const uint8_t *__restrict__ clevel,
                           const unsigned long long *ncellp, int *csub, int *cnkids)
{
    const int ncell = (int)*ncellp;
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < ncell; c += gridDim.x * blockDim.x)
        children_of(c, keys, cpl, cbase, cfirst, clast, clevel, ncell, csub, cnkids);
}

__device__ void children_of(int c, const uint64_t *__restrict__ keys, const int8_t *__restrict__ cpl,
                            const uint32_t *__restrict__ cbase, const int *__restrict__ cfirst,
                            const int *__restrict__ clast, const uint8_t *__restrict__ clevel, int ncell, int *csub,
                            int *cnkids)
{
    int s = cfirst[c], e = clast[c], L = clevel[c];
    int r[8];
#pragma unroll
    for (int o = 0; o < 8; o++) r[o] = kNoChild;
    int nk = 0;
    if (L == kMaxLevel) {            // bucket: bodies s..e are the children
        nk = e - s + 1;
    } else {
        const int shift = 60 - 3 * L;    // octant digit of level L+1
        int t = s;
        while (t <= e) {
            const int o = (int)((keys[t] >> shift) & 7);
            int cc;
            const int tn = child_end(c, s, e, t, L, cpl, cbase, clast, ncell, &cc);
            const int ref = cc >= 0 ? ~cc : t;
#pragma unroll
            for (int k = 0; k < 8; k++)
                if (k == o) r[k] = ref;
            t = tn;
            nk++;
        }
    }
    int4 *sub = reinterpret_cast<int4 *>(csub + 8 * (int64_t)c);
    sub[0] = make_int4(r[0], r[1], r[2], r[3]);
    sub[1] = make_int4(r[4], r[5], r[6], r[7]);
    cnkids[c] = nk;
}

// Moments of cell c from its children's (computed before it: the levels
// are processed bottom-up).  Accumulation follows
// hackcofm's order: children in octant order, FP64, no FMA (this file is
// compiled with -ffp-contract=off).
// The records of cell c from its FP64 mass, c.o.m. and (QUAD) traceless
// quadrupole q (xx xy xz yy yz zz): the FP64 moments (10 per cell), the split
// FP32 c.o.m. with the mass, the size^2 (infinite above a batch's system
// roots, Lg < 0), the FP32 quadrupole (cQ0 non-null) and the learned
// features (cF0 non-null): c.o.m. offset from the geometric centre and the
// quadrupole, in units of the cell size.
template <bool QUAD>
__device__ __forceinline__ void cell_records(int c, double M, double px, double py, double pz, const double *q, int L,
                                             int Lg, double rsize, double4 box, const uint64_t *__restrict__ keys,
                                             const int *__restrict__ cfirst, double *mo, float4 *cA, float4 *cB,
                                             float4 *cQ0, float4 *cQ1, float4 *cF0, float4 *cF1)
{
    double *me = mo + 10 * (int64_t)c;
    me[0] = M;
    me[1] = px;
    me[2] = py;
    me[3] = pz;
    if (QUAD) {
#pragma unroll
        for (int u = 0; u < 6; u++) me[4 + u] = q[u];
    }
    float4 A, B;
    split2(px, A.x, B.x);
    split2(py, A.y, B.y);
    split2(pz, A.z, B.z);
    A.w = (float)M;
    B.w = Lg >= 0 ? (float)(rsize * rsize * ldexp(1.0, -2 * Lg)) : 3.0e38f;
    cA[c] = A;
    cB[c] = B;
    if (QUAD && cQ0) {
        cQ0[c] = make_float4((float)q[0], (float)q[1], (float)q[2], (float)q[3]);
        cQ1[c] = make_float4((float)q[4], (float)q[5], 0.f, 0.f);
    }
    if (cF0 && Lg >= 0) {
        // geometric centre of the cell from the key prefix of its first body
        const uint64_t key = keys[cfirst[c]];
        uint64_t ix = 0, iy = 0, iz = 0;
        for (int l = L - Lg; l < L; l++) {
            const int sh = 60 - 3 * l;
            ix = (ix << 1) | ((key >> (sh + 2)) & 1);
            iy = (iy << 1) | ((key >> (sh + 1)) & 1);
            iz = (iz << 1) | ((key >> sh) & 1);
        }
        const double size = rsize * ldexp(1.0, -Lg);
        const double gx = box.x + ((double)ix + 0.5) * size, gy = box.y + ((double)iy + 0.5) * size,
                     gz = box.z + ((double)iz + 0.5) * size;
        const double qn = 1.0 / (M * size * size);
        cF0[c] = make_float4((float)((px - gx) / size), (float)((py - gy) / size), (float)((pz - gz) / size),
                             (float)(q[4] * qn));
        cF1[c] = make_float4((float)(q[0] * qn), (float)(q[3] * qn), (float)(q[1] * qn), (float)(q[2] * qn));
    }
}

template <bool QUAD>
__device__ void cell_moments(int c, const int *__restrict__ csub, const int *__restrict__ cfirst,
                             const int *__restrict__ clast, const uint8_t *__restrict__ clevel,
                             const double4 *__restrict__ bd, double *mo,
                             float4 *cA, float4 *cB, float4 *cQ0, float4 *cQ1, double rsize,
                             const uint64_t *__restrict__ keys, double4 box, float4 *cF0, float4 *cF1, int lsys,
                             int sysbits, const double4 *__restrict__ sysbox)
{
    // rsize: side of the root box (the cell at level L has side rsize / 2^L).
    // Batches: the box of the cell's system, levels counted from its root
    // (lsys); the cells above the system roots get an infinite size.
    const int L = clevel[c];
    int Lg = L;
    if (sysbox) {
        const int sys = sysbits ? (int)(keys[cfirst[c]] >> (63 - sysbits)) : 0;
        box = sysbox[sys];
        rsize = box.w;
        Lg = L - lsys;
    }
    double M = 0.0, px = 0.0, py = 0.0, pz = 0.0;
    double q[6] = {0, 0, 0, 0, 0, 0};
    if (L == kMaxLevel) {
        // bucket: the bodies s..e are the children, in sorted order
        const int s = cfirst[c], e = clast[c];
        for (int t = s; t <= e; t++) {
            double4 b = bd[t];
            M += b.w;
            double t0 = b.x * b.w, t1 = b.y * b.w, t2 = b.z * b.w;
            px = px + t0;
            py = py + t1;
            pz = pz + t2;
        }
        px = px / M;
        py = py / M;
        pz = pz / M;
        if (QUAD) {
            for (int t = s; t <= e; t++) {
                double4 b = bd[t];
                double dr[3] = {b.x - px, b.y - py, b.z - pz};
                double drsq = dr[0] * dr[0];
                drsq += dr[1] * dr[1];
                drsq += dr[2] * dr[2];
                const int ia[6] = {0, 0, 0, 1, 1, 2}, ib[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
                for (int u = 0; u < 6; u++) {
                    double idr = (ia[u] == ib[u] ? 1.0 : 0.0) * drsq;
                    double tm = (dr[ia[u]] * dr[ib[u]]) * 3.0;
                    tm = tm - idr;
                    tm = tm * b.w;
                    q[u] = q[u] + tm;
                }
            }
        }
    } else {
        // up to 8 children in octant order; all loads issued before the sums
        const int4 s0 = reinterpret_cast<const int4 *>(csub)[2 * (int64_t)c];
        const int4 s1 = reinterpret_cast<const int4 *>(csub)[2 * (int64_t)c + 1];
        const int r[8] = {s0.x, s0.y, s0.z, s0.w, s1.x, s1.y, s1.z, s1.w};
        double m[8], x[8], y[8], z[8];
#pragma unroll
        for (int k = 0; k < 8; k++) {
            m[k] = x[k] = y[k] = z[k] = 0.0;
            if (r[k] == kNoChild) continue;
            if (r[k] >= 0) {
                double4 b = bd[r[k]];
                m[k] = b.w; x[k] = b.x; y[k] = b.y; z[k] = b.z;
            } else {
                const double2 *mc = reinterpret_cast<const double2 *>(mo + 10 * (int64_t)(~r[k]));
                double2 a = mc[0], b = mc[1];
                m[k] = a.x; x[k] = a.y; y[k] = b.x; z[k] = b.y;
            }
        }
#pragma unroll
        for (int k = 0; k < 8; k++) {
            if (r[k] == kNoChild) continue;
            M += m[k];
            double t0 = x[k] * m[k], t1 = y[k] * m[k], t2 = z[k] * m[k];
            px = px + t0;
            py = py + t1;
            pz = pz + t2;
        }
        px = px / M;
        py = py / M;
        pz = pz / M;
        if (QUAD) {
#pragma unroll
            for (int k = 0; k < 8; k++) {
                if (r[k] == kNoChild) continue;
                double qc[6] = {0, 0, 0, 0, 0, 0};
                const bool cell = r[k] < 0;
                if (cell) {
                    const double2 *mq = reinterpret_cast<const double2 *>(mo + 10 * (int64_t)(~r[k]) + 4);
                    double2 a = mq[0], b = mq[1], d = mq[2];
                    qc[0] = a.x; qc[1] = a.y; qc[2] = b.x; qc[3] = b.y; qc[4] = d.x; qc[5] = d.y;
                }
                double dr[3] = {x[k] - px, y[k] - py, z[k] - pz};
                double drsq = dr[0] * dr[0];
                drsq += dr[1] * dr[1];
                drsq += dr[2] * dr[2];
                const int ia[6] = {0, 0, 0, 1, 1, 2}, ib[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
                for (int u = 0; u < 6; u++) {
                    double idr = (ia[u] == ib[u] ? 1.0 : 0.0) * drsq;
                    double tm = (dr[ia[u]] * dr[ib[u]]) * 3.0;
                    tm = tm - idr;
                    tm = tm * m[k];
                    if (cell) tm = tm + qc[u];
                    q[u] = q[u] + tm;
                }
            }
        }
    }
    cell_records<QUAD>(c, M, px, py, pz, q, L, Lg, rsize, box, keys, cfirst, mo, cA, cB, cQ0, cQ1, cF0, cF1);
}

// ---- fast monopole trees: moments from prefix sums --------------------------
//
// A cell is a contiguous range [first, last] of the sorted bodies, so its mass
// and first moments are differences of the FP64 prefix sums of (m x, m y,
// m z, m): one pass over the bodies (beside the topology, on the second
// stream) and one O(1) step per cell, instead of 22 dependent level launches.
// The sums round differently from hackcofm's child-by-child order, so this
// path serves the fast walk only (monopole cells; quadrupole cells and learned
// features also from prefix sums of the second moments); strict mode keeps
// the level-by-level pass.  The difference
// of two prefixes carries an absolute error of ~1e-16 of the running sums:
// relative 1e-9 of a two-body cell's mass at N = 4M, far below the FP32 of
// the walk.
constexpr int kPreThr = 256, kPrePer = 8;
static_assert(kPreThr * kPrePer == kPreTile, "prefix tile");

__device__ __forceinline__ double4 d4add(double4 a, double4 b) { return make_double4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w); }

__device__ __forceinline__ double4 body_mom(const double4 b) { return make_double4(b.x * b.w, b.y * b.w, b.z * b.w, b.w); }

// Bodies [lo, hi) of tile t: kPreTile-body tiles from 0 (one system), or the
// batch's tiles, which restart at every system's first body.
__device__ __forceinline__ int2 pre_tile(const int2 *__restrict__ tiles, int t, int n)
{
    return tiles ? tiles[t] : make_int2(t * kPreTile, min(n, (t + 1) * kPreTile));
}

// Block-wide exclusive scan of one double4 per thread (kPreThr threads).
__device__ double4 block_excl_scan(double4 v, double4 *tot)
{
    __shared__ double4 sw[kPreThr / 64];
    const int lane = lane_id(), w = threadIdx.x >> 6;
    double4 x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        double4 t;
        t.x = __shfl_up(x.x, d, 64);
        t.y = __shfl_up(x.y, d, 64);
        t.z = __shfl_up(x.z, d, 64);
        t.w = __shfl_up(x.w, d, 64);
        if (lane >= d) x = d4add(x, t);
    }
    if (lane == 63) sw[w] = x;
    __syncthreads();
    double4 off = make_double4(0, 0, 0, 0), all = make_double4(0, 0, 0, 0);
#pragma unroll
    for (int k = 0; k < kPreThr / 64; k++) {
        if (k < w) off = d4add(off, sw[k]);
        all = d4add(all, sw[k]);
    }
    __syncthreads();
    if (tot) *tot = all;
    return d4add(off, make_double4(x.x - v.x, x.y - v.y, x.z - v.z, x.w - v.w));
}

// Gather of the mass set into sorted order: bd (FP64 x, y, z, m) and the
// split FP32 records bhi/blo.  pk: packed (x, y, z, m) records in input
// order, or nullptr to read pos/mass.  One block per tile of kPreTile sorted
// bodies (blockIdx.x, blockIdx.x + gridDim.x, ...); in round k thread j takes
// body k * kPreThr + j of the tile (coalesced), all rounds' loads issued
// together.  loc (prefix-moment builds): the tile-local inclusive sums of
// (m x, m y, m z, m), one block scan per round, and part[t] the tile's total,
// from the records just gathered (no second pass over bd).  NV = 3 (cells
// with the quadrupole): also the second moments m (xx, xy, xz, yy) and
// m (yz, zz), as vectors 1 and 2 (loc[nb + i] and the double2 array at
// loc + 2 nb; part[v ntile + t], the second as (yz, zz, 0, 0)).  NV = 1
// keeps the plain gather's registers low (4 waves per SIMD against 2).
template <int NV>
__global__ __launch_bounds__(kPreThr) void k_gather_tree(const uint32_t *__restrict__ perm, int nb,
                                                         const double *__restrict__ pos, const double *__restrict__ mass,
                                                         const double4 *__restrict__ pk, double4 *bd, float4 *bhi,
                                                         float4 *blo, const int2 *__restrict__ tiles, int ntile,
                                                         double4 *loc, double4 *part)
{
    static_assert(NV == 1 || NV == 3, "prefix vectors");
    for (int t = blockIdx.x; t < ntile; t += gridDim.x) {
        const int2 r = pre_tile(tiles, t, nb);
        double4 v[kPrePer];
#pragma unroll
        for (int k = 0; k < kPrePer; k++) {
            const int i = r.x + k * kPreThr + (int)threadIdx.x;
            v[k] = make_double4(0, 0, 0, 0);
            if (i < r.y) {
                const int64_t o = perm[i];
                if (pk) {
                    v[k] = pk[o];
                } else {
                    v[k] = make_double4(pos[3 * o], pos[3 * o + 1], pos[3 * o + 2], mass[o]);
                }
            }
        }
        double4 carry = make_double4(0, 0, 0, 0), carry1 = carry, carry2 = carry;
#pragma unroll
        for (int k = 0; k < kPrePer; k++) {
            const int i = r.x + k * kPreThr + (int)threadIdx.x;
            const double4 b = v[k];
            if (i < r.y) {
                bd[i] = b;
                float4 h, l;
                split2(b.x, h.x, l.x);
                split2(b.y, h.y, l.y);
                split2(b.z, h.z, l.z);
                h.w = (float)b.w;
                l.w = 0.f;
                bhi[i] = h;
                blo[i] = l;
            }
            if (loc) {
                const double4 mm = body_mom(b);
                double4 tot;
                const double4 ex = block_excl_scan(mm, &tot);
                if (i < r.y) loc[i] = d4add(d4add(carry, ex), mm);
                carry = d4add(carry, tot);
                if (NV == 3) {
                    const double4 m1 = make_double4(mm.x * b.x, mm.x * b.y, mm.x * b.z, mm.y * b.y);
                    const double4 m2 = make_double4(mm.y * b.z, mm.z * b.z, 0.0, 0.0);
                    const double4 e1 = block_excl_scan(m1, &tot);
                    if (i < r.y) loc[nb + i] = d4add(d4add(carry1, e1), m1);
                    carry1 = d4add(carry1, tot);
                    const double4 e2 = block_excl_scan(m2, &tot);
                    const double4 s2 = d4add(d4add(carry2, e2), m2);
                    if (i < r.y) reinterpret_cast<double2 *>(loc + 2 * (int64_t)nb)[i] = make_double2(s2.x, s2.y);
                    carry2 = d4add(carry2, tot);
                }
            }
        }
        if (loc && threadIdx.x == 0) {
            part[t] = carry;
            if (NV == 3) {
                part[ntile + t] = carry1;
                part[2 * ntile + t] = carry2;
            }
        }
    }
}

// Exclusive scan of the tile sums of one system per block.
// (vector blockIdx.y of the tile sums, part[v ntile + t])
__global__ __launch_bounds__(kPreThr) void k_pre_top(double4 *part, const int *__restrict__ seg, int ntile)
{
    const int a = seg ? seg[blockIdx.x] : 0, b = seg ? seg[blockIdx.x + 1] : ntile;
    double4 *pv = part + (int64_t)blockIdx.y * ntile;
    double4 carry = make_double4(0, 0, 0, 0);
    for (int base = a; base < b; base += kPreThr) {
        const int i = base + threadIdx.x;
        const double4 v = i < b ? pv[i] : make_double4(0, 0, 0, 0);
        double4 tot;
        const double4 e = block_excl_scan(v, &tot);
        if (i < b) pv[i] = d4add(carry, e);
        carry = d4add(carry, tot);
    }
}

__device__ __forceinline__ double4 d4sub(double4 a, double4 b) { return make_double4(a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w); }

// Mass, c.o.m. and the FP32 walk records of every cell from the prefixes
// (the records cell_moments writes for a monopole build).  The sums over the
// cell's bodies [s, e] are loc[e] - loc[s-1] inside one tile, else the
// difference of the tiles' exclusive prefixes part[] plus that of the
// tile-local sums.  Tile of body i: seg[sys] + (i - s0) / kPreTile, s0 the
// system's first body (one system: i / kPreTile).  In a batch, the cells
// above the systems' roots are never walked and get no moments.
template <bool QUAD>
__global__ __launch_bounds__(kTpb) void k_moments_prefix(const int *__restrict__ cfirst, const int *__restrict__ clast,
                                                         const uint8_t *__restrict__ clevel,
                                                         const unsigned long long *ncellp, int nb, int ntile,
                                                         const double4 *__restrict__ loc, const double4 *__restrict__ part,
                                                         const int *__restrict__ seg, const uint64_t *__restrict__ keys,
                                                         double *mo, float4 *cA, float4 *cB, float4 *cQ0, float4 *cQ1,
                                                         float4 *cF0, float4 *cF1, const double4 *__restrict__ box,
                                                         int lsys, int sysbits, const double4 *__restrict__ sysbox,
                                                         const int64_t *__restrict__ boff)
{
    const int ncell = (int)*ncellp;
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < ncell; c += gridDim.x * blockDim.x) {
        const int s = cfirst[c], e = clast[c];
        const int L = clevel[c];
        int Lg = L, s0 = 0, t0 = 0;
        double4 bx = *box;
        if (sysbox) {
            const int sys = sysbits ? (int)(keys[s] >> (63 - sysbits)) : 0;
            bx = sysbox[sys];
            Lg = L - lsys;
            s0 = (int)boff[sys];
            t0 = seg[sys];
        }
        double *me = mo + 10 * (int64_t)c;
        if (Lg < 0) {
#pragma unroll
            for (int u = 0; u < (QUAD ? 10 : 4); u++) me[u] = 0.0;
            cA[c] = make_float4(0.f, 0.f, 0.f, 0.f);
            cB[c] = make_float4(0.f, 0.f, 0.f, 3.0e38f);
            if (QUAD && cQ0) cQ0[c] = cQ1[c] = make_float4(0.f, 0.f, 0.f, 0.f);
            continue;
        }
        const int tb = t0 + (e - s0) / kPreTile;
        const int ta = s > s0 ? t0 + (s - 1 - s0) / kPreTile : t0;
        // last body of tile ta (its tile-local sum is the tile's total)
        const int ea = s0 + (ta - t0 + 1) * kPreTile - 1;
        // the sums of vector v over the bodies s..e: inside one tile, and
        // across one tile boundary, from tile-local sums only (small cells
        // keep the precision of their own magnitudes); across more, the
        // difference of the tiles' prefixes (cells of over kPreTile bodies)
        auto at = [&](int v, int i) {
            if (v < 2) return loc[(int64_t)v * nb + i];
            const double2 q2 = reinterpret_cast<const double2 *>(loc + 2 * (int64_t)nb)[i];
            return make_double4(q2.x, q2.y, 0.0, 0.0);
        };
        auto span = [&](int v) {
            const double4 *pv = part + (int64_t)v * ntile;
            double4 d = at(v, e);
            if (s > s0) {
                const double4 la = at(v, s - 1);
                if (ta == tb) d = d4sub(d, la);
                else if (tb == ta + 1) d = d4add(d4sub(at(v, ea), la), d);
                else d = d4add(d4sub(pv[tb], pv[ta]), d4sub(d, la));
            } else if (tb != t0) {
                d = d4add(pv[tb], d);
            }
            return d;
        };
        const double4 d = span(0);
        const double M = d.w;
        const double px = d.x / M, py = d.y / M, pz = d.z / M;
        double q[6] = {0, 0, 0, 0, 0, 0};
        if (QUAD) {
            // second moments about the c.o.m. (parallel axis), then the
            // traceless quadrupole sum m (3 d d - |d|^2 I) hackcofm builds
            const double4 d1 = span(1), d2 = span(2);
            const double t[6] = {d1.x - d.x * px, d1.y - d.x * py, d1.z - d.x * pz,
                                 d1.w - d.y * py, d2.x - d.y * pz, d2.y - d.z * pz};
            const double tr = (t[0] + t[3]) + t[5];
            q[0] = 3.0 * t[0] - tr;
            q[1] = 3.0 * t[1];
            q[2] = 3.0 * t[2];
            q[3] = 3.0 * t[3] - tr;
            q[4] = 3.0 * t[4];
            q[5] = 3.0 * t[5] - tr;
        }
        cell_records<QUAD>(c, M, px, py, pz, q, L, Lg, bx.w, bx, keys, cfirst, mo, cA, cB, cQ0, cQ1, cF0, cF1);
    }
}

// Cells bucketed by level so the moments can be computed level by level,
// deepest first, with a kernel boundary between levels: a counting sort with
// one global atomic per (block, level).  The order of cells inside a level
// is arbitrary; each cell's moments depend only on its children.
constexpr int kLevels = kMaxLevel + 1;

// Lanes of the wave holding the same level (levels < 32): a 5-ballot match.
__device__ __forceinline__ uint64_t level_peers(int L, bool ok)
{
    uint64_t peers = __ballot(ok);
#pragma unroll
    for (int b = 0; b < 5; b++) {
        const uint64_t bal = __ballot((L >> b) & 1);
        peers &= ((L >> b) & 1) ? bal : ~bal;
    }
    return peers;
}

__global__ __launch_bounds__(kTpb) void k_level_count(const uint8_t *__restrict__ clevel,
                                                      const unsigned long long *ncellp, uint32_t *lcount)
{
    __shared__ uint32_t h[kLevels];
    if (threadIdx.x < kLevels) h[threadIdx.x] = 0;
    __syncthreads();
    const int ncell = (int)*ncellp;
    const int l = lane_id();
    const uint64_t lt = l ? (~0ULL >> (64 - l)) : 0ULL;
    for (int base = blockIdx.x * blockDim.x; base < ncell; base += gridDim.x * blockDim.x) {
        const int c = base + threadIdx.x;
        const bool ok = c < ncell;
        const int L = ok ? clevel[c] : 0;
        const uint64_t peers = level_peers(L, ok);
        if (ok && __popcll(peers & lt) == 0) atomicAdd(&h[L], (uint32_t)__popcll(peers));
    }
    __syncthreads();
    if (threadIdx.x < kLevels && h[threadIdx.x]) atomicAdd(&lcount[threadIdx.x], h[threadIdx.x]);
}

// Each block computes the level offsets (the scan of the 22 counts; block 0
// also stores them in loff[0..kLevels] for the moments), counts its cells per
// level, reserves ranges in the level lists (cursors start at 0), then places
// its cells (same grid-stride order in both passes).
__global__ __launch_bounds__(kTpb) void k_level_place(const uint8_t *__restrict__ clevel,
                                                      const unsigned long long *ncellp,
                                                      const uint32_t *__restrict__ lcount, uint32_t *loff,
                                                      uint32_t *lcursor, int *list)
{
    __shared__ uint32_t h[kLevels], base[kLevels], off[kLevels + 1];
    if (threadIdx.x == 0) {
        uint32_t acc = 0;
        for (int L = 0; L < kLevels; L++) {
            off[L] = acc;
            acc += lcount[L];
        }
        off[kLevels] = acc;
    }
    if (threadIdx.x < kLevels) h[threadIdx.x] = 0;
    __syncthreads();
    if (blockIdx.x == 0 && threadIdx.x <= kLevels) loff[threadIdx.x] = off[threadIdx.x];
    const int ncell = (int)*ncellp;
    const int l = lane_id();
    const uint64_t lt = l ? (~0ULL >> (64 - l)) : 0ULL;
    for (int b0 = blockIdx.x * blockDim.x; b0 < ncell; b0 += gridDim.x * blockDim.x) {
        const int c = b0 + threadIdx.x;
        const bool ok = c < ncell;
        const int L = ok ? clevel[c] : 0;
        const uint64_t peers = level_peers(L, ok);
        if (ok && __popcll(peers & lt) == 0) atomicAdd(&h[L], (uint32_t)__popcll(peers));
    }
    __syncthreads();
    if (threadIdx.x < kLevels) {
        base[threadIdx.x] = off[threadIdx.x] + (h[threadIdx.x] ? atomicAdd(&lcursor[threadIdx.x], h[threadIdx.x]) : 0u);
        h[threadIdx.x] = 0;
    }
    __syncthreads();
    for (int b0 = blockIdx.x * blockDim.x; b0 < ncell; b0 += gridDim.x * blockDim.x) {
        const int c = b0 + threadIdx.x;
        const bool ok = c < ncell;
        const int L = ok ? clevel[c] : 0;
        const uint64_t peers = level_peers(L, ok);
        const uint32_t rank = __popcll(peers & lt);
        const int leader = ok ? __ffsll((unsigned long long)peers) - 1 : 0;
        uint32_t wb = 0;
        if (ok && rank == 0) wb = atomicAdd(&h[L], (uint32_t)__popcll(peers));
        wb = __shfl(wb, leader, 64);
        if (ok) list[base[L] + wb + rank] = c;
    }
}

template <bool QUAD>
__global__ __launch_bounds__(kTpb) void k_moments_level(const int *__restrict__ list, const uint32_t *__restrict__ loff, int L,
                                const uint64_t *__restrict__ keys, float4 *cF0, float4 *cF1,
                                const int *__restrict__ csub,
                                const int *__restrict__ cfirst, const int *__restrict__ clast,
                                const uint8_t *__restrict__ clevel, const double4 *__restrict__ bd, double *mo,
                                float4 *cA, float4 *cB, float4 *cQ0, float4 *cQ1,
                                const double4 *__restrict__ box, int lsys, int sysbits,
                                const double4 *__restrict__ sysbox)
{
    const double rsize = box->w;
    const double4 bx = *box;
    const uint32_t lo = loff[L], hi = loff[L + 1];
    for (uint32_t k = lo + blockIdx.x * blockDim.x + threadIdx.x; k < hi; k += gridDim.x * blockDim.x)
        cell_moments<QUAD>(list[k], csub, cfirst, clast, clevel, bd, mo, cA, cB, cQ0, cQ1, rsize, keys, bx, cF0,
                           cF1, lsys, sysbits, sysbox);
}

inline int nblocks(int64_t n, int tpb = kTpb) { return (int)std::max<int64_t>(1, (n + tpb - 1) / tpb); }
// Grid of the kernels that stride over a device-side count bounded by cap.
constexpr int kGatherBlocksPerCU = 8;   // grid of the tree build's gather (tiles of kPreTile bodies), per CU
constexpr int kBodyBlocksPerCU = 16;    // grid of the per-body topology kernels (grid-stride), per CU
constexpr int kCapBlocksPerCU = 8;      // grid of the per-cell kernels (grid-stride over the device count), per CU
inline int capgrid(const bh_ctx *ctx, int64_t cap)
{
    return (int)std::min<int64_t>(nblocks(cap), (int64_t)ctx->ncu * kCapBlocksPerCU);
}

}  // namespace

// ---- host launchers ------------------------------------------------------------

void exclusive_scan_u32(bh_ctx *ctx, const uint32_t *in, uint32_t *out, int64_t n, const unsigned long long *dn)
{
    // out must hold n + 1 entries; out[n] = total (in != out)
    int64_t nparts = std::max<int64_t>(1, (n + kScanChunk - 1) / kScanChunk);
    uint32_t *partial = ensure<uint32_t>(ctx->scan_tmp, nparts + 1);
    k_scan_reduce<<<nparts, kTpb, 0, ctx->stream>>>(in, n, dn, partial);
    k_scan_top<<<1, kTpb, 0, ctx->stream>>>(partial, nparts, out, n);
    k_scan_down<<<nparts, kTpb, 0, ctx->stream>>>(in, out, n, dn, partial);
    BH_HIP(hipGetLastError());
}

void radix_sort_pairs(bh_ctx *ctx, uint64_t *k0, uint64_t *k1, uint32_t *v0, uint32_t *v1, int64_t n,
                      int begin_bit, int end_bit, uint64_t **kout, uint32_t **vout)
{
    int nblk = (int)std::max<int64_t>(1, (n + kSortTile - 1) / kSortTile);
    uint32_t *hist = ensure<uint32_t>(ctx->hist, 2 * (256 * (size_t)nblk + 1) + 256);
    uint32_t *hscan = hist + 256 * (size_t)nblk + 1;
    uint32_t *dtot = hscan + 256 * (size_t)nblk + 1;
    uint64_t *ka = k0, *kb = k1;
    uint32_t *va = v0, *vb = v1;
    for (int shift = begin_bit; shift < end_bit; shift += 8) {
        k_radix_hist<<<nblk, kTpb, 0, ctx->stream>>>(ka, n, shift, hist, nblk);
        k_hist_rows<<<256, kTpb, 0, ctx->stream>>>(hist, nblk, hscan, dtot);
        k_radix_scatter<<<nblk, kTpb, 0, ctx->stream>>>(ka, va, kb, vb, n, shift, hist, hscan, dtot, nblk);
        std::swap(ka, kb);
        std::swap(va, vb);
    }
    BH_HIP(hipGetLastError());
    *kout = ka;
    *vout = va;
}

// Cell ranges of the sorted key array keys[0..n): cpl, cbase (n+1 entries),
// block minima, and first/last/level of every cell, launched over the cell
// capacity with the cell count kept on the device (stat[slot]).  n >= 2.
struct Topo {
    int8_t *cpl;
    uint32_t *cbase;
    int *first, *last;
    uint8_t *level;
    int cap;
    const unsigned long long *ncellp;
};

static Topo topology(bh_ctx *ctx, const uint64_t *keys, int n, int cap, DevBuf &bcpl, DevBuf &bcnt, DevBuf &bbase,
                     DevBuf &bm1, DevBuf &bfirst, DevBuf &blast, DevBuf &blevel, DevBuf *bskip, DevBuf *bpre,
                     DevBuf *bfnb, DevBuf &bowner, int slot, int slot_raw, unsigned long long flag)
{
    hipStream_t st = ctx->stream;
    unsigned long long *stat = ptr<unsigned long long>(ctx->counters);
    Topo t{};
    t.cap = cap;
    t.ncellp = stat + slot;
    t.cpl = ensure<int8_t>(bcpl, 2 * (size_t)n + 32);
    uint8_t *cq = reinterpret_cast<uint8_t *>(t.cpl) + ((n + 15) & ~7);   // 8-aligned, n + 8 bytes
    uint32_t *ccnt = ensure<uint32_t>(bcnt, n + 1);
    t.cbase = ensure<uint32_t>(bbase, n + 1);
    const int bgrid = (int)std::min<int64_t>(nblocks(n), (int64_t)ctx->ncu * kBodyBlocksPerCU);
    k_cpl<<<bgrid, kTpb, 0, st>>>(keys, n, t.cpl, cq, ccnt);
    exclusive_scan_u32(ctx, ccnt, t.cbase, n);
    int n1 = (n + 63) / 64, n2 = (n1 + 63) / 64, n3 = (n2 + 63) / 64;
    // each level 8-byte aligned with at least one word of padding (the
    // search reads whole words; it stops at the entry holding cpl[n-1] = -1)
    auto words = [](int v) { return (v + 15) & ~7; };
    uint8_t *m1 = reinterpret_cast<uint8_t *>(ensure<int8_t>(bm1, words(n1) + words(n2) + words(n3)));
    uint8_t *m2 = m1 + words(n1), *m3 = m2 + words(n2);
    const CellClamp cc{t.cbase, n, (unsigned long long)cap, stat, slot, slot_raw, flag};
    k_cmin<<<nblocks(8 * (int64_t)n1), kTpb, 0, st>>>(cq, n, m1, n1, cc);
    k_cmin<<<nblocks(8 * (int64_t)n2), kTpb, 0, st>>>(m1, n1, m2, n2);
    k_cmin<<<nblocks(8 * (int64_t)n3), kTpb, 0, st>>>(m2, n2, m3, n3);
    t.first = ensure<int>(bfirst, cap);
    t.last = ensure<int>(blast, cap);
    t.level = ensure<uint8_t>(blevel, cap);
    int *cskip = bskip ? ensure<int>(*bskip, cap) : nullptr;
    int *pre = bpre ? ensure<int>(*bpre, (size_t)n + cap) : nullptr;
    int2 *fnb = bfnb ? ensure<int2>(*bfnb, cap) : nullptr;
    int *owner = ensure<int>(bowner, cap);
    k_owner<<<bgrid, kTpb, 0, st>>>(t.cbase, n, t.ncellp, owner);
    k_cells<<<capgrid(ctx, cap), kTpb, 0, st>>>(t.cpl, cq, t.cbase, owner, m1, m2, m3, t.ncellp, t.first, t.last, t.level,
                                                cskip, pre, fnb, bskip ? stat + S_DEEP : nullptr);
    if (pre) k_body_pre<<<bgrid, kTpb, 0, st>>>(t.cbase, n, t.ncellp, pre);
    BH_HIP(hipGetLastError());
    return t;
}

// Groups of the sorted test set for the fast walk (ctx->gfirst/gcount, count
// in stat[S_NGROUPS]).
// csub/fnb: the tree's child slots and (first, count) per cell, when the
// groups are cut from the tree's own topology.
static void make_groups(bh_ctx *ctx, const Topo *t, const uint64_t *keys, int64_t n, int G, int lsys = 0,
                        const int *csub = nullptr, const int2 *fnb = nullptr)
{
    hipStream_t st = ctx->stream;
    unsigned long long *stat = ptr<unsigned long long>(ctx->counters);
    int *gf = ensure<int>(ctx->gfirst, n), *gc = ensure<int>(ctx->gcount, n);
    if (t == nullptr) {
        int64_t ng = (n + G - 1) / G;
        k_groups_trivial<<<nblocks(ng), kTpb, 0, st>>>(n, G, gf, gc, stat);
        return;
    }
    uint32_t *gcnt = ensure<uint32_t>(ctx->gscan, (size_t)t->cap + 1);
    uint32_t *goff = ensure<uint32_t>(ctx->goff, (size_t)t->cap + 1);
    k_group_count<<<capgrid(ctx, t->cap), kTpb, 0, st>>>(keys, t->cpl, t->cbase, t->first, t->last, t->level,
                                                         t->ncellp, G, lsys, csub, fnb, gcnt);
    exclusive_scan_u32(ctx, gcnt, goff, t->cap, t->ncellp);
    k_group_write<<<capgrid(ctx, t->cap), kTpb, 0, st>>>(keys, t->cpl, t->cbase, t->first, t->last, t->level,
                                                         t->ncellp, G, lsys, csub, fnb, goff, gf, gc);
    k_copy_count<<<1, 1, 0, st>>>(goff + t->cap, stat + S_NGROUPS);
    BH_HIP(hipGetLastError());
}

// Morton keys of the mass set in input order, in the current box.
void debug_keys(bh_ctx *ctx, uint64_t *d_out)
{
    double4 *d_box = ensure<double4>(ctx->dbox, 1);
    uint32_t *tmp;
    BH_HIP(hipMalloc(&tmp, ctx->nmass * sizeof(uint32_t)));
    k_keys<<<nblocks(ctx->nmass), kTpb, 0, ctx->stream>>>(ptr<double>(ctx->pos), ptr<double>(ctx->mass), ctx->nmass,
                                                          d_box, d_out, tmp, nullptr);
    BH_HIP(hipStreamSynchronize(ctx->stream));
    BH_HIP(hipFree(tmp));
}

// Groups of the sorted test set.  When the test set is exactly the tree's
// body set the tree topology is reused; otherwise the test keys get their own.
static void build_groups(bh_ctx *ctx, const bh_params &p, const Topo *tree_topo, const uint64_t *ks,
                         const uint64_t *tks)
{
    const int G = p.group_size > 0 ? p.group_size : 64;
    const int64_t nt = ctx->ntest;
    const Tree &T = ctx->tree;
    // at most G points: one group (a cell only cuts groups for children of a
    // cell holding more than G points); batches keep one group per system
    if (ctx->nsys == 0 && nt <= G) {
        make_groups(ctx, nullptr, nullptr, nt, G);
        return;
    }
    if (ctx->self_test && nt == T.nb && tree_topo) {
        make_groups(ctx, tree_topo, ks, nt, G, ctx->sysbits / 3, T.sub, T.fnb);
        return;
    }
    if (ctx->nsys > 1) throw ApiFail(BH_ERR_STATE, "batch groups need the system-aware topology");
    if (nt <= G || nt < 2) {
        make_groups(ctx, nullptr, nullptr, nt, G);
        return;
    }
    if (ctx->tcell_cap == 0) ctx->tcell_cap = (int)std::min<int64_t>(2 * nt + 64, INT32_MAX / 2);
    Topo t = topology(ctx, tks, (int)nt, ctx->tcell_cap, ctx->gcpl, ctx->gccnt, ctx->gcbase, ctx->gm1, ctx->gcf,
                      ctx->gcl, ctx->gclv, nullptr, nullptr, nullptr, ctx->gowner, S_TNCELL, S_TNCELL_RAW, F_TCELL_OVF);
    make_groups(ctx, &t, tks, nt, G);
}

// The build's second stream.  side_fork(k): ctx->side waits for the work
// queued so far on the main stream.  side_mark(k): marks the work queued so
// far on ctx->side; main_wait(k): the main stream waits for that mark.
// Under graph capture the side stream joins the capture through the forked
// event and every branch is joined back before the build returns.
enum { kForkSort = 0, kMarkGather = 1, kForkTopo = 2, kMarkLevels = 3, kForkKids = 4, kMarkGroups = 5 };

static void side_fork(bh_ctx *ctx, int k)
{
    BH_HIP(hipEventRecord(ctx->fork_ev[k], ctx->stream));
    BH_HIP(hipStreamWaitEvent(ctx->side, ctx->fork_ev[k], 0));
}

static void side_mark(bh_ctx *ctx, int k) { BH_HIP(hipEventRecord(ctx->fork_ev[k], ctx->side)); }

static void main_wait(bh_ctx *ctx, int k) { BH_HIP(hipStreamWaitEvent(ctx->stream, ctx->fork_ev[k], 0)); }

// build_groups on the side stream (its launches and scans use ctx->stream).
static void build_groups_side(bh_ctx *ctx, const bh_params &p, const Topo *tree_topo, const uint64_t *ks,
                              const uint64_t *tks)
{
    hipStream_t main = ctx->stream;
    ctx->stream = ctx->side;
    try {
        build_groups(ctx, p, tree_topo, ks, tks);
    } catch (...) {
        ctx->stream = main;
        throw;
    }
    ctx->stream = main;
}

// Builds the tree of the uploaded mass set and the sorted test set.  No host
// synchronisation: every count stays on the device and capacities come from
// the host (the caller checks stat[S_FLAGS] afterwards).
void build_tree(bh_ctx *ctx, const bh_params &p)
{
    hipStream_t st = ctx->stream;
    const int64_t n = ctx->nmass;
    const int nb = (int)ctx->nb;
    double *d_pos = ptr<double>(ctx->pos), *d_mass = ptr<double>(ctx->mass);
    double4 *d_box = ensure<double4>(ctx->dbox, 1);
    int *d_boxi = ensure<int>(ctx->dboxi, 8);
    unsigned long long *stat = ensure<unsigned long long>(ctx->counters, S_NSLOTS);
    const bool timing = !ctx->capturing;
    static_assert(S_ZERO <= 128, "k_build_start clears the status words with one 128-thread block");

    // -- box: start from the caller's box (or the previous one) and expand in
    //    input order; a fixed number of rounds, checked afterwards
    const bool keep = p.keep_box && ctx->box_valid;
    const double r0 = p.rmin_given ? p.rmin[0] : -p.rsize / 2.0;
    const double r1 = p.rmin_given ? p.rmin[1] : -p.rsize / 2.0;
    const double r2 = p.rmin_given ? p.rmin[2] : -p.rsize / 2.0;
    const int nsys = ctx->nsys;
    const int64_t *sysoff = nsys ? ptr<int64_t>(ctx->boff) : nullptr;
    double4 *sysbox = nsys ? ensure<double4>(ctx->sysbox, nsys) : nullptr;
    // level-list counters and cursors (k_level_count / k_level_place)
    uint32_t *lwork = ensure<uint32_t>(ctx->lhist, 3 * (kLevels + 1));
    static_assert(3 * (kLevels + 1) <= 128, "k_build_start clears the level counters with one 128-thread block");
    if (nsys) {
        k_build_start<<<1, 128, 0, st>>>(stat, S_ZERO, lwork, 3 * (kLevels + 1), nullptr, nullptr, 0.0, 0.0, 0.0, 0.0,
                                        0);
        k_sys_box<<<nsys, kTpb, 0, st>>>(d_pos, sysoff, sysbox, d_box, r0, r1, r2, p.rsize, keep ? 1 : 0,
                                         16 * ctx->box_rounds, stat);
    } else {
    k_build_start<<<1, 128, 0, st>>>(stat, S_ZERO, lwork, 3 * (kLevels + 1), d_box, d_boxi, r0, r1, r2, p.rsize,
                                    keep ? 1 : 0);
    double4 *cand = ensure<double4>(ctx->bcand, n);
    k_box_compact<<<(int)((n + kBoxChunk - 1) / kBoxChunk), kTpb, 0, st>>>(d_pos, d_mass, n, d_box, cand, d_boxi,
                                                                            stat);
    const int rounds = std::max(1, keep ? ctx->box_rounds_keep : ctx->box_rounds);
    for (int r = 0; r <= rounds; r++) k_box_round<<<ctx->ncu, kTpb, 0, st>>>(d_pos, cand, d_box, d_boxi);
    }

    // -- keys and sort of the mass set (massless bodies sort last)
    uint64_t *k0 = ensure<uint64_t>(ctx->keys0, n), *k1 = ensure<uint64_t>(ctx->keys1, n);
    uint32_t *v0 = ensure<uint32_t>(ctx->vals0, n), *v1 = ensure<uint32_t>(ctx->vals1, n);
    const int grid_n = (int)std::min<int64_t>(nblocks(n), (int64_t)ctx->ncu * 16);
    double4 *pk = nullptr;
    if (nsys) {
        k_keys_batch<<<grid_n, kTpb, 0, st>>>(d_pos, n, sysoff, nsys, sysbox, ctx->sysbits, k0, v0);
        k_sysroot_init<<<nblocks(nsys), kTpb, 0, st>>>(sysoff, nsys, ensure<int>(ctx->sysroot, nsys));
    } else {
        pk = ensure<double4>(ctx->bpk, n);
        k_keys<<<grid_n, kTpb, 0, st>>>(d_pos, d_mass, n, d_box, k0, v0, pk, d_boxi, stat);
    }
    uint64_t *ks;
    uint32_t *vs;
    if (timing) BH_HIP(hipEventRecord(ctx->ev[6], st));
    radix_sort_pairs(ctx, k0, k1, v0, v1, n, 0, 64, &ks, &vs);
    if (timing) BH_HIP(hipEventRecord(ctx->ev[7], st));

    Tree &T = ctx->tree;
    T.nb = nb;
    double4 *bd = ensure<double4>(ctx->bd, std::max(nb, 1));
    float4 *bhi = ensure<float4>(ctx->bhi, std::max(nb, 1));
    float4 *blo = ensure<float4>(ctx->blo, std::max(nb, 1));
    // fast builds take the moments from prefix sums of the gathered bodies
    // (on the second stream, beside the topology): the first moments, and
    // the second moments too when the walk's cells or the learned features
    // carry the quadrupole
    const bool feat_b = p.learned != 0 || ctx->mlp_sampling;
    const bool prefix_moments = p.mode == BH_MODE_FAST && nb > 1;
    const bool prefix_quad = prefix_moments && (p.usequad || feat_b);
    const int nvec = prefix_quad ? 3 : 1;
    double4 *mpre = prefix_moments ? ensure<double4>(ctx->mpre, (size_t)nvec * nb) : nullptr;
    // the gather (HBM-bound; with, for prefix-moment builds, the tile-local
    // sums) on the second stream beside the topology (latency-bound); only
    // the moments and the walk read its records.  A bounded grid, so that
    // the main stream's small launches beside it find free slots.  A batch's
    // prefix tiles restart at every system.
    const int2 *tiles = prefix_moments && nsys ? ptr<int2>(ctx->ptile) : nullptr;
    const int ntile = tiles ? ctx->nptile : (int)((nb + kPreTile - 1) / kPreTile);
    double4 *part = prefix_moments ? ensure<double4>(ctx->mpart, (size_t)nvec * ntile) : nullptr;
    side_fork(ctx, kForkSort);
    if (nb > 0) {
        const int ggrid = std::min(ntile, ctx->ncu * kGatherBlocksPerCU);
        if (nvec == 3)
            k_gather_tree<3><<<ggrid, kPreThr, 0, ctx->side>>>(vs, nb, d_pos, d_mass, pk, bd, bhi, blo, tiles, ntile,
                                                               mpre, part);
        else
            k_gather_tree<1><<<ggrid, kPreThr, 0, ctx->side>>>(vs, nb, d_pos, d_mass, pk, bd, bhi, blo, tiles, ntile,
                                                               mpre, part);
    }
    if (prefix_moments)
        k_pre_top<<<dim3(nsys ? nsys : 1, nvec), kPreThr, 0, ctx->side>>>(part, nsys ? ptr<int>(ctx->pseg) : nullptr,
                                                                         ntile);
    side_mark(ctx, kMarkGather);
    T.bd = bd;
    T.bhi = bhi;
    T.blo = blo;
    T.keys = ks;
    T.dbox = d_box;

    // -- test set, sorted by its own keys in the same box
    const int64_t nt = ctx->ntest;
    const bool alias = ctx->self_test && nt == nb;
    float4 *thi = alias ? nullptr : ensure<float4>(ctx->thi, nt), *tlo = alias ? nullptr : ensure<float4>(ctx->tlo, nt);
    double4 *td = alias ? nullptr : ensure<double4>(ctx->td, nt);
    int *tself = alias ? nullptr : ensure<int>(ctx->tself, nt);
    uint32_t *tperm = alias ? nullptr : ensure<uint32_t>(ctx->tperm, nt);
    const uint64_t *tks_for_groups = ks;
    T.thi = thi;
    T.tlo = tlo;
    T.td = td;
    T.tperm = tperm;
    T.tself = tself;
    T.tinv = nullptr;
    T.selfmode = ctx->self_test ? 2 : 0;
    if (alias) {
        // the test set is the tree's body set in the same order
        T.thi = bhi;
        T.tlo = blo;
        T.td = bd;
        T.tperm = vs;
        T.tself = nullptr;
        T.selfmode = 1;
    } else {
        // the test set sorted by its own keys in the tree's box.  When it is
        // the mass set with massless bodies among it, those bodies sort with
        // their neighbours here (not after every massive body, as in the
        // tree's order), so each group stays compact in space; a test point
        // finds its tree body (self pair) through the inverse of the tree's
        // sort, and each tree body its test index through tinv.
        uint64_t *tk0 = ensure<uint64_t>(ctx->tkeys0, nt), *tk1 = ensure<uint64_t>(ctx->tkeys1, nt);
        uint32_t *tv0 = ensure<uint32_t>(ctx->tvals0, nt), *tv1 = ensure<uint32_t>(ctx->tvals1, nt);
        const double *d_tpos = ctx->self_test ? d_pos : ptr<double>(ctx->tpos);
        k_keys<<<(int)std::min<int64_t>(nblocks(nt), (int64_t)ctx->ncu * 16), kTpb, 0, st>>>(d_tpos, nullptr, nt, d_box,
                                                                                            tk0, tv0, nullptr);
        uint64_t *tks;
        uint32_t *tvs;
        radix_sort_pairs(ctx, tk0, tk1, tv0, tv1, nt, 0, 64, &tks, &tvs);
        tks_for_groups = tks;
        BH_HIP(hipMemcpyAsync(tperm, tvs, nt * sizeof(uint32_t), hipMemcpyDeviceToDevice, st));
        int *tmap = nullptr, *tinv = nullptr;
        if (ctx->self_test) {
            tmap = ensure<int>(ctx->tmap, nt);
            tinv = ensure<int>(ctx->tinv, std::max(nb, 1));
            k_tree_index<<<nblocks(nt), kTpb, 0, st>>>(vs, nt, nb, tmap);
        }
        k_gather_test<<<nblocks(nt), kTpb, 0, st>>>(tperm, nt, d_tpos, thi, tlo, td, tself, tmap, tinv);
        T.tinv = tinv;
    }

    // -- topology over the cell capacity
    int *pre = ensure<int>(ctx->pre, std::max(nb, 1) + (size_t)std::max(ctx->cell_cap, 1));
    T.pre = pre;
    T.cap = 0;
    T.ncellp = stat + S_NCELL;
    if (nb <= 1) {
        main_wait(ctx, kMarkGather);
        if (nb == 1) k_set_int<<<1, 1, 0, st>>>(pre, 0);
        if (p.mode == BH_MODE_FAST) build_groups(ctx, p, nullptr, ks, tks_for_groups);
        return;
    }
    const int cap = ctx->cell_cap;
    Topo tt = topology(ctx, ks, nb, cap, ctx->cpl, ctx->ccnt, ctx->cbase, ctx->cmin1, ctx->cfirst, ctx->clast,
                       ctx->clevel, &ctx->cskip, &ctx->pre, &ctx->cfnb, ctx->cowner, S_NCELL, S_NCELL_RAW, F_CELL_OVF);
    T.cap = cap;
    int8_t *cpl = tt.cpl;
    uint32_t *cbase = tt.cbase;
    int *cfirst = tt.first, *clast = tt.last;
    uint8_t *clevel = tt.level;
    T.pre = ptr<int>(ctx->pre);
    int *csub = ensure<int>(ctx->csub, 8 * (size_t)cap);
    int *cnkids = ensure<int>(ctx->cnkids, cap);
    double *mo = ensure<double>(ctx->cmo, 10 * (size_t)cap);
    float4 *cA = ensure<float4>(ctx->cA, cap), *cB = ensure<float4>(ctx->cB, cap);
    float4 *cQ0 = ensure<float4>(ctx->cQ0, cap), *cQ1 = ensure<float4>(ctx->cQ1, cap);
    float4 *cF0 = ensure<float4>(ctx->cF0, cap), *cF1 = ensure<float4>(ctx->cF1, cap);
    // cells bucketed by level (side stream) beside the child slots (main)
    uint32_t *lcount = lwork, *loffd = lwork + (kLevels + 1), *lcursor = lwork + 2 * (kLevels + 1);
    int *llist = ensure<int>(ctx->llist, cap);
    const int lsys = ctx->sysbits / 3;
    side_fork(ctx, kForkTopo);
    hipStream_t sd = ctx->side;
    if (prefix_moments) {
        // the moments on the second stream (after its prefix scan), beside
        // the child slots and the groups on the main stream
        const int *pseg = nsys ? ptr<int>(ctx->pseg) : nullptr;
        const int64_t *boff = nsys ? ptr<int64_t>(ctx->boff) : nullptr;
        if (prefix_quad)
            k_moments_prefix<true><<<capgrid(ctx, cap), kTpb, 0, sd>>>(
                cfirst, clast, clevel, tt.ncellp, nb, ntile, mpre, part, pseg, ks, mo, cA, cB,
                p.usequad ? cQ0 : nullptr, cQ1, feat_b ? cF0 : nullptr, cF1, d_box, lsys, ctx->sysbits, sysbox, boff);
        else
            k_moments_prefix<false><<<capgrid(ctx, cap), kTpb, 0, sd>>>(
                cfirst, clast, clevel, tt.ncellp, nb, ntile, mpre, part, pseg, ks, mo, cA, cB, nullptr, nullptr,
                nullptr, nullptr, d_box, lsys, ctx->sysbits, sysbox, boff);
    } else {
        k_level_count<<<capgrid(ctx, cap), kTpb, 0, sd>>>(clevel, tt.ncellp, lcount);
        k_level_place<<<capgrid(ctx, cap), kTpb, 0, sd>>>(clevel, tt.ncellp, lcount, loffd, lcursor, llist);
    }
    side_mark(ctx, kMarkLevels);
    k_children<<<capgrid(ctx, cap), kTpb, 0, st>>>(ks, cpl, cbase, cfirst, clast, clevel, tt.ncellp, csub, cnkids);
    T.first = cfirst;
    T.last = clast;
    T.fnb = ptr<int2>(ctx->cfnb);
    T.level = clevel;
    T.skip = ptr<int>(ctx->cskip);
    T.sub = csub;
    T.nkids = cnkids;
    // the groups need the child slots and the topology only: on the second
    // stream (after the level lists there) beside the level-by-level
    // moments, or on the main stream while the second computes the moments
    // from the prefix sums
    const bool groups = p.mode == BH_MODE_FAST;
    const bool groups_side = groups && !prefix_moments;
    if (groups_side) {
        side_fork(ctx, kForkKids);
        build_groups_side(ctx, p, &tt, ks, tks_for_groups);
        side_mark(ctx, kMarkGroups);
    } else if (groups) {
        build_groups(ctx, p, &tt, ks, tks_for_groups);
    }
    main_wait(ctx, kMarkGather);
    main_wait(ctx, kMarkLevels);
    const int mgrid = ctx->ncu * 4;
    // the learned features are written only for the learned walk and sampling
    const bool feat = p.learned != 0 || ctx->mlp_sampling;
    // the learned features carry the quadrupole (the leading error of a
    // monopole cell), so a learned or sampling build computes it whatever
    // the walk's cells are
    for (int L = prefix_moments ? -1 : kMaxLevel; L >= 0; L--) {
        // level L holds at most 8^L cells: the top levels need a block or two
        const int lgrid = L <= 3 ? nblocks((int64_t)1 << (3 * L)) : mgrid * 2;
        if (p.usequad || feat)
            k_moments_level<true><<<lgrid, kTpb, 0, st>>>(llist, loffd, L, ks, feat ? cF0 : nullptr, cF1, csub,
                                                              cfirst, clast, clevel, bd, mo, cA, cB,
                                                              p.usequad ? cQ0 : nullptr, cQ1, d_box, lsys,
                                                              ctx->sysbits, sysbox);
        else
            k_moments_level<false><<<lgrid, kTpb, 0, st>>>(llist, loffd, L, ks, feat ? cF0 : nullptr, cF1, csub,
                                                               cfirst, clast, clevel, bd, mo, cA, cB, cQ0, cQ1, d_box,
                                                               lsys, ctx->sysbits, sysbox);
    }
    T.rc2 = p.criterion != BH_CRIT_HACK ? cell_rcrit(ctx, p, ks, cfirst, clevel, mo, tt.ncellp, cap, d_box, cB)
                                         : nullptr;
    if (nsys)
        k_sysroot_cells<<<nblocks(cap), kTpb, 0, st>>>(clevel, cfirst, ks, lsys, ctx->sysbits, tt.ncellp,
                                                       ptr<int>(ctx->sysroot));
    BH_HIP(hipGetLastError());
    T.mo = mo;
    T.cA = cA;
    T.cB = cB;
    T.cQ0 = cQ0;
    T.cQ1 = cQ1;
    T.cF0 = cF0;
    T.cF1 = cF1;
    if (groups_side) main_wait(ctx, kMarkGroups);
}

}  // namespace bh
