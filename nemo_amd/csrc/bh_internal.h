// Internal declarations shared by the engine's HIP translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/bh.h"

namespace bh {

constexpr int kPreTile = 2048;         // bodies per tile of the prefix-sum moments (a system starts a tile)
constexpr int kMaxLevel = 21;           // 63-bit Morton keys: 21 levels below the root
constexpr int kWaveSize = 64;

// Device buffer that only grows.
struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
};

// Status carried through the host code; set by BH_TRY / fail().
struct Err {
    bh_status code = BH_OK;
    std::string msg;
};

// Arrays describing one built tree (all device pointers).
struct Tree {
    int nb = 0;                 // massive bodies loaded in the tree (sorted Morton order)
    int ncell = 0;              // cells; cell 0 is the root when nb >= 2
    double box[4] = {0, 0, 0, 0};   // rmin.xyz, rsize
    // bodies, sorted
    const double4 *bd = nullptr;    // x, y, z, m (FP64)
    const float4 *bhi = nullptr;    // FP32 hi part of x, y, z; w = mass
    const float4 *blo = nullptr;    // FP32 lo part (x - hi)
    // cells
    const int *first = nullptr, *last = nullptr;   // body range [first, last]
    const int2 *fnb = nullptr;      // (first body, body count) per cell
    const int *sub = nullptr;       // [ncell][8] child refs: >= 0 body, < 0 ~cell, -1... see kNoChild
    const int *nkids = nullptr;     // entries pushed when a cell is opened
    const uint8_t *level = nullptr;
    const int *skip = nullptr;      // pre-order position after the subtree
    const int *pre = nullptr;       // pre-order node refs (nb + ncell entries)
    const double *mo = nullptr;     // FP64 moments, 10 per cell: m, x, y, z, qxx qxy qxz qyy qyz qzz
    const float4 *cA = nullptr;     // FP32 com hi xyz, mass
    const float4 *cB = nullptr;     // FP32 com lo xyz, size^2 (treecode1 criteria: Rcrit2)
    const double *rc2 = nullptr;    // treecode1 criteria: FP64 Rcrit2 per cell
    const float4 *cQ0 = nullptr;    // qxx qxy qxz qyy
    const float4 *cQ1 = nullptr;    // qyz qzz 0 0
    // test set in its sorted order
    const float4 *thi = nullptr, *tlo = nullptr;
    const double4 *td = nullptr;
    const uint32_t *tperm = nullptr;   // sorted test index -> input index
    const int *tself = nullptr;        // selfmode 2: tree body of each test point or -1
    const int *tinv = nullptr;         // selfmode 2: test index of each tree body
    int selfmode = 0;                  // 0 no self pairs, 1 test index == tree body, 2 tself[] / tinv[]
    const uint64_t *keys = nullptr;    // sorted Morton keys of the mass set
    int cap = 0;                       // cell capacity of the arrays
    const unsigned long long *ncellp = nullptr;   // device cell count
    const double4 *dbox = nullptr;     // device root box (rmin.xyz, rsize)
    const float4 *cF0 = nullptr;       // learned features: (com - centre)/size xyz, Q_yz/(M size^2)
    const float4 *cF1 = nullptr;       // Q_xx, Q_yy, Q_xy, Q_xz over M size^2
};

constexpr int kNoChild = INT32_MIN;  // empty octant in Tree::sub

// Device status words (unsigned long long) written by the kernels and read
// back once at the end of an evaluation, so the evaluation itself has no
// host synchronisation (and can be captured in a graph).
enum StatSlot {
    S_N2B = 0, S_NBC = 1, S_DEEP = 2, S_NMLP = 3,
    S_NCELL = 4,        // cells of the tree (clamped to the capacity)
    S_NCELL_RAW = 5,    // cells needed
    S_NGROUPS = 6,
    S_FLAGS = 7,
    S_TNCELL = 8, S_TNCELL_RAW = 9,   // topology of a separate test set (grouping only)
    S_NLOC = 10,        // body-cell interactions evaluated through group local expansions
    S_ZERO = 11,        // slots below are cleared at the start of every evaluation
    S_STICKY = 11,      // flags or-ed over the steps of a leapfrog batch
    // the fast walk's group counter (low 32 bits; cleared at the start of
    // every tree build, one walk per build), on a cache line of its own: its
    // atomics would otherwise serialise the walk's reads of S_NGROUPS
    S_GCTR = 32,
    S_NSLOTS = 48
};
enum StatFlag : unsigned long long {
    F_CELL_OVF = 1, F_BOX_OPEN = 2, F_NONFINITE = 4, F_FRONTIER = 8, F_TCELL_OVF = 16, F_LREC_OVF = 32, F_ECAP_OVF = 64
};

}  // namespace bh

#define BH_HIP(call)                                                                        \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess) {                                                             \
            throw bh::HipFail(e_, #call, __FILE__, __LINE__);                              \
        }                                                                                   \
    } while (0)

namespace bh {
struct HipFail {
    hipError_t e;
    const char *what;
    const char *file;
    int line;
    HipFail(hipError_t e_, const char *w, const char *f, int l) : e(e_), what(w), file(f), line(l) {}
};
struct ApiFail {
    bh_status code;
    std::string msg;
    ApiFail(bh_status c, std::string m) : code(c), msg(std::move(m)) {}
};
}  // namespace bh
