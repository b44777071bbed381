// Engine context: device buffers, stream, timing events, last error.
#pragma once

#include "bh_internal.h"
#include "bh_mlp.h"

#include <vector>

namespace bh {
// Learned-correction state: parameters (FP32 master, BF16 MFMA copy, biases),
// optimiser moments, training samples and the per-group corrections.
struct Mlp {
    DevBuf params, wb, bias, gpart, lpart, adam_m, adam_v, grad, loss;
    DevBuf sx, sy, snew_x, snew_y, soff, corr, ecell, egrp, gstart, gcnt, gcentre, res;
    DevBuf smeta;               // per sample of the last collection: cell, group centre (tests)
    int64_t nsamples = 0;
    int64_t nlast = 0;          // samples added by the last collection
    int64_t nentries = -1;      // entries of the last evaluation if it was a learned one, else -1
    bool keep_res = false;      // store each entry's MLP output (res) in learned evaluations (tests)
    bool res_valid = false;     // res holds the last learned evaluation's outputs
    int64_t ecap = 0;           // capacity of the flattened correction entries
    float norm[kMlpNorm] = {1.f, 0.25f, 1.6f, 1.f};
    bool ready = false;
    int adam_step = 0;
};
}  // namespace bh

struct bh_ctx {
    int device = 0;
    int ncu = 256;
    hipStream_t stream = nullptr;
    // second stream of the tree build (same priority): branches that no
    // kernel of the main stream needs until the matching join
    hipStream_t side = nullptr;
    hipEvent_t fork_ev[6] = {};
    std::string err;
    int inject = 0;

    // uploaded problem
    int64_t nmass = 0;          // bodies in the mass set (input order)
    int64_t ntest = 0;          // test points; == nmass when the mass set is the test set
    bool self_test = true;      // test set is the mass set
    bool uploaded = false;
    bool box_valid = false;     // a previous box exists (keep_box)
    double box[4] = {0, 0, 0, 0};

    // inputs (device, input order)
    bh::DevBuf mass, pos, tpos;
    // box state: double4 box, ints {first, start, changed, nonfinite}
    bh::DevBuf dbox, dboxi, bcand, bcount;
    // sorting
    bh::DevBuf keys0, keys1, vals0, vals1, hist, scan_tmp;
    bh::DevBuf tkeys0, tkeys1, tvals0, tvals1;
    bh::DevBuf nmassive;        // int
    // sorted bodies of the tree and of the test set
    bh::DevBuf bd, bhi, blo, bpk;   // bpk: (x, y, z, m) records in input order
    bh::DevBuf thi, tlo, td, tperm, tself, tinv, tmap;
    // topology
    bh::DevBuf cpl, ccnt, cbase, cmin1, cmin2;
    bh::DevBuf cfirst, clast, clevel, csub, cnkids, cskip, pre, cowner, gowner;
    bh::DevBuf crc2;            // treecode1 Rcrit2 per cell (criterion != BH_CRIT_HACK)
    bh::DevBuf cmo, cfnb, cA, cB, cQ0, cQ1, lhist, llist, lkeys0, lkeys1, lvals0, lvals1;
    bh::DevBuf mpre, mpart;     // fast monopole trees: tile-local prefix sums of (m x, m y, m z, m), tile prefixes
    bh::DevBuf counters;        // u64: n2b, nbc, deep, n_mlp, err flags ...
    // groups of the test set (fast walk)
    bh::DevBuf gfirst, gcount, gscan, goff;
    bh::DevBuf gcpl, gccnt, gcbase, gm1, gm2, gcf, gcl, gclv;
    int ngroups = 0;          // host copy after an evaluation
    // walk scratch
    bh::DevBuf frontier, gctr, werr;
    bh::DevBuf shlist, shid;    // super-group walk: shared lists per wave
    int64_t fcap = 0;
    // outputs (input order of the test set)
    bh::DevBuf phi, acc;

    // leapfrog state (input order): velocities, previous accelerations,
    // device step counters {nstep, record}, per-step outputs, step graph
    bh::DevBuf vel, acc_old, lfstep, diag_partial, diag_buf, n2b_buf, nbc_buf;
    double lf_dt = 0.0;
    bh_params lfp{};
    bool lf_ready = false;
    int64_t lf_nstep = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    int64_t graph_cap = 0;         // steps the captured output buffers hold
    int graph_key[4] = {0, 0, 0, 0};

    bh::Mlp mlp;
    bh::DevBuf lrec;             // per-wave learned-correction records of the walk
    bh::DevBuf sacc;             // recording walks: (a, phi) per test point in sorted order
    int64_t lcap = 0;            // ints per wave region
    int lnwave = 0;              // waves (regions) of the last recording walk
    bh::DevBuf cF0, cF1;         // per-cell learned features
    // batched ensemble (nsys > 0): system s owns bodies [boff[s], boff[s+1])
    // of the mass set, in input order and in sorted order (its id is the top
    // sysbits bits of the Morton key); per-system box, (tol^2, eps^2) and the
    // node the walk of its groups starts from
    int nsys = 0;
    int sysbits = 0;
    bool sys_eps0 = false;      // some system has eps = 0
    bh::DevBuf boff, sysbox, sysp, sysroot, sysdiag;
    bh::DevBuf ptile, pseg;     // prefix-sum moments of a batch: (first, end) body per tile, first tile per system
    int nptile = 0;
    bh::Tree tree;
    int64_t nb = 0;             // massive bodies (counted at upload)
    int cell_cap = 0;           // cell capacity of the tree arrays
    bool mlp_sampling = false;  // a sample-collection walk is running (the tree writes learned features)
    int tuned_group = 128;      // defaults for group_size = 0 / nleaf = -1 / supergroup = 0 (set by bh_autotune)
    int tuned_nleaf = 2;
    int tuned_sg = 1;
    int tcell_cap = 0;          // cell capacity of a separate test-set topology
    int box_rounds = 6;         // expansion rounds from a fresh box
    int box_rounds_keep = 2;    // expansion rounds from the previous box
    bool capturing = false;     // a graph is being captured: no events, no host syncs
    int last_mode = 0;
    // 0-2 evaluation phases, 4-5 walk kernel, 6-7 sort, 8-9 MLP inference
    hipEvent_t ev[10] = {};
    float t_walk_kernel_ms = 0.f, t_sort_ms = 0.f;
    bh_stats last{};
};

namespace bh {

template <class T>
T *ensure(DevBuf &b, size_t count)
{
    size_t bytes = count * sizeof(T);
    if (bytes == 0) bytes = sizeof(T);
    if (b.cap < bytes) {
        if (b.p) BH_HIP(hipFree(b.p));
        b.p = nullptr;
        b.cap = 0;
        size_t want = bytes + bytes / 8;   // slack so slowly growing N does not reallocate each call
        hipError_t e = hipMalloc(&b.p, want);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            throw ApiFail(BH_ERR_NOMEM, "device allocation of " + std::to_string(want) + " bytes failed");
        }
        b.cap = want;
    }
    return static_cast<T *>(b.p);
}

template <class T>
T *ptr(DevBuf &b) { return static_cast<T *>(b.p); }

// bh_tree.hip
void build_tree(bh_ctx *ctx, const bh_params &p);
void exclusive_scan_u32(bh_ctx *ctx, const uint32_t *in, uint32_t *out, int64_t n,
                        const unsigned long long *dn = nullptr);
void radix_sort_pairs(bh_ctx *ctx, uint64_t *k0, uint64_t *k1, uint32_t *v0, uint32_t *v1, int64_t n,
                      int begin_bit, int end_bit, uint64_t **kout, uint32_t **vout);
// bh_exact.hip
void walk_strict(bh_ctx *ctx, const bh_params &p);
double *cell_rcrit(bh_ctx *ctx, const bh_params &p, const uint64_t *keys, const int *cfirst, const uint8_t *level,
                   const double *mo, const unsigned long long *ncellp, int cap, const double4 *box, float4 *cB);
void direct_sum(bh_ctx *ctx, double eps, int64_t nidx, const int64_t *d_idx, double *d_phi, double *d_acc);
// bh_walk.hip
void walk_fast(bh_ctx *ctx, const bh_params &p, int lmode = 0, int lstride = 1);
// bh_mlp.hip
void mlp_set(bh_ctx *ctx, const float *host_params, const float *norm);
void mlp_pack(bh_ctx *ctx);
void mlp_forward(bh_ctx *ctx, const float *d_x, int64_t n, float *d_y);
void mlp_grad(bh_ctx *ctx, const float *d_x, const float *d_y, int64_t n, float *d_grad, float *d_loss);
void mlp_adam(bh_ctx *ctx, const float *d_grad, float lr, int step);
void mlp_gather(bh_ctx *ctx, uint64_t a, uint64_t c, int64_t first, int64_t cnt, float *bx, float *by);
void mlp_correct(bh_ctx *ctx, const bh_params &p);
int64_t mlp_collect(bh_ctx *ctx, const bh_params &p);
// bh_lf.hip
void lf_reset(bh_ctx *ctx);
void lf_reset_record(bh_ctx *ctx);
void lf_after_force(bh_ctx *ctx, double *diag_out, long long *n2b_out, long long *nbc_out);
void lf_drift(bh_ctx *ctx);

}  // namespace bh
