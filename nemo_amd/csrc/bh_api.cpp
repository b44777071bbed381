// C ABI of the engine (include/bh.h): argument checking, device buffers,
// the force_calc sequence (tree build then walk) and error mapping.  No entry
// point exits the process; failures come back as bh_status + bh_last_error.
#include "bh_ctx.h"

#include <algorithm>
#include <cmath>
#include <cstring>

#include <mutex>

namespace bh {
void debug_keys(bh_ctx *ctx, uint64_t *d_out);
}

// Staging of bh_direct_noalloc: target indices (int64) then FP64 results
// (phi, then acc[3]) for at most BH_DIRECT_NOALLOC_MAX targets, 40 bytes each
// (1.25 MiB), one copy per device for the life of the process.
static __device__ __align__(256) unsigned char scratch[(size_t)BH_DIRECT_NOALLOC_MAX * 40];

namespace {

// Device address of each device's copy of `scratch`, looked up once per
// device (with that device current) by bh_create.  The host-side name of the
// array is not a device address and is never passed to the device.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch[kMaxDevices] = {};

hipError_t scratch_setup(int device)
{
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch[device]) return hipSuccess;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return e;
    }
    g_scratch[device] = p;
    return hipSuccess;
}

void *scratch_of(int device)
{
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return g_scratch[device];
}

}  // namespace

namespace {

using bh::ApiFail;
using bh::HipFail;

template <class F>
bh_status guarded(bh_ctx *ctx, F &&f)
{
    if (!ctx) return BH_ERR_ARG;
    try {
        if (ctx->inject) {
            int code = ctx->inject;
            ctx->inject = 0;
            throw ApiFail((bh_status)code, "injected failure");
        }
        BH_HIP(hipSetDevice(ctx->device));
        f();
        ctx->err.clear();
        return BH_OK;
    } catch (const HipFail &e) {
        char buf[512];
        snprintf(buf, sizeof buf, "HIP error %d (%s) in %s at %s:%d", (int)e.e, hipGetErrorString(e.e), e.what, e.file,
                 e.line);
        ctx->err = buf;
        (void)hipGetLastError();
        return BH_ERR_DEVICE;
    } catch (const ApiFail &e) {
        ctx->err = e.msg;
        return e.code;
    } catch (const std::exception &e) {
        ctx->err = e.what();
        return BH_ERR_NOMEM;
    }
}

void check_params(const bh_params *p)
{
    if (!p) throw ApiFail(BH_ERR_ARG, "params is NULL");
    if (!(p->tol >= 0.0) || !std::isfinite(p->tol)) throw ApiFail(BH_ERR_ARG, "tol must be finite and >= 0");
    if (!(p->eps >= 0.0) || !std::isfinite(p->eps)) throw ApiFail(BH_ERR_ARG, "eps must be finite and >= 0");
    if (!(p->rsize > 0.0) || !std::isfinite(p->rsize)) throw ApiFail(BH_ERR_ARG, "rsize must be finite and > 0");
    if (p->mode < BH_MODE_FAST || p->mode > BH_MODE_DIRECT) throw ApiFail(BH_ERR_ARG, "unknown mode");
    if (p->group_size != 0 && p->group_size != 64 && p->group_size != 128 && p->group_size != 192 &&
        p->group_size != 256)
        throw ApiFail(BH_ERR_ARG, "group_size must be 0, 64, 128, 192 or 256");
    if (p->rmin_given)
        for (int k = 0; k < 3; k++)
            if (!std::isfinite(p->rmin[k])) throw ApiFail(BH_ERR_ARG, "rmin must be finite");
    if (p->criterion < BH_CRIT_HACK || p->criterion > BH_CRIT_SW94) throw ApiFail(BH_ERR_ARG, "unknown criterion");
    if (p->supergroup < 0 || p->supergroup > 4) throw ApiFail(BH_ERR_ARG, "supergroup must be 0 (tuned), 1 (off) or 2..4");
    if (p->criterion != BH_CRIT_HACK && p->learned)
        throw ApiFail(BH_ERR_ARG, "the learned correction is trained for the hackcode1 criterion only");
    if (!(p->local_alpha >= 0.0 && p->local_alpha <= 0.5))
        throw ApiFail(BH_ERR_ARG, "local_alpha must lie in [0, 0.5]");
    if (p->local_alpha > 0.0 && p->mode == BH_MODE_FAST && p->usequad)
        throw ApiFail(BH_ERR_ARG, "local expansions take monopole cells");
}

void lf_drop_graph(bh_ctx *ctx)
{
    if (ctx->graph_exec) (void)hipGraphExecDestroy(ctx->graph_exec);
    if (ctx->graph) (void)hipGraphDestroy(ctx->graph);
    ctx->graph_exec = nullptr;
    ctx->graph = nullptr;
    ctx->graph_cap = 0;
}

// Device temporary of one API call, freed on every exit path.
struct TmpDev {
    void *p = nullptr;
    explicit TmpDev(size_t bytes)
    {
        if (hipMalloc(&p, bytes ? bytes : 1) != hipSuccess) {
            (void)hipGetLastError();
            p = nullptr;
            throw ApiFail(BH_ERR_NOMEM, "device allocation of " + std::to_string(bytes) + " bytes failed");
        }
    }
    ~TmpDev()
    {
        if (p) (void)hipFree(p);
    }
    TmpDev(const TmpDev &) = delete;
    TmpDev &operator=(const TmpDev &) = delete;
    template <class T>
    T *as() const { return static_cast<T *>(p); }
};

// A new upload replaces the bodies a leapfrog integrates: its state (vel,
// acc_old sized for the old set) and captured step graph are void, and the
// bh_lf_* calls report BH_ERR_STATE until bh_lf_init runs again.
void upload(bh_ctx *ctx, int64_t nmass, const double *mass, const double *pos, int64_t ntest, const double *tpos)
{
    if (nmass < 1) throw ApiFail(BH_ERR_ARG, "nmass must be >= 1");
    if (nmass > (int64_t)INT32_MAX / 2) throw ApiFail(BH_ERR_ARG, "nmass too large (max 2^30)");
    if (!mass || !pos) throw ApiFail(BH_ERR_ARG, "mass/pos is NULL");
    const bool self = (tpos == nullptr || ntest == 0);
    if (!self && (ntest < 0 || ntest > (int64_t)INT32_MAX / 2)) throw ApiFail(BH_ERR_ARG, "bad ntest");
    ctx->lf_ready = false;
    lf_drop_graph(ctx);
    ctx->nsys = 0;
    ctx->sysbits = 0;
    ctx->nmass = nmass;
    int64_t nb = 0;
    for (int64_t i = 0; i < nmass; i++) nb += mass[i] != 0.0;
    if (nb != ctx->nb) ctx->cell_cap = 0;
    ctx->nb = nb;
    ctx->self_test = self;
    ctx->ntest = self ? nmass : ntest;
    double *dm = bh::ensure<double>(ctx->mass, nmass);
    double *dp = bh::ensure<double>(ctx->pos, 3 * nmass);
    BH_HIP(hipMemcpyAsync(dm, mass, nmass * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    BH_HIP(hipMemcpyAsync(dp, pos, 3 * nmass * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    if (!self) {
        double *dt = bh::ensure<double>(ctx->tpos, 3 * ntest);
        BH_HIP(hipMemcpyAsync(dt, tpos, 3 * ntest * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    }
    bh::ensure<double>(ctx->phi, ctx->ntest);
    bh::ensure<double>(ctx->acc, 3 * ctx->ntest);
    BH_HIP(hipStreamSynchronize(ctx->stream));
    ctx->uploaded = true;
}

// One evaluation without host synchronisation, then one read-back of the
// device status.  A capacity overflow (cells, walk frontier) or a box that
// needed more expansion rounds grows the capacity and evaluates again.
void eval(bh_ctx *ctx, const bh_params *p, bh_stats *st)
{
    check_params(p);
    if (!ctx->uploaded) throw ApiFail(BH_ERR_STATE, "bh_eval before bh_upload");
    bh_params q = *p;
    if (q.group_size == 0) q.group_size = ctx->tuned_group;
    if (q.nleaf < 0) q.nleaf = ctx->tuned_nleaf;
    if (q.supergroup == 0) q.supergroup = ctx->tuned_sg;
    if (q.mode == BH_MODE_DIRECT) throw ApiFail(BH_ERR_ARG, "BH_MODE_DIRECT is evaluated with bh_direct");
    if (ctx->nsys && (q.mode != BH_MODE_FAST || q.learned))
        throw ApiFail(BH_ERR_ARG, "batches are evaluated by the fast walk without the learned correction");
    if (ctx->nsys && q.criterion != BH_CRIT_HACK)
        throw ApiFail(BH_ERR_ARG, "batches use the hackcode1 criterion");
    if (ctx->cell_cap == 0) ctx->cell_cap = (int)std::min<int64_t>(2 * ctx->nb + 64, INT32_MAX / 2);
    unsigned long long stat[bh::S_NSLOTS];
    double box[4];
    // plain fast walk: the tree-end event also starts the walk kernel and the
    // walk kernel's end event also ends the phase (fewer event records)
    const bool plain_fast = q.mode == BH_MODE_FAST && !q.learned;
    for (int attempt = 0;; attempt++) {
        BH_HIP(hipEventRecord(ctx->ev[0], ctx->stream));
        bh::build_tree(ctx, q);
        BH_HIP(hipEventRecord(ctx->ev[1], ctx->stream));
        if (q.mode == BH_MODE_STRICT) {
            bh::walk_strict(ctx, q);
        } else if (q.learned) {
            if (!ctx->mlp.ready && q.learned == 1)
                throw ApiFail(BH_ERR_STATE, "learned correction requested but no model is loaded");
            bh::walk_fast(ctx, q, 1, 1);
            bh::mlp_correct(ctx, q);
        } else {
            bh::walk_fast(ctx, q);
        }
        if (!plain_fast) BH_HIP(hipEventRecord(ctx->ev[2], ctx->stream));
        BH_HIP(hipMemcpyAsync(stat, ctx->counters.p, sizeof stat, hipMemcpyDeviceToHost, ctx->stream));
        BH_HIP(hipMemcpyAsync(box, ctx->dbox.p, sizeof box, hipMemcpyDeviceToHost, ctx->stream));
        BH_HIP(hipStreamSynchronize(ctx->stream));
        ctx->box_valid = true;
        const unsigned long long f = stat[bh::S_FLAGS];
        if (f & bh::F_NONFINITE)
            throw ApiFail(BH_ERR_NONFINITE, "non-finite position of a massive body (expandbox cannot fit it)");
        if (!(f & (bh::F_CELL_OVF | bh::F_TCELL_OVF | bh::F_BOX_OPEN | bh::F_FRONTIER | bh::F_LREC_OVF |
                   bh::F_ECAP_OVF)))
            break;
        if (f & bh::F_LREC_OVF) ctx->lcap *= 4;
        if (f & bh::F_ECAP_OVF) ctx->mlp.ecap *= 2;
        if (attempt >= 12) throw ApiFail(BH_ERR_CAPACITY, "evaluation did not converge (capacities)");
        if (f & bh::F_CELL_OVF)
            ctx->cell_cap = (int)std::min<uint64_t>(stat[bh::S_NCELL_RAW] + stat[bh::S_NCELL_RAW] / 4 + 64, INT32_MAX / 2);
        if (f & bh::F_TCELL_OVF)
            ctx->tcell_cap = (int)std::min<uint64_t>(stat[bh::S_TNCELL_RAW] + stat[bh::S_TNCELL_RAW] / 4 + 64, INT32_MAX / 2);
        if (f & bh::F_BOX_OPEN) {
            ctx->box_rounds *= 4;
            ctx->box_rounds_keep *= 4;
        }
        if (f & bh::F_FRONTIER) {
            ctx->fcap *= 4;
            if ((size_t)ctx->fcap * sizeof(int) * 2 * 4 * 8 * ctx->ncu > ((size_t)64 << 30))
                throw ApiFail(BH_ERR_CAPACITY, "walk frontier would exceed 64 GiB; use a larger tol");
        }
    }
    for (int k = 0; k < 4; k++) ctx->box[k] = box[k];
    ctx->tree.ncell = (int)stat[bh::S_NCELL];
    ctx->ngroups = (int)stat[bh::S_NGROUPS];
    bh_stats s{};
    s.n2btot = (int64_t)stat[bh::S_N2B];
    s.nbctot = (int64_t)stat[bh::S_NBC];
    s.deep_buckets = (int64_t)stat[bh::S_DEEP];
    s.n_mlp = (int64_t)stat[bh::S_NMLP];
    s.n_local = (int64_t)stat[bh::S_NLOC];
    s.ncell = ctx->tree.ncell;
    for (int k = 0; k < 3; k++) s.rmin[k] = box[k];
    s.rsize = box[3];
    float t1 = 0, t2 = 0, tw = 0, ts = 0;
    BH_HIP(hipEventElapsedTime(&t1, ctx->ev[0], ctx->ev[1]));
    BH_HIP(hipEventElapsedTime(&t2, ctx->ev[1], plain_fast ? ctx->ev[5] : ctx->ev[2]));
    BH_HIP(hipEventElapsedTime(&ts, ctx->ev[6], ctx->ev[7]));
    if (plain_fast) tw = t2;
    else if (q.mode == BH_MODE_FAST) BH_HIP(hipEventElapsedTime(&tw, ctx->ev[4], ctx->ev[5]));
    else tw = t2;
    s.t_tree_ms = t1;
    s.t_walk_ms = t2;
    s.t_walk_kernel_ms = tw;
    s.t_sort_ms = ts;
    s.ngroups = q.mode == BH_MODE_FAST ? ctx->ngroups : 0;
    if (q.mode == BH_MODE_FAST && q.learned == 1) {
        float tm = 0;
        BH_HIP(hipEventElapsedTime(&tm, ctx->ev[8], ctx->ev[9]));
        s.t_mlp_ms = tm;
    }
    ctx->last = s;
    ctx->last_mode = q.mode;
    ctx->mlp.nentries = (q.mode == BH_MODE_FAST && q.learned == 1) ? std::min<int64_t>(s.n_mlp, ctx->mlp.ecap) : -1;
    ctx->mlp.res_valid = ctx->mlp.nentries >= 0 && ctx->mlp.keep_res;
    if (st) *st = s;
}

// Batch upload: per-system offsets, eps and tol; every body massive.
void batch_upload(bh_ctx *ctx, int nsys, const int64_t *off, const double *mass, const double *pos, const double *eps,
                  const double *tol)
{
    if (nsys < 1 || nsys > (1 << 18)) throw ApiFail(BH_ERR_ARG, "nsys must be in [1, 2^18]");
    if (!off || !mass || !pos || !eps || !tol) throw ApiFail(BH_ERR_ARG, "NULL batch argument");
    if (off[0] != 0) throw ApiFail(BH_ERR_ARG, "off[0] must be 0");
    for (int s = 0; s < nsys; s++) {
        if (off[s + 1] <= off[s]) throw ApiFail(BH_ERR_ARG, "every system needs at least one body");
        if (!(eps[s] >= 0.0) || !std::isfinite(eps[s]) || !(tol[s] >= 0.0) || !std::isfinite(tol[s]))
            throw ApiFail(BH_ERR_ARG, "per-system eps and tol must be finite and >= 0");
    }
    const int64_t n = off[nsys];
    for (int64_t i = 0; i < n; i++)
        if (!(mass[i] != 0.0)) throw ApiFail(BH_ERR_ARG, "batch bodies must all be massive");
    upload(ctx, n, mass, pos, 0, nullptr);
    int bits = 0;
    while ((1 << bits) < nsys) bits += 3;
    ctx->sysbits = bits;
    std::vector<float> sp(2 * (size_t)nsys);
    bool eps0 = false;
    for (int s = 0; s < nsys; s++) {
        sp[2 * s] = (float)(tol[s] * tol[s]);
        sp[2 * s + 1] = (float)(eps[s] * eps[s]);
        eps0 = eps0 || eps[s] == 0.0;
    }
    int64_t *doff = bh::ensure<int64_t>(ctx->boff, nsys + 1);
    float *dsp = bh::ensure<float>(ctx->sysp, 2 * (size_t)nsys);
    BH_HIP(hipMemcpyAsync(doff, off, (nsys + 1) * sizeof(int64_t), hipMemcpyHostToDevice, ctx->stream));
    BH_HIP(hipMemcpyAsync(dsp, sp.data(), sp.size() * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
    bh::ensure<double4>(ctx->sysbox, nsys);
    bh::ensure<int>(ctx->sysroot, nsys);
    // prefix-sum moments: every system starts a tile, so its sums round as
    // in a run of that system alone
    std::vector<int2> tiles;
    std::vector<int> seg(nsys + 1);
    for (int sy = 0; sy < nsys; sy++) {
        seg[sy] = (int)tiles.size();
        for (int64_t b = off[sy]; b < off[sy + 1]; b += bh::kPreTile)
            tiles.push_back(make_int2((int)b, (int)std::min<int64_t>(b + bh::kPreTile, off[sy + 1])));
    }
    seg[nsys] = (int)tiles.size();
    int2 *dt = bh::ensure<int2>(ctx->ptile, tiles.size());
    int *dseg = bh::ensure<int>(ctx->pseg, nsys + 1);
    BH_HIP(hipMemcpyAsync(dt, tiles.data(), tiles.size() * sizeof(int2), hipMemcpyHostToDevice, ctx->stream));
    BH_HIP(hipMemcpyAsync(dseg, seg.data(), seg.size() * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
    ctx->nptile = (int)tiles.size();
    BH_HIP(hipStreamSynchronize(ctx->stream));
    ctx->nsys = nsys;
    ctx->sys_eps0 = eps0;
    ctx->box_valid = false;
}

void lf_upload_state(bh_ctx *ctx, int64_t n, const double *vel)
{
    double *dv = bh::ensure<double>(ctx->vel, 3 * n);
    double *da = bh::ensure<double>(ctx->acc_old, 3 * n);
    BH_HIP(hipMemcpyAsync(dv, vel, 3 * n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    BH_HIP(hipMemsetAsync(da, 0, 3 * n * sizeof(double), ctx->stream));
    bh::ensure<int>(ctx->lfstep, 2);
    bh::lf_reset(ctx);
    BH_HIP(hipStreamSynchronize(ctx->stream));
}

// Parameters of one leapfrog force pass: after the first step the box is kept.
bh_params lf_params(const bh_ctx *ctx)
{
    bh_params q = ctx->lfp;
    if (q.group_size == 0) q.group_size = ctx->tuned_group;
    if (q.nleaf < 0) q.nleaf = ctx->tuned_nleaf;
    if (q.supergroup == 0) q.supergroup = ctx->tuned_sg;
    q.keep_box = ctx->box_valid ? 1 : 0;
    return q;
}

}  // namespace

extern "C" {

void bh_default_params(bh_params *p)
{
    if (!p) return;
    memset(p, 0, sizeof *p);
    p->tol = 1.0;
    p->eps = 0.05;
    p->usequad = 0;
    p->mode = BH_MODE_FAST;
    p->rsize = 4.0;
    p->rmin_given = 0;
    p->group_size = 0;
    p->learned = 0;
    p->keep_box = 0;
    p->nleaf = -1;
    p->criterion = BH_CRIT_HACK;
    p->supergroup = 0;
    p->local_alpha = 0.0;
}

const char *bh_version(void) { return "nemo_amd-bh 0.1 (gfx950)"; }

bh_status bh_create(bh_ctx **out, int device)
{
    if (!out) return BH_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        (void)hipGetLastError();
        return BH_ERR_DEVICE;
    }
    if (device < 0 || device >= ndev) return BH_ERR_ARG;
    bh_ctx *ctx = new bh_ctx();
    ctx->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return BH_ERR_DEVICE;
    }
    if (hipStreamCreateWithFlags(&ctx->side, hipStreamNonBlocking) != hipSuccess) {
        (void)hipStreamDestroy(ctx->stream);
        delete ctx;
        return BH_ERR_DEVICE;
    }
    for (auto &e : ctx->fork_ev) (void)hipEventCreateWithFlags(&e, hipEventDisableTiming);
    // the device's copy of the bh_direct_noalloc staging array (failure is
    // reported by that entry point; the others do not need it)
    (void)scratch_setup(device);
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
        ctx->ncu = prop.multiProcessorCount;
    for (auto &e : ctx->ev) (void)hipEventCreate(&e);
    *out = ctx;
    return BH_OK;
}

void bh_destroy(bh_ctx *ctx)
{
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipStreamSynchronize(ctx->side);
    lf_drop_graph(ctx);
    bh::DevBuf *bufs[] = {&ctx->mass,   &ctx->pos,    &ctx->tpos,   &ctx->dbox,    &ctx->dboxi,  &ctx->keys0,
                          &ctx->keys1,  &ctx->vals0,  &ctx->vals1,  &ctx->hist,    &ctx->scan_tmp, &ctx->tkeys0,
                          &ctx->tkeys1, &ctx->tvals0, &ctx->tvals1, &ctx->nmassive, &ctx->bd,    &ctx->bpk,   &ctx->bhi,
                          &ctx->blo,    &ctx->thi,    &ctx->tlo,    &ctx->td,      &ctx->tperm,  &ctx->tself, &ctx->tinv, &ctx->tmap,
                          &ctx->cpl,    &ctx->ccnt,   &ctx->cbase,  &ctx->cmin1,   &ctx->cmin2,  &ctx->cfirst,
                          &ctx->clast,  &ctx->clevel, &ctx->csub,   &ctx->cnkids,  &ctx->crc2,
                          &ctx->cskip,  &ctx->pre,    &ctx->cmo,    &ctx->mpre,    &ctx->mpart,  &ctx->ptile, &ctx->pseg,
                          &ctx->cA,     &ctx->cB,     &ctx->cQ0,    &ctx->cQ1,     &ctx->counters, &ctx->frontier,
                          &ctx->gctr,   &ctx->werr, &ctx->shlist, &ctx->shid,   &ctx->phi,    &ctx->acc,    &ctx->gfirst, &ctx->gcount,
                          &ctx->gscan,  &ctx->goff,   &ctx->gcpl,   &ctx->gccnt,   &ctx->gcbase, &ctx->gm1,
                          &ctx->gm2,    &ctx->gcf,    &ctx->gcl,    &ctx->gclv,   &ctx->lhist,  &ctx->llist,  &ctx->bcand,  &ctx->bcount, &ctx->lkeys0, &ctx->lkeys1,
                          &ctx->lvals0, &ctx->lvals1, &ctx->cfnb, &ctx->cowner, &ctx->gowner,
                          &ctx->vel,    &ctx->acc_old, &ctx->lfstep, &ctx->diag_partial, &ctx->diag_buf,
                          &ctx->n2b_buf, &ctx->nbc_buf, &ctx->lrec, &ctx->sacc,   &ctx->cF0,    &ctx->cF1,
                          &ctx->mlp.params, &ctx->mlp.wb, &ctx->mlp.bias, &ctx->mlp.gpart, &ctx->mlp.lpart,
                          &ctx->mlp.adam_m, &ctx->mlp.adam_v, &ctx->mlp.grad, &ctx->mlp.loss, &ctx->mlp.sx,
                          &ctx->mlp.sy, &ctx->mlp.snew_x, &ctx->mlp.snew_y, &ctx->mlp.soff, &ctx->mlp.corr,
                          &ctx->mlp.ecell, &ctx->mlp.egrp, &ctx->mlp.gstart, &ctx->mlp.gcnt, &ctx->mlp.gcentre,
                          &ctx->mlp.res, &ctx->mlp.smeta};
    for (auto *b : bufs)
        if (b->p) (void)hipFree(b->p);
    for (auto &e : ctx->ev)
        if (e) (void)hipEventDestroy(e);
    for (auto &e : ctx->fork_ev)
        if (e) (void)hipEventDestroy(e);
    if (ctx->side) (void)hipStreamDestroy(ctx->side);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char *bh_last_error(const bh_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

void bh_debug_keep_learned(bh_ctx *ctx, int on)
{
    if (ctx) ctx->mlp.keep_res = on != 0;
}

void bh_debug_inject(bh_ctx *ctx, int code)
{
    if (ctx) ctx->inject = code;
}

bh_status bh_upload(bh_ctx *ctx, int64_t nmass, const double *mass, const double *pos, int64_t ntest,
                    const double *tpos)
{
    return guarded(ctx, [&] {
        upload(ctx, nmass, mass, pos, ntest, tpos);
        ctx->box_valid = false;
    });
}

bh_status bh_eval(bh_ctx *ctx, const bh_params *p, bh_stats *st)
{
    return guarded(ctx, [&] { eval(ctx, p, st); });
}

bh_status bh_download(bh_ctx *ctx, double *phi, double *acc)
{
    return guarded(ctx, [&] {
        if (!ctx->uploaded) throw ApiFail(BH_ERR_STATE, "bh_download before bh_upload");
        if (phi) BH_HIP(hipMemcpy(phi, ctx->phi.p, ctx->ntest * sizeof(double), hipMemcpyDeviceToHost));
        if (acc) BH_HIP(hipMemcpy(acc, ctx->acc.p, 3 * ctx->ntest * sizeof(double), hipMemcpyDeviceToHost));
    });
}

bh_status bh_sync(bh_ctx *ctx)
{
    return guarded(ctx, [&] { BH_HIP(hipStreamSynchronize(ctx->stream)); });
}

bh_status bh_force(bh_ctx *ctx, const bh_params *p, int64_t nmass, const double *mass, const double *pos,
                   int64_t ntest, const double *tpos, double *phi, double *acc, bh_stats *st)
{
    return guarded(ctx, [&] {
        check_params(p);
        if (!phi || !acc) throw ApiFail(BH_ERR_ARG, "phi/acc is NULL");
        upload(ctx, nmass, mass, pos, ntest, tpos);
        ctx->box_valid = false;
        bh_params q = *p;
        q.keep_box = 0;
        eval(ctx, &q, st);
        BH_HIP(hipMemcpy(phi, ctx->phi.p, ctx->ntest * sizeof(double), hipMemcpyDeviceToHost));
        BH_HIP(hipMemcpy(acc, ctx->acc.p, 3 * ctx->ntest * sizeof(double), hipMemcpyDeviceToHost));
    });
}

bh_status bh_direct(bh_ctx *ctx, double eps, int64_t nidx, const int64_t *idx, double *phi, double *acc)
{
    return guarded(ctx, [&] {
        if (!ctx->uploaded) throw ApiFail(BH_ERR_STATE, "bh_direct before bh_upload");
        if (nidx < 1 || !idx || !phi || !acc) throw ApiFail(BH_ERR_ARG, "bad direct-sum arguments");
        if (!(eps >= 0.0)) throw ApiFail(BH_ERR_ARG, "eps must be >= 0");
        for (int64_t i = 0; i < nidx; i++)
            if (idx[i] < 0 || idx[i] >= ctx->nmass) throw ApiFail(BH_ERR_ARG, "direct-sum index out of range");
        TmpDev t_idx(nidx * sizeof(int64_t)), t_out(4 * nidx * sizeof(double));
        int64_t *d_idx = t_idx.as<int64_t>();
        double *d_out = t_out.as<double>();
        BH_HIP(hipMemcpyAsync(d_idx, idx, nidx * sizeof(int64_t), hipMemcpyHostToDevice, ctx->stream));
        bh::direct_sum(ctx, eps, nidx, d_idx, d_out, d_out + nidx);
        BH_HIP(hipMemcpyAsync(phi, d_out, nidx * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        BH_HIP(hipMemcpyAsync(acc, d_out + nidx, 3 * nidx * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        BH_HIP(hipStreamSynchronize(ctx->stream));
    });
}

bh_status bh_direct_noalloc(bh_ctx *ctx, double eps, int64_t nidx, const int64_t *idx, double *phi, double *acc)
{
    return guarded(ctx, [&] {
        if (!ctx->uploaded) throw ApiFail(BH_ERR_STATE, "bh_direct_noalloc before bh_upload");
        if (nidx < 1 || !idx || !phi || !acc) throw ApiFail(BH_ERR_ARG, "bad direct-sum arguments");
        if (nidx > BH_DIRECT_NOALLOC_MAX)
            throw ApiFail(BH_ERR_ARG, "bh_direct_noalloc takes at most " + std::to_string(BH_DIRECT_NOALLOC_MAX) +
                                          " targets (got " + std::to_string(nidx) + "); use bh_direct");
        if (!(eps >= 0.0)) throw ApiFail(BH_ERR_ARG, "eps must be >= 0");
        for (int64_t i = 0; i < nidx; i++)
            if (idx[i] < 0 || idx[i] >= ctx->nmass) throw ApiFail(BH_ERR_ARG, "direct-sum index out of range");
        unsigned char *base = static_cast<unsigned char *>(scratch_of(ctx->device));
        if (!base) throw ApiFail(BH_ERR_DEVICE, "no device address for the staging array of this device");
        // [idx: nidx int64][phi: nidx][acc: 3 nidx], all written before they are read
        int64_t *d_idx = reinterpret_cast<int64_t *>(base);
        double *d_out = reinterpret_cast<double *>(base + (size_t)nidx * sizeof(int64_t));
        BH_HIP(hipMemcpyAsync(d_idx, idx, nidx * sizeof(int64_t), hipMemcpyHostToDevice, ctx->stream));
        bh::direct_sum(ctx, eps, nidx, d_idx, d_out, d_out + nidx);
        BH_HIP(hipMemcpyAsync(phi, d_out, nidx * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        BH_HIP(hipMemcpyAsync(acc, d_out + nidx, 3 * nidx * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        BH_HIP(hipStreamSynchronize(ctx->stream));
    });
}

bh_status bh_debug_keys(bh_ctx *ctx, uint64_t *keys)
{
    return guarded(ctx, [&] {
        if (!ctx->uploaded || !ctx->box_valid) throw ApiFail(BH_ERR_STATE, "no tree built");
        if (!keys) throw ApiFail(BH_ERR_ARG, "keys is NULL");
        TmpDev t(ctx->nmass * sizeof(uint64_t));
        bh::debug_keys(ctx, t.as<uint64_t>());
        BH_HIP(hipMemcpy(keys, t.p, ctx->nmass * sizeof(uint64_t), hipMemcpyDeviceToHost));
    });
}

bh_status bh_debug_cells(bh_ctx *ctx, int64_t cap, int *level, int *nbody, double *mass, double *com, double *quad6,
                         int64_t *ncell)
{
    return guarded(ctx, [&] {
        if (!ctx->box_valid) throw ApiFail(BH_ERR_STATE, "no tree built");
        const int64_t nc = std::min<int64_t>(ctx->tree.ncell, ctx->tree.cap);
        if (ncell) *ncell = nc;
        int64_t k = std::min<int64_t>(cap, nc);
        if (k <= 0) return;
        std::vector<int> f(k), l(k);
        std::vector<uint8_t> lv(k);
        BH_HIP(hipMemcpy(f.data(), ctx->tree.first, k * sizeof(int), hipMemcpyDeviceToHost));
        BH_HIP(hipMemcpy(l.data(), ctx->tree.last, k * sizeof(int), hipMemcpyDeviceToHost));
        BH_HIP(hipMemcpy(lv.data(), ctx->tree.level, k, hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < k; i++) {
            if (level) level[i] = lv[i];
            if (nbody) nbody[i] = l[i] - f[i] + 1;
        }
        std::vector<double> mo(10 * k);
        BH_HIP(hipMemcpy(mo.data(), ctx->tree.mo, 10 * k * sizeof(double), hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < k; i++) {
            if (mass) mass[i] = mo[10 * i];
            for (int a = 0; a < 3; a++)
                if (com) com[3 * i + a] = mo[10 * i + 1 + a];
            for (int a = 0; a < 6; a++)
                if (quad6) quad6[6 * i + a] = mo[10 * i + 4 + a];
        }
    });
}

bh_status bh_debug_perm(bh_ctx *ctx, uint32_t *perm)
{
    return guarded(ctx, [&] {
        if (!ctx->box_valid || !perm) throw ApiFail(BH_ERR_STATE, "no tree built");
        BH_HIP(hipMemcpy(perm, ctx->tree.tperm, ctx->ntest * sizeof(uint32_t), hipMemcpyDeviceToHost));
    });
}

bh_status bh_debug_groups(bh_ctx *ctx, int64_t cap, int *first, int *count, int64_t *ngroups)
{
    return guarded(ctx, [&] {
        if (!ctx->box_valid) throw ApiFail(BH_ERR_STATE, "no tree built");
        if (ngroups) *ngroups = ctx->ngroups;
        int64_t k = std::min<int64_t>(cap, ctx->ngroups);
        if (k <= 0) return;
        if (first) BH_HIP(hipMemcpy(first, ctx->gfirst.p, k * sizeof(int), hipMemcpyDeviceToHost));
        if (count) BH_HIP(hipMemcpy(count, ctx->gcount.p, k * sizeof(int), hipMemcpyDeviceToHost));
    });
}

bh_status bh_lf_init(bh_ctx *ctx, const bh_params *p, int64_t n, const double *mass, const double *pos,
                     const double *vel, double dt)
{
    return guarded(ctx, [&] {
        check_params(p);
        if (!vel) throw ApiFail(BH_ERR_ARG, "vel is NULL");
        if (!(dt > 0.0) || !std::isfinite(dt)) throw ApiFail(BH_ERR_ARG, "dt must be finite and > 0");
        if (p->mode == BH_MODE_DIRECT) throw ApiFail(BH_ERR_ARG, "leapfrog needs the fast or strict walk");
        lf_drop_graph(ctx);
        upload(ctx, n, mass, pos, 0, nullptr);
        ctx->box_valid = false;
        lf_upload_state(ctx, n, vel);
        ctx->lfp = *p;
        ctx->lf_dt = dt;
        ctx->lf_nstep = 0;
        ctx->lf_ready = true;
        bh::ensure<double>(ctx->diag_buf, 36);
        bh::ensure<long long>(ctx->n2b_buf, 1);
        bh::ensure<long long>(ctx->nbc_buf, 1);
    });
}

bh_status bh_lf_force(bh_ctx *ctx, double *diag, int64_t *n2b, int64_t *nbc)
{
    return guarded(ctx, [&] {
        if (!ctx->lf_ready) throw ApiFail(BH_ERR_STATE, "bh_lf_force before bh_lf_init");
        bh_params q = lf_params(ctx);
        bh::lf_reset_record(ctx);
        eval(ctx, &q, nullptr);
        const int nd = 36 * std::max(1, ctx->nsys);
        double *dd = bh::ensure<double>(ctx->diag_buf, nd);
        long long *d2 = bh::ensure<long long>(ctx->n2b_buf, 1), *dc = bh::ensure<long long>(ctx->nbc_buf, 1);
        bh::lf_after_force(ctx, dd, d2, dc);
        std::vector<double> h(nd);
        long long c[2];
        BH_HIP(hipMemcpyAsync(h.data(), dd, nd * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        BH_HIP(hipMemcpyAsync(&c[0], d2, sizeof(long long), hipMemcpyDeviceToHost, ctx->stream));
        BH_HIP(hipMemcpyAsync(&c[1], dc, sizeof(long long), hipMemcpyDeviceToHost, ctx->stream));
        BH_HIP(hipStreamSynchronize(ctx->stream));
        if (diag) memcpy(diag, h.data(), nd * sizeof(double));
        if (n2b) *n2b = c[0];
        if (nbc) *nbc = c[1];
    });
}

bh_status bh_lf_drift(bh_ctx *ctx)
{
    return guarded(ctx, [&] {
        if (!ctx->lf_ready) throw ApiFail(BH_ERR_STATE, "bh_lf_drift before bh_lf_init");
        bh::lf_drift(ctx);
        BH_HIP(hipStreamSynchronize(ctx->stream));
        ctx->lf_nstep++;
    });
}

bh_status bh_lf_run(bh_ctx *ctx, int64_t nsteps, int use_graph, double *diag, int64_t *n2b, int64_t *nbc)
{
    return guarded(ctx, [&] {
        if (!ctx->lf_ready) throw ApiFail(BH_ERR_STATE, "bh_lf_run before bh_lf_init");
        if (nsteps < 0) throw ApiFail(BH_ERR_ARG, "nsteps < 0");
        if (nsteps == 0) return;
        hipStream_t st = ctx->stream;
        const int64_t cap = std::max<int64_t>(nsteps, ctx->graph_cap);
        const int64_t nd = 36 * (int64_t)std::max(1, ctx->nsys);
        double *dd = bh::ensure<double>(ctx->diag_buf, nd * (size_t)cap);
        long long *d2 = bh::ensure<long long>(ctx->n2b_buf, cap), *dc = bh::ensure<long long>(ctx->nbc_buf, cap);
        bh::lf_reset_record(ctx);
        unsigned long long *stat = bh::ptr<unsigned long long>(ctx->counters);
        int64_t done = 0;
        // one eager step first when the box or the capacities are not settled
        if (!ctx->box_valid || !use_graph) {
            bh_params q = lf_params(ctx);
            eval(ctx, &q, nullptr);
            bh::lf_after_force(ctx, dd, d2, dc);
            bh::lf_drift(ctx);
            done = 1;
        }
        bh_params q = lf_params(ctx);
        q.keep_box = 1;
        if (use_graph) {
            const int key[4] = {ctx->cell_cap, ctx->tcell_cap, (int)ctx->fcap, (int)(ctx->box_rounds_keep)};
            bool stale = ctx->graph_exec == nullptr || ctx->graph_cap < nsteps || memcmp(key, ctx->graph_key, sizeof key);
            if (stale) {
                lf_drop_graph(ctx);
                BH_HIP(hipStreamSynchronize(st));
                BH_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
                ctx->capturing = true;
                try {
                    bh::build_tree(ctx, q);
                    if (q.mode == BH_MODE_STRICT) bh::walk_strict(ctx, q);
                    else bh::walk_fast(ctx, q);
                    bh::lf_after_force(ctx, dd, d2, dc);
                    bh::lf_drift(ctx);
                } catch (...) {
                    ctx->capturing = false;
                    hipGraph_t g = nullptr;
                    (void)hipStreamEndCapture(st, &g);
                    if (g) (void)hipGraphDestroy(g);
                    throw;
                }
                ctx->capturing = false;
                BH_HIP(hipStreamEndCapture(st, &ctx->graph));
                BH_HIP(hipGraphInstantiate(&ctx->graph_exec, ctx->graph, nullptr, nullptr, 0));
                ctx->graph_cap = cap;
                memcpy(ctx->graph_key, key, sizeof key);
            }
            for (int64_t s = done; s < nsteps; s++) BH_HIP(hipGraphLaunch(ctx->graph_exec, st));
        } else {
            for (int64_t s = done; s < nsteps; s++) {
                bh::build_tree(ctx, q);
                if (q.mode == BH_MODE_STRICT) bh::walk_strict(ctx, q);
                else bh::walk_fast(ctx, q);
                bh::lf_after_force(ctx, dd, d2, dc);
                bh::lf_drift(ctx);
            }
        }
        unsigned long long h[bh::S_NSLOTS];
        double box[4];
        BH_HIP(hipMemcpyAsync(h, stat, sizeof h, hipMemcpyDeviceToHost, st));
        BH_HIP(hipMemcpyAsync(box, ctx->dbox.p, sizeof box, hipMemcpyDeviceToHost, st));
        if (diag) BH_HIP(hipMemcpyAsync(diag, dd, nd * nsteps * sizeof(double), hipMemcpyDeviceToHost, st));
        if (n2b) BH_HIP(hipMemcpyAsync(n2b, d2, nsteps * sizeof(int64_t), hipMemcpyDeviceToHost, st));
        if (nbc) BH_HIP(hipMemcpyAsync(nbc, dc, nsteps * sizeof(int64_t), hipMemcpyDeviceToHost, st));
        BH_HIP(hipStreamSynchronize(st));
        for (int k = 0; k < 4; k++) ctx->box[k] = box[k];
        ctx->lf_nstep += nsteps;
        const unsigned long long f = h[bh::S_STICKY] | h[bh::S_FLAGS];
        if (f & bh::F_NONFINITE) throw ApiFail(BH_ERR_NONFINITE, "non-finite position of a massive body during the run");
        if (f & (bh::F_CELL_OVF | bh::F_TCELL_OVF | bh::F_BOX_OPEN | bh::F_FRONTIER)) {
            lf_drop_graph(ctx);
            throw ApiFail(BH_ERR_CAPACITY, "a step of the batch exceeded a capacity (cells, frontier or box rounds); "
                                           "the state is invalid: restart with bh_lf_init");
        }
    });
}

bh_status bh_lf_get(bh_ctx *ctx, double *pos, double *vel, double *acc, double *phi, int64_t *nstep)
{
    return guarded(ctx, [&] {
        if (!ctx->lf_ready) throw ApiFail(BH_ERR_STATE, "bh_lf_get before bh_lf_init");
        const int64_t n = ctx->nmass;
        if (pos) BH_HIP(hipMemcpy(pos, ctx->pos.p, 3 * n * sizeof(double), hipMemcpyDeviceToHost));
        if (vel) BH_HIP(hipMemcpy(vel, ctx->vel.p, 3 * n * sizeof(double), hipMemcpyDeviceToHost));
        if (acc) BH_HIP(hipMemcpy(acc, ctx->acc.p, 3 * n * sizeof(double), hipMemcpyDeviceToHost));
        if (phi) BH_HIP(hipMemcpy(phi, ctx->phi.p, n * sizeof(double), hipMemcpyDeviceToHost));
        if (nstep) *nstep = ctx->lf_nstep;
    });
}

// ---- batched ensembles and the autotuner ------------------------------------------

bh_status bh_batch_upload(bh_ctx *ctx, int nsys, const int64_t *off, const double *mass, const double *pos,
                          const double *eps, const double *tol)
{
    return guarded(ctx, [&] { batch_upload(ctx, nsys, off, mass, pos, eps, tol); });
}

bh_status bh_batch_force(bh_ctx *ctx, const bh_params *p, int nsys, const int64_t *off, const double *mass,
                         const double *pos, const double *eps, const double *tol, double *phi, double *acc,
                         bh_stats *st)
{
    return guarded(ctx, [&] {
        check_params(p);
        if (!phi || !acc) throw ApiFail(BH_ERR_ARG, "phi/acc is NULL");
        batch_upload(ctx, nsys, off, mass, pos, eps, tol);
        bh_params q = *p;
        q.keep_box = 0;
        eval(ctx, &q, st);
        BH_HIP(hipMemcpy(phi, ctx->phi.p, ctx->ntest * sizeof(double), hipMemcpyDeviceToHost));
        BH_HIP(hipMemcpy(acc, ctx->acc.p, 3 * ctx->ntest * sizeof(double), hipMemcpyDeviceToHost));
    });
}

bh_status bh_batch_lf_init(bh_ctx *ctx, const bh_params *p, int nsys, const int64_t *off, const double *mass,
                           const double *pos, const double *vel, const double *eps, const double *tol, double dt)
{
    return guarded(ctx, [&] {
        check_params(p);
        if (!vel) throw ApiFail(BH_ERR_ARG, "vel is NULL");
        if (!(dt > 0.0) || !std::isfinite(dt)) throw ApiFail(BH_ERR_ARG, "dt must be finite and > 0");
        if (p->mode != BH_MODE_FAST || p->learned) throw ApiFail(BH_ERR_ARG, "batches use the fast walk");
        if (p->criterion != BH_CRIT_HACK) throw ApiFail(BH_ERR_ARG, "batches use the hackcode1 criterion");
        lf_drop_graph(ctx);
        batch_upload(ctx, nsys, off, mass, pos, eps, tol);
        lf_upload_state(ctx, off[nsys], vel);
        ctx->lfp = *p;
        ctx->lf_dt = dt;
        ctx->lf_nstep = 0;
        ctx->lf_ready = true;
    });
}

bh_status bh_autotune(bh_ctx *ctx, const bh_params *p, int reps, bh_params *best, double *ms)
{
    return guarded(ctx, [&] {
        check_params(p);
        if (!ctx->uploaded) throw ApiFail(BH_ERR_STATE, "bh_autotune before an upload");
        if (p->mode != BH_MODE_FAST) throw ApiFail(BH_ERR_ARG, "the autotuner tunes the fast walk");
        static const int groups[4] = {64, 128, 192, 256}, leaves[5] = {0, 1, 2, 4, 8};
        static_assert(4 * 5 == BH_AUTOTUNE_NCAND, "candidate count");
        double tbest = 1e300;
        bh_params qbest = *p;
        int k = 0;
        for (int g : groups)
            for (int nl : leaves) {
                bh_params q = *p;
                q.group_size = g;
                q.nleaf = nl;
                q.keep_box = 0;
                bh_stats st{};
                eval(ctx, &q, &st);      // warm-up; capacities settle here
                double t = 1e300;
                for (int r = 0; r < std::max(1, reps); r++) {
                    eval(ctx, &q, &st);
                    t = std::min(t, st.t_tree_ms + st.t_walk_ms);
                }
                if (ms) ms[k] = t;
                k++;
                if (t < tbest) {
                    tbest = t;
                    qbest = q;
                }
            }
        // then super-groups for the chosen group and leaf size (single
        // systems, hackcode1 criterion, no local expansions: the only walks
        // that use them)
        qbest.supergroup = 1;
        if (ctx->nsys == 0 && p->criterion == BH_CRIT_HACK && p->local_alpha == 0.0) {
            for (int K : {2, 3, 4}) {
                bh_params q = qbest;
                q.supergroup = K;
                bh_stats st{};
                eval(ctx, &q, &st);
                double t = 1e300;
                for (int r = 0; r < std::max(1, reps); r++) {
                    eval(ctx, &q, &st);
                    t = std::min(t, st.t_tree_ms + st.t_walk_ms);
                }
                if (t < tbest) {
                    tbest = t;
                    qbest = q;
                }
            }
        }
        ctx->tuned_group = qbest.group_size;
        ctx->tuned_nleaf = qbest.nleaf;
        ctx->tuned_sg = qbest.supergroup;
        if (best) *best = qbest;
    });
}

bh_status bh_lf_box(bh_ctx *ctx, double *rmin, double *rsize)
{
    return guarded(ctx, [&] {
        if (!ctx->lf_ready) throw ApiFail(BH_ERR_STATE, "bh_lf_box before bh_lf_init");
        if (!ctx->box_valid) throw ApiFail(BH_ERR_STATE, "no force pass yet");
        if (rmin)
            for (int k = 0; k < 3; k++) rmin[k] = ctx->box[k];
        if (rsize) *rsize = ctx->box[3];
    });
}

bh_status bh_lf_restore(bh_ctx *ctx, const double *acc, int64_t nstep, const double *rmin, double rsize)
{
    return guarded(ctx, [&] {
        if (!ctx->lf_ready) throw ApiFail(BH_ERR_STATE, "bh_lf_restore before bh_lf_init");
        if (ctx->nsys) throw ApiFail(BH_ERR_ARG, "bh_lf_restore handles single systems");
        if (nstep < 0 || nstep > INT32_MAX) throw ApiFail(BH_ERR_ARG, "bad nstep");
        if (rmin && !(rsize > 0.0 && std::isfinite(rsize) && std::isfinite(rmin[0]) && std::isfinite(rmin[1]) &&
                      std::isfinite(rmin[2])))
            throw ApiFail(BH_ERR_ARG, "bad box");
        lf_drop_graph(ctx);
        const int64_t n = ctx->nmass;
        if (acc)
            BH_HIP(hipMemcpy(ctx->acc_old.p, acc, 3 * n * sizeof(double), hipMemcpyHostToDevice));
        const int ns = (int)nstep;
        BH_HIP(hipMemcpy(ctx->lfstep.p, &ns, sizeof ns, hipMemcpyHostToDevice));
        ctx->lf_nstep = nstep;
        if (rmin) {
            const double b[4] = {rmin[0], rmin[1], rmin[2], rsize};
            BH_HIP(hipMemcpy(bh::ensure<double4>(ctx->dbox, 1), b, sizeof b, hipMemcpyHostToDevice));
            for (int k = 0; k < 4; k++) ctx->box[k] = b[k];
            ctx->box_valid = true;
        }
    });
}

// ---- learned correction ----------------------------------------------------------

bh_status bh_mlp_init(bh_ctx *ctx, uint64_t seed)
{
    return guarded(ctx, [&] {
        std::vector<float> P(bh::kMlpParams, 0.f);
        uint64_t st = seed * 0x9e3779b97f4a7c15ULL + 1;
        auto u = [&]() {
            st ^= st << 13;
            st ^= st >> 7;
            st ^= st << 17;
            return (double)(st >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
        };
        auto fill = [&](int off, int rows, int cols) {
            const double lim = std::sqrt(6.0 / cols);
            for (int i = 0; i < rows * cols; i++) P[off + i] = (float)(lim * u());
        };
        fill(bh::kOffW1, bh::kMlpHid, bh::kMlpIn);
        fill(bh::kOffW2, bh::kMlpHid, bh::kMlpHid);
        fill(bh::kOffW3, bh::kMlpHid, bh::kMlpHid);
        fill(bh::kOffW4, bh::kMlpOut, bh::kMlpHid);
        for (int i = 0; i < bh::kMlpOut * bh::kMlpHid; i++) P[bh::kOffW4 + i] *= 0.1f;
        bh::mlp_set(ctx, P.data(), nullptr);
        ctx->mlp.adam_step = 0;
    });
}

bh_status bh_mlp_set(bh_ctx *ctx, const float *params, const float *norm)
{
    return guarded(ctx, [&] {
        if (!params) throw ApiFail(BH_ERR_ARG, "params is NULL");
        bh::mlp_set(ctx, params, norm);
        ctx->mlp.adam_step = 0;
    });
}

bh_status bh_mlp_get(bh_ctx *ctx, float *params, float *norm)
{
    return guarded(ctx, [&] {
        if (!ctx->mlp.ready) throw ApiFail(BH_ERR_STATE, "no model");
        if (params)
            BH_HIP(hipMemcpy(params, ctx->mlp.params.p, bh::kMlpParams * sizeof(float), hipMemcpyDeviceToHost));
        if (norm) memcpy(norm, ctx->mlp.norm, sizeof ctx->mlp.norm);
    });
}

bh_status bh_mlp_forward(bh_ctx *ctx, int64_t n, const float *x, float *y)
{
    return guarded(ctx, [&] {
        if (!ctx->mlp.ready) throw ApiFail(BH_ERR_STATE, "no model");
        if (n < 1 || !x || !y) throw ApiFail(BH_ERR_ARG, "bad arguments");
        TmpDev tx(n * bh::kMlpIn * sizeof(float)), ty(n * bh::kMlpOut * sizeof(float));
        float *dx = tx.as<float>(), *dy = ty.as<float>();
        BH_HIP(hipMemcpy(dx, x, n * bh::kMlpIn * sizeof(float), hipMemcpyHostToDevice));
        bh::mlp_forward(ctx, dx, n, dy);
        BH_HIP(hipMemcpyAsync(y, dy, n * bh::kMlpOut * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
        BH_HIP(hipStreamSynchronize(ctx->stream));
    });
}

bh_status bh_mlp_grad(bh_ctx *ctx, int64_t n, const float *x, const float *y, float *grad, float *loss)
{
    return guarded(ctx, [&] {
        if (!ctx->mlp.ready) throw ApiFail(BH_ERR_STATE, "no model");
        if (n < 1 || !x || !y || !grad) throw ApiFail(BH_ERR_ARG, "bad arguments");
        TmpDev tx(n * bh::kMlpIn * sizeof(float)), ty(n * bh::kMlpOut * sizeof(float)),
            tg((bh::kMlpParams + 1) * sizeof(float));
        float *dx = tx.as<float>(), *dy = ty.as<float>(), *dg = tg.as<float>();
        BH_HIP(hipMemcpy(dx, x, n * bh::kMlpIn * sizeof(float), hipMemcpyHostToDevice));
        BH_HIP(hipMemcpy(dy, y, n * bh::kMlpOut * sizeof(float), hipMemcpyHostToDevice));
        bh::mlp_grad(ctx, dx, dy, n, dg, dg + bh::kMlpParams);
        BH_HIP(hipMemcpyAsync(grad, dg, bh::kMlpParams * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
        float l = 0;
        BH_HIP(hipMemcpyAsync(&l, dg + bh::kMlpParams, sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
        BH_HIP(hipStreamSynchronize(ctx->stream));
        if (loss) *loss = l;
    });
}

bh_status bh_mlp_collect(bh_ctx *ctx, const bh_params *p, int stride, double rlo, double rhi, int64_t *nadded)
{
    return guarded(ctx, [&] {
        check_params(p);
        if (!ctx->uploaded) throw ApiFail(BH_ERR_STATE, "bh_mlp_collect before bh_upload");
        if (!(rlo >= 0.0 && rhi > rlo)) throw ApiFail(BH_ERR_ARG, "bad size/distance range");
        ctx->mlp.nentries = -1;   // the tree (and its cell numbering) is rebuilt
        if (p->criterion != BH_CRIT_HACK)
            throw ApiFail(BH_ERR_ARG, "training samples are collected with the hackcode1 criterion");
        ctx->mlp.norm[1] = (float)rlo;
        ctx->mlp.norm[2] = (float)rhi;
        bh_params q = *p;
        if (q.group_size == 0) q.group_size = ctx->tuned_group;
        if (q.nleaf < 0) q.nleaf = ctx->tuned_nleaf;
        q.mode = BH_MODE_FAST;
        q.learned = 0;
        if (ctx->cell_cap == 0) ctx->cell_cap = (int)std::min<int64_t>(2 * ctx->nb + 64, INT32_MAX / 2);
        for (int attempt = 0;; attempt++) {
            ctx->mlp_sampling = true;
            bh::build_tree(ctx, q);
            ctx->mlp_sampling = false;
            bh::walk_fast(ctx, q, 2, std::max(1, stride));
            unsigned long long stat[bh::S_NSLOTS];
            BH_HIP(hipMemcpyAsync(stat, ctx->counters.p, sizeof stat, hipMemcpyDeviceToHost, ctx->stream));
            BH_HIP(hipStreamSynchronize(ctx->stream));
            ctx->box_valid = true;
            const unsigned long long f = stat[bh::S_FLAGS];
            if (f & bh::F_NONFINITE) throw ApiFail(BH_ERR_NONFINITE, "non-finite position");
            ctx->ngroups = (int)stat[bh::S_NGROUPS];   // the sampled groups (mlp_collect orders by group)
            if (!f) break;
            if (attempt > 12) throw ApiFail(BH_ERR_CAPACITY, "sample collection did not converge");
            if (f & bh::F_CELL_OVF) ctx->cell_cap = (int)std::min<uint64_t>(stat[bh::S_NCELL_RAW] * 5 / 4 + 64, INT32_MAX / 2);
            if (f & bh::F_TCELL_OVF) ctx->tcell_cap = (int)std::min<uint64_t>(stat[bh::S_TNCELL_RAW] * 5 / 4 + 64, INT32_MAX / 2);
            if (f & bh::F_BOX_OPEN) ctx->box_rounds *= 4;
            if (f & bh::F_FRONTIER) ctx->fcap *= 4;
            if (f & bh::F_LREC_OVF) ctx->lcap *= 4;
        }
        const int64_t got = bh::mlp_collect(ctx, q);
        if (nadded) *nadded = got;
    });
}

bh_status bh_mlp_samples(bh_ctx *ctx, int64_t cap, float *x, float *y, int64_t *n)
{
    return guarded(ctx, [&] {
        const int64_t ns = ctx->mlp.nsamples;
        if (n) *n = ns;
        const int64_t k = std::min(cap, ns);
        if (k > 0 && x) BH_HIP(hipMemcpy(x, ctx->mlp.sx.p, k * bh::kMlpIn * sizeof(float), hipMemcpyDeviceToHost));
        if (k > 0 && y) BH_HIP(hipMemcpy(y, ctx->mlp.sy.p, k * bh::kMlpOut * sizeof(float), hipMemcpyDeviceToHost));
    });
}

bh_status bh_debug_samples(bh_ctx *ctx, int64_t cap, int *cell, float *centre, int64_t *n)
{
    return guarded(ctx, [&] {
        const int64_t ns = ctx->mlp.nlast;
        if (n) *n = ns;
        const int64_t k = std::min(cap, ns);
        if (k <= 0) return;
        std::vector<float> m(4 * (size_t)k);
        BH_HIP(hipMemcpy(m.data(), ctx->mlp.smeta.p, m.size() * sizeof(float), hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < k; i++) {
            if (cell) memcpy(cell + i, &m[4 * i + 3], sizeof(int));
            if (centre)
                for (int a = 0; a < 3; a++) centre[3 * i + a] = m[4 * i + a];
        }
    });
}

bh_status bh_debug_learned(bh_ctx *ctx, int64_t cap, int *cell, int *group, float *res, int64_t *nent, int64_t gcap,
                           float *centre, int64_t *ngroups)
{
    return guarded(ctx, [&] {
        bh::Mlp &M = ctx->mlp;
        if (M.nentries < 0) throw ApiFail(BH_ERR_STATE, "the last evaluation was not a learned one");
        if (res && !M.res_valid)
            throw ApiFail(BH_ERR_STATE, "the entries' outputs were not kept (bh_debug_keep_learned before the evaluation)");
        const int64_t ne = M.nentries;
        if (nent) *nent = ne;
        if (ngroups) *ngroups = ctx->ngroups;
        const int64_t k = std::min(cap, ne);
        if (k > 0) {
            if (cell) BH_HIP(hipMemcpy(cell, M.ecell.p, k * sizeof(int), hipMemcpyDeviceToHost));
            if (group) BH_HIP(hipMemcpy(group, M.egrp.p, k * sizeof(int), hipMemcpyDeviceToHost));
            if (res) BH_HIP(hipMemcpy(res, M.res.p, 12 * k * sizeof(float), hipMemcpyDeviceToHost));
        }
        const int64_t kg = std::min<int64_t>(gcap, ctx->ngroups);
        if (kg > 0 && centre) BH_HIP(hipMemcpy(centre, M.gcentre.p, 4 * kg * sizeof(float), hipMemcpyDeviceToHost));
    });
}

bh_status bh_debug_cell_ranges(bh_ctx *ctx, int64_t cap, int *first, int *count, int64_t *ncell)
{
    return guarded(ctx, [&] {
        if (!ctx->box_valid) throw ApiFail(BH_ERR_STATE, "no tree built");
        const int64_t nc = std::min<int64_t>(ctx->tree.ncell, ctx->tree.cap);
        if (ncell) *ncell = nc;
        const int64_t k = std::min(cap, nc);
        if (k <= 0) return;
        std::vector<int2> fn(k);
        BH_HIP(hipMemcpy(fn.data(), ctx->tree.fnb, k * sizeof(int2), hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < k; i++) {
            if (first) first[i] = fn[i].x;
            if (count) count[i] = fn[i].y;
        }
    });
}

bh_status bh_mlp_clear_samples(bh_ctx *ctx)
{
    return guarded(ctx, [&] { ctx->mlp.nsamples = 0; });
}

bh_status bh_mlp_train(bh_ctx *ctx, int epochs, int batch, float lr, uint64_t seed, float *epoch_loss)
{
    return guarded(ctx, [&] {
        bh::Mlp &M = ctx->mlp;
        if (!M.ready) throw ApiFail(BH_ERR_STATE, "bh_mlp_train before bh_mlp_init");
        const int64_t n = M.nsamples;
        if (n < 64) throw ApiFail(BH_ERR_STATE, "too few training samples");
        if (epochs < 1 || batch < 64 || !(lr > 0.f)) throw ApiFail(BH_ERR_ARG, "bad training arguments");
        hipStream_t st = ctx->stream;
        // output scales of the residual and Jacobian blocks: the rms of the
        // targets below the block's 99th percentile of |y| on a sample of the
        // set (a few cells reach bodies close to the group centre, where the
        // field is far from linear; those targets are also clipped when the
        // batches are gathered, so they cannot dominate the loss)
        if (M.adam_step == 0) {
            const int64_t k = std::min<int64_t>(n, 1 << 20);
            std::vector<float> h((size_t)bh::kMlpOut * k);
            BH_HIP(hipMemcpy(h.data(), M.sy.p, h.size() * sizeof(float), hipMemcpyDeviceToHost));
            for (int blk = 0; blk < 2; blk++) {
                std::vector<float> a;
                a.reserve(h.size());
                for (size_t i = 0; i < h.size(); i++)
                    if (((int)(i % bh::kMlpOut) >= bh::kMlpOutR) == (blk == 1)) a.push_back(std::fabs(h[i]));
                const size_t q = (size_t)(0.99 * (double)(a.size() - 1));
                std::nth_element(a.begin(), a.begin() + q, a.end());
                const float lim = a[q];
                double s2 = 0.0;
                size_t cnt = 0;
                for (float v : a)
                    if (v <= lim) {
                        s2 += (double)v * v;
                        cnt++;
                    }
                M.norm[blk == 0 ? 0 : 3] = (float)std::max(1e-12, std::sqrt(s2 / (double)std::max<size_t>(cnt, 1)));
            }
        }
        float *bx = bh::ensure<float>(M.snew_x, (size_t)batch * bh::kMlpIn);
        float *by = bh::ensure<float>(M.snew_y, (size_t)batch * bh::kMlpOut);
        float *grad = bh::ensure<float>(M.grad, bh::kMlpParams);
        const int64_t nb = n / batch;
        float *lossb = bh::ensure<float>(M.loss, std::max<int64_t>(nb, 1));
        uint64_t rs = seed * 0x9e3779b97f4a7c15ULL + 7;
        auto rnd = [&]() {
            rs ^= rs << 13;
            rs ^= rs >> 7;
            rs ^= rs << 17;
            return rs;
        };
        auto gcd = [](uint64_t a, uint64_t b) {
            while (b) {
                uint64_t t = a % b;
                a = b;
                b = t;
            }
            return a;
        };
        // learning rate: cosine decay from lr to lr/10 over the steps of this call
        const int64_t total_steps = (int64_t)epochs * nb;
        int64_t step = 0;
        for (int e = 0; e < epochs; e++) {
            uint64_t a = (rnd() % (uint64_t)n) | 1;
            while (gcd(a, (uint64_t)n) != 1) a = (a + 2) % (uint64_t)n;
            const uint64_t c = rnd() % (uint64_t)n;
            for (int64_t b = 0; b < nb; b++) {
                bh::mlp_gather(ctx, a, c, b * batch, batch, bx, by);
                bh::mlp_grad(ctx, bx, by, batch, grad, lossb + b);
                const double f = total_steps > 1 ? (double)step / (double)(total_steps - 1) : 0.0;
                const float lr_t = (float)(lr * (0.1 + 0.45 * (1.0 + std::cos(3.141592653589793 * f))));
                step++;
                bh::mlp_adam(ctx, grad, lr_t, ++M.adam_step);
            }
            if (epoch_loss) {
                std::vector<float> lh(std::max<int64_t>(nb, 1));
                BH_HIP(hipMemcpyAsync(lh.data(), lossb, nb * sizeof(float), hipMemcpyDeviceToHost, st));
                BH_HIP(hipStreamSynchronize(st));
                double s = 0.0;
                for (int64_t b = 0; b < nb; b++) s += lh[b];
                epoch_loss[e] = (float)(s / std::max<int64_t>(nb, 1));
            }
        }
        BH_HIP(hipStreamSynchronize(st));
    });
}

}  // extern "C"
