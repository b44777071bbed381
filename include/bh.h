/*
 * bh.h -- C ABI of the MI355X Barnes-Hut gravity engine (libnemo_bh.so).
 *
 * Drop-in for the force path of NEMO's hackforce/hackcode1.  Each entry
 * point names the reference interface it replaces (paths relative to the
 * NEMO source tree):
 *
 *   bh_force        hackforce force_calc()            src/nbody/evolve/hackcode/hackcode1/hackforce.c:177-215
 *                   = maketree() + hackgrav() loop     load.c:32-62, grav.c:22-32
 *   bh_upload/bh_eval/bh_download
 *                   the same force_calc() split so the snapshot stays resident on the device
 *   bh_direct       directcode hackgrav()              src/nbody/evolve/directcode/grav.c:23-65
 *   bh_lf_*         hackcode1 stepsystem()/diagnostics()  code.c:159-191, code_io.c:351-388
 *   bh_batch_*      many independent hackforce calls in one launch (no reference counterpart:
 *                   the reference runs independent processes, src/scripts/nemo.bench:69-90)
 *   bh_autotune     group size / leaf size chosen by timing (no reference counterpart)
 *   bh_mlp_*        learned far-field correction (no reference counterpart)
 *   bh_mkplummer    mkplummer's sampling recipe     src/nbody/init/mkplummer.c:257-422 (own RNG)
 *
 * Conventions: caller-owned host buffers, FP64 at the boundary (the snapshot
 * RealType), positions as [n][3].  G = 1.  Functions return a bh_status and
 * never exit(); bh_last_error() gives the message.  One context per host
 * thread; a context is not thread safe; contexts are independent.
 */
#ifndef NEMO_BH_H
#define NEMO_BH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    BH_OK = 0,
    BH_ERR_ARG = 1,          /* bad argument (null pointer, n < 1, tol < 0, ...)        */
    BH_ERR_DEVICE = 2,       /* HIP runtime error or no device                          */
    BH_ERR_NOMEM = 3,        /* device allocation failed                                 */
    BH_ERR_NONFINITE = 4,    /* NaN/Inf position: the reference loops forever in expandbox */
    BH_ERR_CAPACITY = 5,     /* an internal capacity (walk frontier) was exceeded        */
    BH_ERR_STATE = 6,        /* call out of order (eval before upload, ...)              */
    BH_ERR_INJECTED = 7      /* failure forced by bh_debug_inject (tests of error paths) */
} bh_status;

typedef enum {
    BH_MODE_FAST = 0,    /* group-wise walk, FP32 split-precision interactions (the product path) */
    BH_MODE_STRICT = 1,  /* per-particle walk in FP64 with the reference's exact decisions/order  */
    BH_MODE_DIRECT = 2   /* FP64 direct sum over all massive bodies (tol -> 0 limit)              */
} bh_mode;

/* Cell opening criterion.  BH_CRIT_HACK is hackcode1's (grav.c:122-131): a
 * cell opens when tol^2 |com - x|^2 < size^2 for the body x.  The others are
 * treecode1's (treeload.c:232-252, treegrav.c:137-156): each cell gets a
 * critical radius Rcrit2 (tol = theta; tol = 0 opens everything), and a cell
 * c of side s is accepted for body x only when the squared distance from c's
 * c.o.m. to the side-s cube that holds x exceeds Rcrit2 and that cube is not
 * adjacent (some |com - centre| component > 1.5 s).  Strict mode applies it
 * per body exactly; fast mode applies it per group to the union of the
 * side-s cubes that hold the group's points (never accepts a cell that a
 * body of the group would open).  No counterpart in hackcode1: treecode1
 * selects them with options=bh86 / sw94 (treeload.c:51-54). */
typedef enum {
    BH_CRIT_HACK = 0,    /* hackcode1: (size/tol)^2 against the body's distance to the c.o.m. */
    BH_CRIT_RCRIT = 1,   /* treecode1 default: Rcrit = size/theta + |com - centre|           */
    BH_CRIT_BH86 = 2,    /* treecode1 options=bh86: Rcrit = size/theta                       */
    BH_CRIT_SW94 = 3     /* treecode1 options=sw94: Rcrit = max corner distance from com / theta */
} bh_criterion;

typedef struct {
    double tol;          /* opening parameter: open iff tol^2 |com-x|^2 < size^2 (grav.c:122-131) */
    double eps;          /* Plummer softening */
    int usequad;         /* 1: quadrupole cells (hackforce_qp) */
    int mode;            /* bh_mode */
    double rsize;        /* initial box side (hackforce rsize=4.0) */
    double rmin[3];      /* initial lower corner, used when rmin_given */
    int rmin_given;      /* 0: rmin = -rsize/2 in every dimension (hackforce.c:189-193) */
    int group_size;      /* fast mode: particles per group (0 = tuned default, else 64, 128, 192 or 256) */
    int learned;         /* fast mode: apply the loaded MLP correction to accepted cells */
    int keep_box;        /* bh_eval: start from the box of the previous call (hackcode1 semantics) */
    int nleaf;           /* fast mode: a cell holding <= nleaf bodies interacts body by body (exact);
                            0 keeps the reference's cell interactions; -1 = tuned default */
    int criterion;       /* bh_criterion (default BH_CRIT_HACK); not with learned or batches */
    int supergroup;      /* fast mode: K consecutive groups share one traversal (1 = off, 2..4;
                            0 = tuned default); the interactions are those of K = 1.  Used by the
                            plain single-system walk with the hackcode1 criterion */
    double local_alpha;  /* fast mode, monopole cells (also under the learned correction, whose
                            residuals are added for the cells in its range either way): an accepted cell whose
                            c.o.m. lies farther than R/local_alpha from the group centre (R: the
                            group's radius about that centre) enters the group's second-order local
                            expansion (the cell's softened monopole potential expanded to third
                            order, its field to second, about the centre) instead of one
                            interaction per member.  0 = off (default), at most 0.5; runs with
                            supergroup 1 */
} bh_params;

typedef struct {
    int64_t n2btot;      /* body-body interactions (64-bit; the reference's int overflows) */
    int64_t nbctot;      /* body-cell interactions */
    int64_t ncell;       /* cells in the tree */
    int64_t deep_buckets;/* depth-21 cells holding > 1 body (deeper splits of the reference are merged) */
    double rmin[3];      /* final box */
    double rsize;
    double t_tree_ms;    /* device time: box, keys, sort, topology, moments */
    double t_walk_ms;    /* device time: walk + interactions (+ MLP) */
    int64_t n_mlp;       /* learned corrections applied */
    double t_walk_kernel_ms; /* device time of the walk kernel launch alone */
    double t_sort_ms;    /* device time of the mass-set key sort */
    int64_t ngroups;     /* fast mode: groups of test points walked */
    double t_mlp_ms;     /* learned mode: device time of the learned-correction kernel (MLP inference
                            and the per-group apply) */
    int64_t n_local;     /* fast mode: the body-cell interactions of nbctot evaluated through
                            group local expansions (local_alpha > 0) */
} bh_stats;

typedef struct bh_ctx bh_ctx;

void bh_default_params(bh_params *p);

bh_status bh_create(bh_ctx **ctx, int device);
void bh_destroy(bh_ctx *ctx);
const char *bh_last_error(const bh_ctx *ctx);
const char *bh_version(void);

/* hackforce force_calc on host buffers.  ntest == 0 / tpos == NULL: the
 * mass set is the test set and each body skips itself (pointer identity in
 * grav.c:356); otherwise every test point sees every massive body.
 * Zero-mass bodies are not loaded into the tree (load.c:56-60). */
bh_status bh_force(bh_ctx *ctx, const bh_params *p, int64_t nmass, const double *mass, const double *pos,
                   int64_t ntest, const double *tpos, double *phi, double *acc, bh_stats *st);

/* Resident variant: upload once, evaluate many times (timing, leapfrog). */
bh_status bh_upload(bh_ctx *ctx, int64_t nmass, const double *mass, const double *pos,
                    int64_t ntest, const double *tpos);
bh_status bh_eval(bh_ctx *ctx, const bh_params *p, bh_stats *st);
bh_status bh_download(bh_ctx *ctx, double *phi, double *acc);
/* Evaluate FP64 direct sum on the device for a subset of the test set. */
bh_status bh_direct(bh_ctx *ctx, double eps, int64_t nidx, const int64_t *idx, double *phi, double *acc);
/* bh_direct without any device allocation.  No counterpart in the reference:
 * for callers that cannot allocate device memory (code that must not call
 * hipMalloc, e.g. around HIP graph capture or in a serving loop).  The
 * per-call staging of bh_direct (target indices and FP64 results) lives in a
 * fixed array of the library, one per device, held until the process ends;
 * at most BH_DIRECT_NOALLOC_MAX targets (BH_ERR_ARG beyond).  Results equal
 * bh_direct's bit for bit.  Every call on one device shares that array:
 * calls that overlap on one device (from several threads or contexts) race. */
#define BH_DIRECT_NOALLOC_MAX 32768
bh_status bh_direct_noalloc(bh_ctx *ctx, double eps, int64_t nidx, const int64_t *idx, double *phi, double *acc);
/* Synchronise and return the device time of the last bh_eval kernels. */
bh_status bh_sync(bh_ctx *ctx);

/* Device tree inspection (tests): Morton keys of the uploaded mass set in
 * input order, and the cells in DFS pre-order (quad6 only after an
 * evaluation with usequad). */
bh_status bh_debug_keys(bh_ctx *ctx, uint64_t *keys);
bh_status bh_debug_cells(bh_ctx *ctx, int64_t cap, int *level, int *nbody, double *mass, double *com,
                         double *quad6, int64_t *ncell);
/* Sorted order of the test set (test index -> input index) and the group
 * partition of the last fast-mode evaluation (first, count per group). */
bh_status bh_debug_perm(bh_ctx *ctx, uint32_t *perm);
bh_status bh_debug_groups(bh_ctx *ctx, int64_t cap, int *first, int *count, int64_t *ngroups);
/* Force the next call to fail with BH_ERR_INJECTED (error-path tests). */
void bh_debug_inject(bh_ctx *ctx, int code);

/* hackcode1 time integration on the device (stepsystem, code.c:159-191;
 * diagnostics, code_io.c:351-388).  State in FP64, input order.  The box
 * starts from p's rmin/rsize (hackcode1: -2, 4; code.c:119-120) and then only
 * grows.  A step = force pass (tree + walk), second-order velocity
 * correction when nstep > 0, diagnostics of that state, drift.
 * diag: 36 doubles per step: Energy[3] (T+U, T, U), KETensor[9],
 * PETensor[9], AMTensor[9], CMPhaseSpace[6]. */
bh_status bh_lf_init(bh_ctx *ctx, const bh_params *p, int64_t n, const double *mass, const double *pos,
                     const double *vel, double dt);
/* Force pass + velocity correction + diagnostics; the state is then at output time. */
bh_status bh_lf_force(bh_ctx *ctx, double *diag, int64_t *n2b, int64_t *nbc);
/* Half-kick, drift, half-kick; nstep += 1. */
bh_status bh_lf_drift(bh_ctx *ctx);
/* nsteps full steps with one synchronisation at the end; use_graph: after one
 * eager step, the step is captured in a hipGraph and replayed.  diag
 * [nsteps][36], n2b/nbc [nsteps] may be NULL. */
bh_status bh_lf_run(bh_ctx *ctx, int64_t nsteps, int use_graph, double *diag, int64_t *n2b, int64_t *nbc);
/* Copy the current state to the host; any pointer may be NULL. */
bh_status bh_lf_get(bh_ctx *ctx, double *pos, double *vel, double *acc, double *phi, int64_t *nstep);
/* Root box of the last force pass (hackcode1 keeps it across steps). */
bh_status bh_lf_box(bh_ctx *ctx, double *rmin, double *rsize);
/* Resume a saved run (restorestate, code_io.c:258-284) after bh_lf_init:
 * the accelerations of the saved state (the next force pass corrects the
 * velocities with their change), the step count and the box (rmin NULL:
 * keep the box logic of a fresh start).  Single systems only. */
bh_status bh_lf_restore(bh_ctx *ctx, const double *acc, int64_t nstep, const double *rmin, double rsize);

/* Batched ensembles: nsys independent systems in one tree build and one
 * persistent walk launch (the reference runs one hackforce process per
 * system, src/scripts/nemo.bench:69-90).  System s owns bodies
 * [off[s], off[s+1]) of mass/pos (off[0] = 0, every system non-empty, every
 * body massive, nsys <= 2^18) and has its own eps[s] and tol[s]; each system
 * gets its own box (expandbox from p's rsize/rmin) and its own tree, and its
 * bodies see only its own bodies (self skipped).  p: usequad, group_size,
 * nleaf, rsize/rmin; p->eps/tol are not used.  Fast mode only.  After
 * bh_batch_upload, bh_eval evaluates every system and bh_download returns
 * phi/acc in input order. */
bh_status bh_batch_upload(bh_ctx *ctx, int nsys, const int64_t *off, const double *mass, const double *pos,
                          const double *eps, const double *tol);
bh_status bh_batch_force(bh_ctx *ctx, const bh_params *p, int nsys, const int64_t *off, const double *mass,
                         const double *pos, const double *eps, const double *tol, double *phi, double *acc,
                         bh_stats *st);
/* Leapfrog of a batch: as bh_lf_init, then bh_lf_force / bh_lf_drift /
 * bh_lf_run / bh_lf_get work on every system at once; every system keeps its
 * own box; diag holds 36 doubles per system per step ([step][system][36]). */
bh_status bh_batch_lf_init(bh_ctx *ctx, const bh_params *p, int nsys, const int64_t *off, const double *mass,
                           const double *pos, const double *vel, const double *eps, const double *tol, double dt);

/* Autotuner (no reference counterpart): times the fast walk on the uploaded
 * problem (single or batch) for group_size in {64, 128, 192, 256} x nleaf in
 * {0, 1, 2, 4, 8}, best of reps evaluations each (device time of tree +
 * walk), then (single systems, hackcode1 criterion, local_alpha 0) supergroup 2, 3, 4 for
 * the fastest of those, and makes the fastest the default used when
 * group_size = 0, nleaf = -1 and supergroup = 0.  ms (may be NULL) receives
 * the BH_AUTOTUNE_NCAND times of the first stage in that order; best (may be
 * NULL) the chosen parameters. */
#define BH_AUTOTUNE_NCAND 20
bh_status bh_autotune(bh_ctx *ctx, const bh_params *p, int reps, bh_params *best, double *ms);

/* Learned far-field correction (no reference counterpart).  A 12-64-64-64-9
 * MLP (smooth hard-SiLU activations) on BF16 MFMA predicts, for a cell accepted by the group walk whose
 * size/distance lies in the trained range, the residual between the exact
 * force of the cell's bodies and its multipole force at the group centre,
 * and that residual's Jacobian there; the group's sums R and J give each
 * member at offset dx from the centre the correction R + J dx
 * (params.learned).  params: kMlpParams = 9737 floats (W1[64][12] b1
 * W2[64][64] b2 W3[64][64] b3 W4[9][64] b4; outputs: residual x y z, then
 * the symmetric Jacobian xx xy xz yy yz zz); norm[4] = residual scale,
 * size/distance range lo, hi, Jacobian scale. */
bh_status bh_mlp_init(bh_ctx *ctx, uint64_t seed);
bh_status bh_mlp_set(bh_ctx *ctx, const float *params, const float *norm);
bh_status bh_mlp_get(bh_ctx *ctx, float *params, float *norm);
/* Batch inference / mean-squared-error gradient on host buffers (tests);
 * x: [n][12], y: [n][9]. */
bh_status bh_mlp_forward(bh_ctx *ctx, int64_t n, const float *x, float *y);
bh_status bh_mlp_grad(bh_ctx *ctx, int64_t n, const float *x, const float *y, float *grad, float *loss);
/* Training samples from the uploaded snapshot: a group walk that records, for
 * every stride-th group, its accepted cells with size/distance in [rlo, rhi];
 * each gives (features[12], targets[9]: the exact FP64 residual at the group
 * centre in units of M/r^2 and its symmetric Jacobian in units of M/r^3). */
bh_status bh_mlp_collect(bh_ctx *ctx, const bh_params *p, int stride, double rlo, double rhi, int64_t *nadded);
bh_status bh_mlp_samples(bh_ctx *ctx, int64_t cap, float *x, float *y, int64_t *n);
/* Tests: the cell (DFS pre-order index, as bh_debug_cells) and the group
 * centre ([3]) of each sample added by the last bh_mlp_collect. */
bh_status bh_debug_samples(bh_ctx *ctx, int64_t cap, int *cell, float *centre, int64_t *n);
/* Tests: the entries of the last learned evaluation (cell, group, predicted
 * field res[12] = residual xyz, Jacobian xx xy xz yy yz zz, 0 0 0, in force
 * units; only when bh_debug_keep_learned was on for it, else BH_ERR_STATE
 * for a non-null res) and each group's centre ([4], w unused). */
bh_status bh_debug_learned(bh_ctx *ctx, int64_t cap, int *cell, int *group, float *res, int64_t *nent, int64_t gcap,
                           float *centre, int64_t *ngroups);
/* Tests: keep (on != 0) each entry's MLP output in the following learned
 * evaluations for bh_debug_learned; off by default, when the learned kernel
 * stores only the corrected results. */
void bh_debug_keep_learned(bh_ctx *ctx, int on);
/* Tests: each cell's first body (index in the tree's sorted order, see
 * bh_debug_perm) and body count. */
bh_status bh_debug_cell_ranges(bh_ctx *ctx, int64_t cap, int *first, int *count, int64_t *ncell);
bh_status bh_mlp_clear_samples(bh_ctx *ctx);
/* Adam on the collected samples (mean squared error of the scaled targets),
 * the learning rate decaying from lr to lr/10 (cosine) over the call's steps. */
bh_status bh_mlp_train(bh_ctx *ctx, int epochs, int batch, float lr, uint64_t seed, float *epoch_loss);

/* Synthetic inputs: equal-mass Plummer sphere in virial units (G=M=1,
 * E=-1/4) following mkplummer's recipe with mfrac=0.999, rfrac=22.8042468,
 * zero centre of mass; counter-based RNG keyed by seed, so the result does
 * not depend on thread count.  vel may be NULL. */
void bh_mkplummer(int64_t n, uint64_t seed, double *mass, double *pos, double *vel);
/* Synthetic galaxy (no reference counterpart; the time-integration
 * workload): exponential disk (mass mdisk, scale length rdisk, sech^2 scale
 * height zdisk, Toomre q) in a Hernquist halo (mass mhalo, scale radius
 * ahalo, isotropic, truncated at 95 % of its mass), G = 1, approximate
 * equilibrium; bodies 0..ndisk-1 are the disk.  Counter-based RNG keyed by
 * seed. */
void bh_mkdiskhalo(int64_t ndisk, int64_t nhalo, uint64_t seed, double mdisk, double rdisk, double zdisk, double q,
                   double mhalo, double ahalo, double *mass, double *pos, double *vel);

#ifdef __cplusplus
}
#endif
#endif
