#!/usr/bin/env python3
"""Benchmark: particle-force evaluations per second at N = 2^22 Plummer on
one MI355X, at the opening angle where the median |da|/|a| against an FP64
direct sum (8192-body seeded subset) is 1e-3: the learned-correction walk
against the plain tree.

A step is one full force evaluation of the resident snapshot: root box,
Morton keys, radix sort, topology, FP64 moments, group walk with the
interactions and, for the learned variant, the MLP correction.  Inputs are
resident in HBM before timing starts; results stay on the device.

Plain tree: monopole and quadrupole cells, and monopole cells with the
group local expansion of far cells (local_alpha in --local), are each
calibrated to the target median error and the fastest is the plain figure.  Learned: for walks with
quadrupole and with monopole cells, an MLP is trained on the device from
residual samples of three other Plummer snapshots (seeds 41, 43 and 47, tol
0.9/1.1/1.3, cells with size/r in [0.45, 1.5]; the targets are the exact
cell force minus that walk's multipole force at the group centre and its
Jacobian, applied to each member to first order), then that learned walk is
calibrated to the same target on the seed-123 snapshot; the walk with
monopole cells also runs with the group local expansion of far cells
(local_alpha in --learned-local, same model: the correction covers the same
recorded cells); the fastest is the learned figure (all are in
`config.learned_variants`).
`value` is the evals/s of the faster of the two (`config.value_is` names it;
both variants and their ratio `learned_over_plain` are in `config`); with
--no-learned it is the plain figure.

Prints one JSON line.  `roofline` is for the dominant kernel, the walk: 20
flop per interaction (P2P and P2C, counted by the kernel; with local
expansions only the interactions evaluated one by one, nbctot - n_local of
the cell interactions) over the walk kernel's own event-timed duration, against MI355X's FP32 vector peak (the
walk issues no MFMA: its bound is labelled "valu", the vector ALU).
`roofline_mlp` is the learned correction's kernel (MLP inference of every
entry and, in the same kernel, each group's correction added to its
members): 19,072 flop per MLP evaluation against the dense BF16 MFMA peak.  `cpu_baseline` times the CPU restatement
of hackforce (FP64, single thread) on a bounded sample of the same snapshot
at the oracle's own tol for the same median error (bisected on the error
subset, monopole and quadrupole cells, the faster reported): one full tree
build plus the walk of a subset of bodies, extrapolated to N.
`cpu_baseline_allcores` is the same with the walk on every host thread
(at most 16) over the one tree.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle-force evals/s, N=4M, median |da|/|a|=1e-3: learned vs plain tree"
FP32_PEAK_TFLOPS = 157.3     # MI355X FP32 (vector packed / dense MFMA figure)
BF16_PEAK_TFLOPS = 2500.0    # MI355X dense BF16 MFMA
FLOP_PER_INTERACTION = 20
FLOP_PER_MLP = 2 * (12 * 64 + 2 * 64 * 64 + 64 * 9)   # 19,072: residual and its Jacobian


def rel_err(a, ref):
    return np.linalg.norm(a - ref, axis=1) / np.linalg.norm(ref, axis=1)


def measure_err(eng, E, p, idx, dacc):
    eng.eval(p)
    _, acc = eng.download()
    e = rel_err(acc[idx], dacc)
    return float(np.median(e)), float(np.percentile(e, 99))


def calibrate_tol(eng, E, base, idx, dacc, target, lo=0.2, hi=1.4, iters=9):
    """Largest tol whose median error is <= target (bisection; error grows with tol)."""
    best = None
    for _ in range(iters):
        mid = 0.5 * (lo + hi)
        p = E.default_params(**dict(base, tol=mid))
        med, p99 = measure_err(eng, E, p, idx, dacc)
        if med <= target:
            best = (mid, med, p99)
            lo = mid
        else:
            hi = mid
    if best is None:
        p = E.default_params(**dict(base, tol=lo))
        med, p99 = measure_err(eng, E, p, idx, dacc)
        best = (lo, med, p99)
    return best


def time_steps(eng, p, steps, warmup):
    for _ in range(warmup):
        eng.eval(p)
    stats = []
    t0 = time.perf_counter()
    for _ in range(steps):
        stats.append(eng.eval(p))
    t1 = time.perf_counter()
    return (t1 - t0) / steps, stats


def host_threads():
    """Host threads for the all-core CPU baseline: the CPUs this process may
    run on, at most 16 (one GPU's share of a shared host)."""
    return max(1, min(len(os.sched_getaffinity(0)), 16))


def cpu_tol_at_target(tree, idx_err, dacc, eps, target, nthreads, lo=0.2, hi=1.6, iters=9):
    """Largest tol at which the oracle's own walk (per-body opening test, FP64)
    has median |da|/|a| <= target on the error subset (bisection over one tree)."""
    best = None
    for _ in range(iters):
        mid = 0.5 * (lo + hi)
        _, acc, _, _, _ = tree.walk(idx_err, mid, eps, nthreads=nthreads)
        med = float(np.median(rel_err(acc, dacc)))
        if med <= target:
            best, lo = (mid, med), mid
        else:
            hi = mid
    return best if best is not None else (lo, float("nan"))


def cpu_baseline(m, x, idx_err, dacc, eps, target, budget_s, nthreads_list=(1,)):
    """The oracle (hackforce restated, FP64) on the host: for monopole and
    quadrupole cells, tol bisected on the error subset to median |da|/|a| =
    target, then one tree build and the walk of a bounded sample at that tol,
    extrapolated to N; the faster cell type is reported, per thread count."""
    from oracle import oracle

    n = len(m)
    nt_cal = host_threads()
    runs = {}
    for usequad in (0, 1):
        tree = oracle.Tree(m, x, usequad)
        try:
            tol, med = cpu_tol_at_target(tree, idx_err, dacc, eps, target, nt_cal)
            for nthreads in nthreads_list:
                # pilot: size the sample for the time budget
                pilot = idx_err[:64]
                per = tree.walk(pilot, tol, eps)[4] / len(pilot)
                ns = int(max(64, min(n, (budget_s / 4 - tree.t_build) / max(per, 1e-9) * nthreads)))
                rest = np.setdiff1d(np.random.default_rng(7).permutation(n)[: ns + len(idx_err)], idx_err)
                sample = np.sort(np.concatenate([idx_err, rest])[:ns])
                _, acc, n2b, nbc, t_walk = tree.walk(sample, tol, eps, nthreads=nthreads)
                t_full = tree.t_build + t_walk / ns * n
                have = np.isin(idx_err, sample)
                e = rel_err(acc[np.searchsorted(sample, idx_err[have])], dacc[have])
                runs[(usequad, nthreads)] = dict(value=n / t_full, tol=tol, med_cal=med, t_build=tree.t_build,
                                                 ns=ns, t_walk=t_walk, ipb=(n2b + nbc) / ns,
                                                 med=float(np.median(e)), p99=float(np.percentile(e, 99)))
        finally:
            tree.close()
    out = {}
    for nthreads in nthreads_list:
        uq = max((0, 1), key=lambda q: runs[(q, nthreads)]["value"])
        r, o = runs[(uq, nthreads)], runs[(1 - uq, nthreads)]
        how = ("1 thread" if nthreads == 1 else
               "%d threads walking slices of the sample over one tree, as %d processes that each build the tree "
               "and walk 1/%d of the bodies would" % (nthreads, nthreads, nthreads))
        out[nthreads] = {
            "value": r["value"],
            "unit": "evals/s",
            "cores": nthreads,
            "kind": "port",
            "sample": "oracle hackforce restatement (FP64, %s), %s cells at its own tol=%.4f, bisected on the "
                      "%d-body error subset to median |da|/|a| %.3g (target %g; p99 %.3g); full tree build of "
                      "N=%d (%.2f s) + walk of %d bodies (%.2f s, %.0f interactions/body), extrapolated to N; "
                      "%s cells at tol=%.4f: %.3g evals/s" % (
                          how, "quadrupole" if uq else "monopole", r["tol"], len(idx_err), r["med"], target,
                          r["p99"], n,
                          r["t_build"], r["ns"], r["t_walk"], r["ipb"], "monopole" if uq else "quadrupole",
                          o["tol"], o["value"]),
        }
    return out


def train_learned(eng, E, n, eps, seeds=(41, 43, 47), tols=(0.9, 1.1, 1.3), stride=2, epochs=10, rlo=0.45, rhi=1.5,
                  usequad=1, group=192):
    """Residual samples of other snapshots, then training on the device.
    The correction (and the samples) cover accepted cells with size/r in
    [rlo, rhi]: below 0.45 the residual is small and correcting it does not
    move the median error, while it would triple the MLP evaluations."""
    t0 = time.perf_counter()
    eng.mlp_init(5)
    eng.mlp_clear_samples()
    for s in seeds:
        m, x, _ = E.mkplummer(n, seed=s, with_vel=False)
        eng.upload(m, x)
        for tol in tols:
            # a fixed group size, so that the samples do not follow the last
            # autotuner's near-ties (192-point groups give a third more samples
            # than 256-point ones, and a model that reaches a larger tol)
            eng.mlp_collect(E.default_params(tol=tol, eps=eps, usequad=usequad, group_size=group, nleaf=1),
                            stride=stride, rlo=rlo, rhi=rhi)
    nsamp = eng.mlp_nsamples()
    t1 = time.perf_counter()
    loss = eng.mlp_train(epochs=epochs, batch=8192, lr=1e-3)
    t2 = time.perf_counter()
    eng.mlp_clear_samples()
    return {"samples": nsamp, "snapshots": "Plummer N=%d seeds %s, tol %s, %d-point groups, every %d-th group" % (
                n, list(seeds), list(tols), group, stride), "size_over_r_range": [rlo, rhi],
            "epochs": epochs, "loss_first": float(loss[0]), "loss_last": float(loss[-1]),
            "t_collect_s": t1 - t0, "t_train_s": t2 - t1}


def profiled_traffic(kernel, profiles=None):
    """HBM bytes per dispatch of `kernel` from the committed rocprofv3
    summaries of the N = 4M walks (profiles/walk_plummer4M_*.md: FETCH_SIZE
    and WRITE_SIZE runs, each scaled by a device copy of known size), from the
    first summary whose dominant kernel is this kernel; else None.  Counters
    cannot be read from inside the timed process, so the figure is the one
    measured on the same workload by the profile's command."""
    pdir = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles")
    if profiles is None:
        try:
            profiles = sorted(f for f in os.listdir(pdir) if f.startswith("walk_plummer4M_") and f.endswith(".md"))
        except OSError:
            return None, None
    for prof in profiles:
        got = _profile_traffic(os.path.join(pdir, prof), kernel)
        if got is not None:
            return got, "profiles/%s (rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE; scaled)" % prof
    return None, None


def _profile_traffic(path, kernel):
    try:
        text = open(path).read()
    except OSError:
        return None
    want = kernel.replace(" ", "")
    dom, in_traffic = None, False
    for line in text.splitlines():
        if line.startswith("Dominant kernel:") and "`" in line:
            dom = line.split("`")[1].replace(" ", "")
        if line.startswith("## "):
            in_traffic = line.startswith("## HBM traffic")
        if (in_traffic and dom == want and line.startswith("| `")
                and line.split("`")[1].replace(" ", "") == want):
            cells = [c.strip() for c in line.strip().strip("|").split("|")]
            try:
                return float(cells[-1])
            except ValueError:
                return None
    return None


def walk_roofline(n_int, t_walk_s, kernel):
    achieved = FLOP_PER_INTERACTION * n_int / t_walk_s / 1e12
    return {"bound": "valu", "compute_unit": "FP32 vector ALU (the walk issues no MFMA)", "kernel": kernel,
            "achieved": achieved, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": achieved / FP32_PEAK_TFLOPS,
            "traffic": None,
            "flop_convention": "20 flop per interaction (P2P and P2C), interactions counted by the kernel"}


def run_ensemble(args, E, log):
    """512 independent Plummer systems, N log-uniform in [256, 65536]
    (seeded), eps_i = 1/(4 N_i^(1/3)), tol 0.7, quadrupole cells: one tree
    build + one walk launch for all systems (group/leaf size from the
    autotuner), then 32 graph-captured leapfrog steps at dt = 1/32."""
    rng = np.random.default_rng(args.seed)
    nsys = args.nsys
    ns = np.round(np.exp(rng.uniform(np.log(256), np.log(65536), nsys))).astype(np.int64)
    ms, xs, vs = [], [], []
    for s, n in enumerate(ns):
        m, x, v = E.mkplummer(int(n), seed=args.seed * 100003 + s)
        ms.append(m)
        xs.append(x)
        vs.append(v)
    off = np.concatenate([[0], np.cumsum(ns)]).astype(np.int64)
    eps = 1.0 / (4.0 * ns ** (1.0 / 3.0))
    tol = np.full(nsys, 0.7)
    m, x, v = np.concatenate(ms), np.concatenate(xs), np.concatenate(vs)
    ntot = int(off[-1])
    eng = E.Engine(0)
    # FP64 direct sums per system on the device (256 seeded bodies each), before the batch upload
    ref = []
    for s in range(nsys):
        a, b = int(off[s]), int(off[s + 1])
        idx = np.sort(np.random.default_rng(s).permutation(b - a)[: min(256, b - a)])
        eng.upload(m[a:b], x[a:b])
        ref.append((idx, eng.direct(float(eps[s]), idx)[1]))
    eng.batch_upload(off, m, x, eps, tol)
    base = E.default_params(usequad=1)
    t0 = time.perf_counter()
    best, ms_c = eng.autotune(base, reps=2)
    t_tune = time.perf_counter() - t0
    log("autotune: group %d nleaf %d (%.1f s; candidates %s ms)" % (best.group_size, best.nleaf, t_tune,
                                                                   np.round(ms_c, 3).tolist()))
    p = E.default_params(usequad=1)          # tuned defaults
    sec, stats = time_steps(eng, p, args.steps, args.warmup)
    _, acc = eng.download()
    med = []
    p99 = []
    for s in range(nsys):
        a = int(off[s])
        idx, da = ref[s]
        e = np.linalg.norm(acc[a + idx] - da, axis=1) / np.linalg.norm(da, axis=1)
        med.append(np.median(e))
        p99.append(np.percentile(e, 99))
    last = stats[-1]
    n_int = int(last.n2btot + last.nbctot)
    t_walk = float(np.mean([st.t_walk_kernel_ms for st in stats])) * 1e-3
    # leapfrog: one eager step, then graph replays
    lf = E.Leapfrog(eng, m, x, v, 1.0 / 32.0, batch=(off, eps, tol), usequad=1)
    lf.run(1, graph=True)
    t0 = time.perf_counter()
    d, n2b, nbc = lf.run(args.lf_steps, graph=True)
    t_lf = time.perf_counter() - t0
    e0 = d[0, :, 0]
    drift = np.abs(d[-1, :, 0] - e0) / np.abs(e0)
    eng.close()
    return {
        "metric": "aggregate particle-force evals/s, 512-system Plummer ensemble in one launch",
        "value": ntot / sec, "unit": "evals/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": sec * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic Plummer ensemble (own counter RNG, seed %d)" % args.seed,
        "config": {"workload": "ensemble_%d_plummer_logU_256_65536" % nsys, "nsys": nsys, "bodies": ntot,
                   "n_min": int(ns.min()), "n_max": int(ns.max()), "tol": 0.7, "usequad": 1,
                   "eps": "1/(4 N^(1/3)) per system", "tuned_group": best.group_size, "tuned_nleaf": best.nleaf,
                   "autotune_ms": ms_c.tolist(), "interactions_per_s": n_int / sec,
                   "interactions_per_step": n_int, "t_tree_ms": float(np.mean([st.t_tree_ms for st in stats])),
                   "t_walk_kernel_ms": t_walk * 1e3,
                   "median_rel_err": {"median_over_systems": float(np.median(med)), "max": float(np.max(med))},
                   "p99_rel_err": {"median_over_systems": float(np.median(p99)), "max": float(np.max(p99))},
                   "leapfrog": {"steps": args.lf_steps, "dt": 1.0 / 32.0, "graph": True,
                                "ms_per_step": t_lf / args.lf_steps * 1e3,
                                "interactions_per_s": float((n2b + nbc).sum() / t_lf),
                                "energy_drift": {"median": float(np.median(drift)), "max": float(np.max(drift))}}},
        "roofline": walk_roofline(n_int, t_walk, "k_walk<%d,true,false,false,false,false>" % (best.group_size // 64)),
    }


def run_plummer1m(args, E, log):
    """Plummer N = 2^20, eps 0.01, monopole cells, tol in {0.5, 0.7, 1.0}:
    one force evaluation per tol (group and leaf size from the autotuner at
    that tol), its median and 99th-percentile |da|/|a| against the FP64
    direct sum on a seeded 8192-body subset, interactions/s and the walk's
    fraction of the FP32 roofline."""
    n = 1 << 20
    m, x, _ = E.mkplummer(n, seed=args.seed, with_vel=False)
    eng = E.Engine(0)
    eng.upload(m, x)
    idx = np.sort(np.random.default_rng(args.seed).permutation(n)[: args.nsub])
    _, dacc = eng.direct(0.01, idx)
    rows = []
    for tol in (0.5, 0.7, 1.0):
        best, _ = eng.autotune(E.default_params(tol=tol, eps=0.01, usequad=0), reps=2)
        p = E.default_params(tol=tol, eps=0.01, usequad=0, group_size=best.group_size, nleaf=best.nleaf)
        med, p99 = measure_err(eng, E, p, idx, dacc)
        sec, stats = time_steps(eng, p, args.steps, args.warmup)
        st = stats[-1]
        n_int = int(st.n2btot + st.nbctot)
        t_walk = float(np.mean([s.t_walk_kernel_ms for s in stats])) * 1e-3
        rows.append({"tol": tol, "group_size": best.group_size, "nleaf": best.nleaf, "median_rel_err": med,
                     "p99_rel_err": p99, "ms_per_step": sec * 1e3, "evals_per_s": n / sec,
                     "interactions_per_step": n_int, "interactions_per_s": n_int / sec,
                     "t_tree_ms": float(np.mean([s.t_tree_ms for s in stats])), "t_walk_kernel_ms": t_walk * 1e3,
                     "roofline": walk_roofline(n_int, t_walk, "k_walk<%d,false,false,false,false,false>" %
                                               (best.group_size // 64))})
        log("plummer 2^20 tol %.1f: group %d nleaf %d median %.3g p99 %.3g  %.3f ms/step  %.3g interactions/s" % (
            tol, best.group_size, best.nleaf, med, p99, sec * 1e3, n_int / sec))
    eng.close()
    head = rows[1]
    return {
        "metric": "interactions/s, Plummer N=2^20, eps 0.01, tol 0.7 (0.5 and 1.0 in config)",
        "value": head["interactions_per_s"], "unit": "interactions/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": head["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32",
        "data": "synthetic Plummer (mkplummer recipe, own counter RNG, seed %d)" % args.seed,
        "config": {"workload": "plummer_n1048576_tol_0.5_0.7_1.0", "n": n, "eps": 0.01, "usequad": 0,
                   "err_subset": len(idx), "per_tol": rows},
        "roofline": head["roofline"],
    }


def run_diskhalo(args, E, log):
    """Exponential disk (N = 2^21, M = 0.1, scale length 1, sech^2 height
    0.1, Q = 1.5) in a Hernquist halo (N = 2^22, M = 1, a = 4), eps = 0.02:
    force error of the first evaluation, then 256 graph-captured leapfrog
    steps at dt = 1/64 (tol 0.7, quadrupole cells, 192-point groups unless
    --group names another size)."""
    nd, nh = args.ndisk, args.nhalo
    group = args.group or 192
    m, x, v = E.mkdiskhalo(nd, nh, seed=args.seed)
    n = len(m)
    eng = E.Engine(0)
    eng.upload(m, x)
    idx = np.sort(np.random.default_rng(args.seed).permutation(n)[: args.nsub])
    _, dacc = eng.direct(0.02, idx)
    p = E.default_params(tol=args.dh_tol, eps=0.02, usequad=1, rsize=4.0, group_size=group)
    st = eng.eval(p)
    _, acc = eng.download()
    e = rel_err(acc[idx], dacc)
    lf = E.Leapfrog(eng, m, x, v, 1.0 / 64.0, tol=args.dh_tol, eps=0.02, usequad=1, group_size=group)
    lf.run(1, graph=True)
    t0 = time.perf_counter()
    d, n2b, nbc = lf.run(args.lf_steps, graph=True)
    t_lf = time.perf_counter() - t0
    drift = abs(d[-1, 0] - d[0, 0]) / abs(d[0, 0])
    eng.close()
    return {
        "metric": "leapfrog steps/s, disk N=2^21 + halo N=2^22, hipGraph-captured",
        "value": args.lf_steps / t_lf, "unit": "steps/s", "n_gpus": 1, "steps": args.lf_steps, "warmup": 1,
        "ms_per_step": t_lf / args.lf_steps * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic disk + Hernquist halo (own recipe and RNG, seed %d)" % args.seed,
        "config": {"workload": "diskhalo_%d_%d_leapfrog" % (nd, nh), "n": n, "eps": 0.02, "tol": args.dh_tol,
                   "group_size": group,
                   "usequad": 1, "dt": 1.0 / 64.0, "evals_per_s": n * args.lf_steps / t_lf,
                   "interactions_per_s": float((n2b + nbc).sum() / t_lf),
                   "relative_energy_drift": float(drift), "energy_first_last": [float(d[0, 0]), float(d[-1, 0])],
                   "first_eval": {"median_rel_err": float(np.median(e)), "p99_rel_err": float(np.percentile(e, 99)),
                                  "t_tree_ms": st.t_tree_ms, "t_walk_kernel_ms": st.t_walk_kernel_ms}},
        "roofline": walk_roofline(int(st.n2btot + st.nbctot), st.t_walk_kernel_ms * 1e-3,
                                  "k_walk<%d,true,false,false,false,false>" % (group // 64)),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="plummer4m", choices=["plummer4m", "plummer1m", "ensemble", "diskhalo"],
                    help="plummer4m: the headline metric (default); plummer1m: interactions/s and roofline at "
                         "N=2^20 for tol 0.5/0.7/1.0; ensemble / diskhalo: the batch and time-integration workloads")
    ap.add_argument("--seed", type=int, default=123)
    ap.add_argument("--nsys", type=int, default=512)
    ap.add_argument("--lf-steps", type=int, default=0, help="leapfrog steps (default 32 ensemble, 256 disk+halo)")
    ap.add_argument("--ndisk", type=int, default=1 << 21)
    ap.add_argument("--nhalo", type=int, default=1 << 22)
    ap.add_argument("--dh-tol", type=float, default=0.7)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=1 << 22)
    ap.add_argument("--eps", type=float, default=0.01)
    ap.add_argument("--target", type=float, default=1e-3)
    ap.add_argument("--nsub", type=int, default=8192)
    ap.add_argument("--tol", type=float, default=None, help="skip the plain calibration and use this tol")
    ap.add_argument("--group", type=int, default=0, help="group size (64, 128, 192, 256); 0: autotuned per variant")
    ap.add_argument("--nleaf", type=int, default=-1, help="with --group: leaf size (-1: the engine's default)")
    ap.add_argument("--quad", type=int, default=-1, help="plain tree: -1 calibrates both and keeps the faster")
    ap.add_argument("--local", default="0.12,0.2",
                    help="plain monopole tree with group local expansions at these local_alpha values "
                         "(comma-separated; 'none' skips)")
    ap.add_argument("--local-only", action="store_true", help="plain tree: only the local-expansion variants")
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--learned-tol", type=float, default=None,
                    help="skip the learned calibration and use this tol (profiling runs)")
    ap.add_argument("--learned-quad", type=int, default=-1,
                    help="learned walk cells: 1 quadrupole, 0 monopole, -1 both (each with its own model; the faster "
                         "is the learned figure)")
    ap.add_argument("--learned-local", default="0.2",
                    help="learned walk with monopole cells: also with group local expansions at these local_alpha "
                         "values (comma-separated; 'none' skips), same model")
    ap.add_argument("--epochs", type=int, default=100)
    ap.add_argument("--train-group", type=int, default=192, help="group size of the training-sample walks")
    ap.add_argument("--train-stride", type=int, default=1, help="training samples from every k-th group")
    ap.add_argument("--train-rlo", type=float, default=0.45,
                    help="lower end of the size/r range the model is trained for and corrects")
    ap.add_argument("--save-model", default=None, help="write the trained MLP (params, norm) to this .npz")
    ap.add_argument("--model", default=None, help="load the MLP from this .npz instead of training it")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="also run one device-to-device copy of this many MiB (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this engine runs on one GPU; --gpus %d is not supported (how the walk would shard "
              "across GPUs is described, not built)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    from nemo_amd import engine as E

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    if args.workload == "plummer1m":
        print(json.dumps(run_plummer1m(args, E, log)))
        return
    if args.workload != "plummer4m":
        if args.lf_steps <= 0:
            args.lf_steps = 32 if args.workload == "ensemble" else 256
        out = (run_ensemble if args.workload == "ensemble" else run_diskhalo)(args, E, log)
        print(json.dumps(out))
        return

    if args.copy_probe:
        # a plain copy kernel of known size, so that FETCH_SIZE/WRITE_SIZE
        # counters of the walk can be scaled (plumbing only)
        import torch

        a = torch.ones(args.copy_probe << 18, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    n = args.n
    eng = E.Engine(0)

    def model_path(path, lq):
        return path[:-4] + "_q%d.npz" % lq if path.endswith(".npz") else path + "_q%d" % lq

    def learned_model(lq):
        """Load (--model) or train the MLP for learned walks with quadrupole
        (lq = 1) or monopole (lq = 0) cells."""
        if args.model:
            with np.load(model_path(args.model, lq)) as z:
                eng.mlp_set(z["params"], z["norm"])
            return {"loaded_from": model_path(args.model, lq)}
        tr = train_learned(eng, E, n, args.eps, epochs=args.epochs, usequad=lq, group=args.train_group,
                           stride=args.train_stride, rlo=args.train_rlo)
        log("learned (usequad=%d): %d samples, collect %.1f s, train %.1f s, loss %.3f -> %.3f" % (
            lq, tr["samples"], tr["t_collect_s"], tr["t_train_s"], tr["loss_first"], tr["loss_last"]))
        if args.save_model:
            params, norm = eng.mlp_get()
            np.savez(model_path(args.save_model, lq), params=params, norm=norm)
        return tr

    m, x, _ = E.mkplummer(n, seed=123, with_vel=False)
    eng.upload(m, x)
    rng = np.random.default_rng(123)
    idx = np.sort(rng.permutation(n)[: args.nsub])
    t0 = time.perf_counter()
    _, dacc = eng.direct(args.eps, idx)
    log("direct-sum reference: %d targets x %d sources in %.2f s" % (len(idx), n, time.perf_counter() - t0))

    def variant(base, fixed_tol=None):
        tune = None
        if base.get("group_size", 0) == 0:
            # group and leaf size from the autotuner, timed at a nominal tol
            nominal = fixed_tol or (0.95 if base.get("learned") else (0.75 if base["usequad"] else 0.67))
            t0 = time.perf_counter()
            best, ms = eng.autotune(E.default_params(**dict(base, tol=nominal)), reps=2)
            # near-ties (within 2 %) go to the larger group: its conservative
            # opening is more accurate at the same tol
            near = [k for k, t in enumerate(ms) if t <= 1.02 * min(ms)]
            g, nl = max((E.AUTOTUNE_CANDIDATES[k] for k in near), key=lambda c: (c[0], -ms[E.AUTOTUNE_CANDIDATES.index(c)]))
            # super-groups as the autotuner's second stage chose them for its best (group, leaf)
            sg = best.supergroup if (g, nl) == (best.group_size, best.nleaf) else 1
            base = dict(base, group_size=g, nleaf=nl, supergroup=sg)
            tune = {"nominal_tol": nominal, "candidates": [list(c) for c in E.AUTOTUNE_CANDIDATES],
                    "ms": [round(float(t), 3) for t in ms], "seconds": time.perf_counter() - t0}
        if fixed_tol is None:
            tol, med, p99 = calibrate_tol(eng, E, base, idx, dacc, args.target, hi=1.8 if base.get("learned") else 1.4)
        else:
            tol = fixed_tol
            med, p99 = measure_err(eng, E, E.default_params(**dict(base, tol=tol)), idx, dacc)
        p = E.default_params(**dict(base, tol=tol))
        sec, stats = time_steps(eng, p, args.steps, args.warmup)
        last = stats[-1]
        return {"usequad": base["usequad"], "learned": base.get("learned", 0), "tol": tol,
                "group_size": base["group_size"], "nleaf": base.get("nleaf", -1),
                "supergroup": base.get("supergroup", 1), "local_alpha": base.get("local_alpha", 0.0),
                "autotune": tune,
                "median_rel_err": med, "p99_rel_err": p99, "ms_per_step": sec * 1e3, "evals_per_s": n / sec,
                "interactions_per_step": int(last.n2btot + last.nbctot), "n2b": int(last.n2btot),
                "nbc": int(last.nbctot), "n_local": int(last.n_local), "n_mlp": int(last.n_mlp),
                "ngroups": int(last.ngroups),
                "deep_buckets": int(last.deep_buckets),
                "t_tree_ms": float(np.mean([s.t_tree_ms for s in stats])),
                "t_sort_ms": float(np.mean([s.t_sort_ms for s in stats])),
                "t_walk_kernel_ms": float(np.mean([s.t_walk_kernel_ms for s in stats])),
                "t_mlp_kernel_ms": float(np.mean([s.t_mlp_ms for s in stats])),
                "t_walk_phase_ms": float(np.mean([s.t_walk_ms for s in stats]))}

    plain = []
    for usequad in (() if args.local_only else (0, 1) if args.quad < 0 else (args.quad,)):
        v = variant(dict(eps=args.eps, usequad=usequad, group_size=args.group, nleaf=args.nleaf), args.tol)
        log("plain tree usequad=%d: group %d nleaf %d tol=%.4f median=%.3g p99=%.3g  %.3f ms/step" % (
            usequad, v["group_size"], v["nleaf"], v["tol"], v["median_rel_err"], v["p99_rel_err"], v["ms_per_step"]))
        plain.append(v)
    alphas = [] if args.local.strip().lower() in ("", "none") else [float(a) for a in args.local.split(",")]
    if args.quad != 1:
        for alpha in alphas:
            v = variant(dict(eps=args.eps, usequad=0, group_size=args.group, nleaf=args.nleaf, local_alpha=alpha),
                        args.tol)
            log("plain tree + local expansions (alpha %.3f): group %d nleaf %d tol=%.4f median=%.3g p99=%.3g  "
                "%.3f ms/step (%.1f %% of the cell interactions by expansion)" % (
                    alpha, v["group_size"], v["nleaf"], v["tol"], v["median_rel_err"], v["p99_rel_err"],
                    v["ms_per_step"], 100.0 * v["n_local"] / max(v["nbc"], 1)))
            plain.append(v)
    best_plain = min(plain, key=lambda v: v["ms_per_step"])
    learned_variants = []
    if not args.no_learned:
        for lq in ((1, 0) if args.learned_quad < 0 else (args.learned_quad,)):
            tr = learned_model(lq)
            eng.upload(m, x)
            v = variant(dict(eps=args.eps, usequad=lq, group_size=args.group, nleaf=args.nleaf, learned=1),
                        args.learned_tol)
            v["training"] = tr
            log("learned walk usequad=%d: group %d nleaf %d tol=%.4f median=%.3g p99=%.3g  %.3f ms/step (MLP kernel "
                "%.3f ms, %d evaluations)" % (lq, v["group_size"], v["nleaf"], v["tol"], v["median_rel_err"],
                                              v["p99_rel_err"], v["ms_per_step"], v["t_mlp_kernel_ms"], v["n_mlp"]))
            learned_variants.append(v)
            lalphas = [] if lq or args.learned_local.strip().lower() in ("", "none") else \
                [float(a) for a in args.learned_local.split(",")]
            for alpha in lalphas:
                v = variant(dict(eps=args.eps, usequad=0, group_size=args.group, nleaf=args.nleaf, learned=1,
                                 local_alpha=alpha), args.learned_tol)
                v["training"] = tr
                log("learned walk + local expansions (alpha %.3f): group %d nleaf %d tol=%.4f median=%.3g p99=%.3g  "
                    "%.3f ms/step (MLP kernel %.3f ms, %d evaluations; %.1f %% of the cell interactions by "
                    "expansion)" % (alpha, v["group_size"], v["nleaf"], v["tol"], v["median_rel_err"],
                                    v["p99_rel_err"], v["ms_per_step"], v["t_mlp_kernel_ms"], v["n_mlp"],
                                    100.0 * v["n_local"] / max(v["nbc"], 1)))
                learned_variants.append(v)
    learned = min(learned_variants, key=lambda v: v["ms_per_step"]) if learned_variants else None
    learned_train = learned["training"] if learned else None
    # value: the faster of the two at the target median error (both are in
    # config, with their ratio)
    head_v = learned if learned is not None and learned["evals_per_s"] > best_plain["evals_per_s"] else best_plain
    t_walk = head_v["t_walk_kernel_ms"] * 1e-3
    tile_int = head_v["interactions_per_step"] - head_v["n_local"]     # evaluated one by one
    achieved = FLOP_PER_INTERACTION * tile_int / t_walk / 1e12
    head_local = head_v["local_alpha"] > 0
    out = {
        "metric": METRIC,
        "value": head_v["evals_per_s"],
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head_v["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic Plummer (mkplummer recipe, own counter RNG, seed 123)",
        "config": {
            "workload": "plummer_n%d_%s_matched_median_err" % (n, "learned_vs_plain" if learned else "plain_tree"),
            "n": n, "eps": args.eps, "group_size": args.group or "autotuned per variant",
            "target_median_err": args.target,
            "err_subset": len(idx),
            "value_is": "learned" if head_v is learned else "plain",
            "learned": learned, "learned_training": learned_train, "learned_variants": learned_variants,
            "plain": best_plain, "plain_variants": plain,
            "learned_over_plain": (learned["evals_per_s"] / best_plain["evals_per_s"]) if learned else None,
        },
        "roofline": {
            "bound": "valu",
            "compute_unit": "FP32 vector ALU (the walk issues no MFMA)",
            "kernel": (("k_walk_sg<%d,%s,false>" % (head_v["group_size"] // 64, "true" if head_v["usequad"] else "false"))
                       if head_v.get("supergroup", 1) > 1 and not head_v["learned"] and not head_local else
                       "k_walk<%d,%s,false,%s,false,%s>" % (head_v["group_size"] // 64,
                                                            "true" if head_v["usequad"] else "false",
                                                            "true" if head_v["learned"] else "false",
                                                            "true" if head_local else "false")),
            "achieved": achieved,
            "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP32_PEAK_TFLOPS,
            "traffic": None,
            "flop_convention": "20 flop per interaction evaluated one by one (P2P and P2C, counted by the "
                               "kernel; cell interactions taken by group local expansions are not counted)",
            "interactions_evaluated": int(tile_int),
            "interactions_by_expansion": int(head_v["n_local"]),
            # the rate a walk evaluating every interaction one by one would need
            # to match this one (not an achieved flop rate)
            "equivalent_tflops_all_interactions": FLOP_PER_INTERACTION * head_v["interactions_per_step"] / t_walk / 1e12,
        },
    }
    traffic, src = profiled_traffic(out["roofline"]["kernel"])
    if traffic is not None:
        out["roofline"]["traffic"] = traffic
        out["roofline"]["traffic_source"] = src
    if learned and learned["t_mlp_kernel_ms"] > 0:
        a_mlp = FLOP_PER_MLP * learned["n_mlp"] / (learned["t_mlp_kernel_ms"] * 1e-3) / 1e12
        out["roofline_mlp"] = {"bound": "mfma", "kernel": "k_learned_group", "achieved": a_mlp,
                               "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": a_mlp / BF16_PEAK_TFLOPS,
                               "traffic": None,
                               "flop_convention": "19,072 flop per MLP evaluation (12-64-64-64-9); the kernel also "
                                                  "applies each group's correction to its members"}
    if not args.no_cpu:
        nt = host_threads()
        cb = cpu_baseline(m, x, idx, dacc, args.eps, args.target, args.cpu_budget,
                          nthreads_list=(1, nt) if nt > 1 else (1,))
        out["cpu_baseline"] = cb[1]
        if nt > 1:
            out["cpu_baseline_allcores"] = cb[nt]
    eng.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
