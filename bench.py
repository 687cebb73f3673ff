"""Benchmark: exact Brandes (GTEPS) and PageRank (GB/s) on one MI355X.

Default run (`--workload all`): the headline workload, rmat20 -- R-MAT scale
20 (1,048,576 nodes, 2^24 distinct arcs, a=.57 b=.19 c=.19 d=.05, seeded),
exact unweighted Brandes from 4,096 seeded sources swept 256 per launch,
graph and sources resident in HBM, output left in HBM; PageRank (FP32 pull
SpMV with Kahan row sums, tol 1e-6 L1, at most 100 iterations) timed on the
same resident graph -- then, under "workloads", a shorter run of each other
workload with its own roofline, traffic and CPU baseline:
  er10k     directed Erdos-Renyi n=10,000, avg out-degree 8, all sources;
            PageRank tol 1e-6
  grid2048  2048x2048 grid with 5% of edges removed, 1,024 seeded sources
            (per-source queue engine); PageRank to 1e-8 L1 in the mixed
            FP32/FP64 mode, at most 1,000 iterations
  ba512     512 undirected Barabasi-Albert graphs (m=4, node counts
            log-uniform in [1,000, 100,000]), all sources of every graph in
            one batched resource
  rmat20w   rmat20's graph with f64 arc weights uniform in (0, 1], weighted
            Brandes (betweennessF64) from the 4,096 sources
  power     eigenvector, Katz, alpha centrality on rmat20's graph
Other single workloads: pivots (learned vs uniform pivots for sampled
Brandes), copy (1 GiB device copy and a 1 GiB streaming sum: the
known-size kernels that calibrate the FETCH_SIZE / WRITE_SIZE counters).

Prints one JSON line.  Units and byte models:
  Brandes GTEPS   = sum over sources of arcs scanned by the forward sweep
                    (m_s) / seconds / 1e9
  64-lane engine  (rmat20, er10k, ba512; 64 sources share one row access)
                  per step, from the sweep's own counters: R row visits
                  (group, level >= 1, vertex on that level for some source of
                  the group), A_R arcs of the visited rows, N lane-vertices
                  reached (includes the S sources), T per-lane tree arcs
                  (depth(w) = depth(v) + 1 >= 2), S sources with A_0 arcs:
                    backward = 36 R + 12 A_R + 16 (N - S) + 8 T
                      (per visit: vertex id 4, lane mask 8, row bounds 8,
                       score read-modify-write 16; per arc: head id 4, the
                       head's level lane mask 8; per lane-vertex: path count
                       read 8, dependency term write 8; per lane tree arc:
                       the child's dependency term 8)
                    forward  = 28 R + 20 S + 12 (A_R + A_0) + 8 N + 8 (T + A_0)
                      (per frontier visit: id 4, lane mask 8, row bounds 8,
                       plus the next-frontier mask write 8 for levels >= 1;
                       per arc: head id 4, head visited mask 8; per
                       lane-vertex: its path count written 8; per lane tree
                       arc, the sources' arcs included: the tail's count
                       carried 8)
  queue engine    (grid2048; one source per access) per source, SURVEY's
                  per-source model: 40 m_s + 64 n_s, forward 16 m_s + 24 n_s,
                  backward 24 m_s + 40 n_s
  PageRank GB/s   = iterations * (8 E + 20 n) bytes / seconds
"""
import argparse
import ctypes
import gc
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def roofline_entry(bytes_per_step, ms, launches, kernel, traffic=None):
    """Algorithmic bytes of one step over the kernel's summed device time
    (HIP events on the launch stream); traffic = counter-measured HBM bytes
    per launch, when a traffic file has them."""
    achieved = bytes_per_step / (ms / 1000.0) / 1e9 if ms > 0 else 0.0
    r = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
         "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": kernel, "launches": launches,
         "avg_launch_us": round(ms * 1000.0 / max(launches, 1), 2),
         "algorithmic_bytes_per_launch": round(bytes_per_step / max(launches, 1))}
    if traffic:
        mgbs = traffic * launches / (ms / 1000.0) / 1e9
        r["measured_hbm_GBs"] = round(mgbs, 1)
        r["measured_frac"] = round(mgbs / HBM_PEAK_GBS, 4)
        r["traffic_over_algorithmic"] = round(traffic / max(bytes_per_step / max(launches, 1), 1), 3)
    return r


def traffic_of(workload, kernel):
    """HBM bytes per dispatch (or per unit of work, where the file gives units) of `kernel` from
    profiles/traffic_<workload>.json."""
    tfile = os.path.join(ROOT, "profiles", f"traffic_{workload}.json")
    if not os.path.exists(tfile):
        return None
    with open(tfile) as fh:
        tj = json.load(fh)
    # "a+b": the kernels' bytes over their dispatches together; a name
    # without its block-shape argument ("pr_iter_kernel<float, 0>") stands
    # for the shape that ran most often
    def entry(k):
        if k in tj:
            return tj[k]
        stem = k[:-1] + ", " if k.endswith(">") else None
        cands = [v for kk, v in tj.items() if stem and kk.startswith(stem) and isinstance(v, dict)]
        return max(cands, key=lambda v: v["dispatches"]) if cands else None

    parts = [entry(k) for k in kernel.split("+")]
    if not all(parts):
        return None
    # per unit of work where the file says how many units the dispatches
    # served (PageRank: iterations), else per dispatch
    n = sum(p.get("units", p["dispatches"]) for p in parts)
    return round(sum(p["hbm_bytes_per_dispatch"] * p["dispatches"] for p in parts) / max(n, 1))


def lane_model(info, n_sources, arcs0):
    """(forward, backward) algorithmic bytes of one step of the 64-lane engine."""
    R, AR, N, T = info["row_visits"], info["row_arcs"], info["nodes_reached"], info["tree_arcs"]
    S = n_sources
    bwd = 36 * R + 12 * AR + 16 * (N - S) + 8 * T
    fwd = 28 * R + 20 * S + 12 * (AR + arcs0) + 8 * N + 8 * (T + arcs0)
    return fwd, bwd


def info_dict(inf):
    from zog_amd import _lib
    return {k: getattr(inf, k) for k, _ in _lib.ZgBcInfo._fields_}


def cleanup():
    import torch
    gc.collect()
    torch.cuda.synchronize()
    torch.cuda.empty_cache()


def bench_copy(args):
    """A plain 1 GiB device-to-device copy and a 1 GiB streaming sum: the
    known-size kernels that scale the FETCH_SIZE / WRITE_SIZE counters to HBM
    bytes (the sum checks the copy's read factor on a kernel that only reads)."""
    import torch
    torch.cuda.set_device(0)
    nbytes = 1 << 30
    a = torch.empty(nbytes, dtype=torch.uint8, device="cuda").fill_(1)
    b = torch.empty_like(a)
    x = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    for _ in range(args.warmup):
        b.copy_(a)
        x.sum()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        b.copy_(a)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / max(args.steps, 1)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        x.sum()
    torch.cuda.synchronize()
    ds = (time.perf_counter() - t0) / max(args.steps, 1)
    return {"metric": "device copy", "value": round(2 * nbytes / dt / 1e9, 1), "unit": "GB/s",
            "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(dt * 1e3, 3),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u8",
            "data": "synthetic", "config": {"workload": "copy: 1 GiB read + 1 GiB written per step; "
                                                        "sum: 1 GiB of FP32 read per step"},
            "sum_GBs": round(nbytes / ds / 1e9, 1)}


def bench_pivots(args):
    """Learned vs uniform pivots for sampled Brandes (see zog_amd.pivots).

    Samplers, each judged against exact Brandes over every source by top-1%
    overlap and Kendall tau at pivot budgets k = 2048, 4096, ...:
      uniform          k sources uniformly from all live vertices (the plain
                       method), weights n/k
      uniform_nonsink  uniformly from the vertices with a successor
      uniform_folded   uniformly from the folded frame (pivots.sampling_frame:
                       single-successor sources folded exactly into their
                       successor's sweep, sources that add nothing left out)
      learned_folded   the folded frame with inclusion probabilities from the
                       trained network's predicted dependency totals
    The value is (sources uniform needs for overlap >= 0.95 on every sample)
    / (sources learned_folded needs)."""
    import numpy as np
    import torch

    from zog_amd import _lib
    from zog_amd import generators as gen
    from zog_amd import pivots as PV
    from zog_amd import resource_graph as rg
    from zog_amd import workloads as W
    from zog_amd.soa import SoA

    torch.cuda.set_device(0)
    sc = PV.Scorer(seed=7)
    t_label = t_train = 0.0
    train_rows = None
    losses = [None]
    if args.model_cache and os.path.exists(args.model_cache):
        with np.load(args.model_cache) as z:
            sc.params = z["params"]
            train_rows = int(z["rows"])
            losses = [float(z["loss"])]
        log(f"[bench] network parameters read from {args.model_cache}")
    else:
        t0 = time.time()
        X, T = PV.rmat_training_set(per_scale=args.train_per_scale, log=log)
        t_label = time.time() - t0
        log(f"[bench] {X.shape[0]} training rows labelled in {t_label:.1f}s")
        p = sc.params
        p[-1] = float(T.mean())  # output bias starts at the mean target
        sc.params = p
        t0 = time.time()
        losses = sc.fit(X, T, epochs=args.epochs, batch=_lib.ZG_MLP_MAX_BATCH, lr=1e-3, log=log)
        t_train = time.time() - t0
        train_rows = int(X.shape[0])
        del X, T
        if args.model_cache:
            np.savez(args.model_cache, params=sc.params, rows=train_rows, loss=losses[-1])
    scale = args.eval_scale
    n, src, dst = gen.rmat(scale, 16 << scale, W.SEEDS["rmat22"] if scale == 22 else W.SEEDS["rmat20"] + scale)
    g = rg.new(SoA.from_arrays("directed", n, src, dst))
    V, n_live, E = g.resource.info()
    live_ids = np.arange(V)
    log(f"[bench] judged graph R-MAT {scale} built: {V} nodes, {E} arcs")
    t0 = time.time()
    if args.exact_cache and os.path.exists(args.exact_cache):
        with np.load(args.exact_cache) as z:
            exact, true_tot, t_exact = z["exact"], z["totals"], float(z["seconds"])
        log(f"[bench] exact scores read from {args.exact_cache} (computed in {t_exact:.1f}s)")
    else:
        exact = np.zeros(V)
        true_tot = np.zeros(V)
        arcs_done = 0
        sl = 1 << 18
        for a in range(0, V, sl):
            info = {}
            ids = live_ids[a:a + sl].astype(np.uint32)
            b, tt = PV.bc_weighted(g, ids, None, want_bc=True, want_totals=True, info=info)
            exact += b
            true_tot[a:a + sl] = tt
            arcs_done += info["arcs_traversed"]
            log(f"[bench] exact Brandes: {min(a + sl, V)}/{V} sources, {time.time() - t0:.1f}s")
        t_exact = time.time() - t0
        log(f"[bench] exact Brandes on R-MAT {scale}: {t_exact:.1f}s, {arcs_done / t_exact / 1e9:.1f} GTEPS")
        if args.exact_cache:
            np.savez_compressed(args.exact_cache, exact=exact, totals=true_tot, seconds=t_exact)
    if args.exact_only:
        rg.destroy(g)
        sc.close()
        return {"metric": "exact Brandes scores for the pivot judge", "value": round(t_exact, 1), "unit": "s",
                "n_gpus": 1, "steps": 1, "warmup": 0, "ms_per_step": round(t_exact * 1000, 1),
                "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
                "data": "synthetic", "config": {"workload": f"R-MAT scale {scale}, exact Brandes over all sources"}}
    pred, _ = PV.predict_totals(g, sc)
    from scipy.stats import spearmanr
    rho = float(spearmanr(pred, true_tot).statistic)
    feats = PV.node_features(g)
    sc.forward(feats)
    reps = 20
    t0 = time.perf_counter()
    for _ in range(reps):
        sc.forward(feats)
    mlp_s = (time.perf_counter() - t0) / reps
    flops = 2.0 * V * (8 * 128 + 128 * 128 + 128)
    rng = np.random.default_rng(W.SEEDS["rmat20"] + 99)
    frame = PV.sampling_frame(g)
    nonsink = np.nonzero(np.bincount(src, minlength=V) > 0)[0]
    if args.ks:
        ks = [int(x) for x in args.ks.split(",")]
    else:
        ks = []
        k = 2048
        while k < min(n_live, args.k_max):
            ks.append(k)
            k *= 2
    samplers = args.samplers.split(",")
    curves = {m: [] for m in samplers}
    for method in samplers:
        for k in ks:
            ov, taus, used = [], [], []
            for r in range(args.reps):
                if method == "uniform":
                    est, m = PV.sampled_betweenness(g, k, rng, None, live_ids, log=log)
                elif method == "uniform_nonsink":
                    est, m = PV.sampled_betweenness(g, k, rng, None, nonsink, log=log)
                elif method == "learned":
                    est, m = PV.sampled_betweenness(g, k, rng, pred, live_ids, log=log)
                elif method == "uniform_folded":
                    est, m = PV.sampled_betweenness_folded(g, k, rng, None, frame, log=log)
                elif method == "learned_folded":
                    est, m = PV.sampled_betweenness_folded(g, k, rng, pred, frame, log=log)
                else:
                    raise SystemExit(f"unknown sampler {method}")
                ov.append(PV.top_overlap(est, exact))
                used.append(m)
                if r == 0:
                    taus.append(PV.kendall_tau(est, exact))
            curves[method].append({"k": k, "sources": float(np.mean(used)), "top1_overlap": float(np.mean(ov)),
                                   "top1_overlap_min": float(np.min(ov)), "kendall_tau": taus[0]})
            log(f"[bench] {method} k={k}: overlap {np.mean(ov):.4f} (min {np.min(ov):.4f}) tau {taus[0]:.4f}")

    def needed(c):
        for row in c:
            if row["top1_overlap_min"] >= 0.95:
                return row["sources"]
        return None

    need = {m: needed(curves[m]) for m in samplers}
    best = "learned_folded" if "learned_folded" in need else samplers[-1]
    nu, nl = need.get("uniform"), need.get(best)
    ratio = (nu / nl) if (nu and nl) else None
    line = {
        "metric": "learned-pivot sampled Brandes: sources to reach top-1% overlap >= 0.95, uniform / learned",
        "value": round(ratio, 3) if ratio else None, "unit": "x", "n_gpus": 1, "steps": 1, "warmup": 0,
        "ms_per_step": round(t_exact * 1000, 1), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "bf16-mfma/f64", "data": "synthetic",
        "config": {"workload": f"pivots: train on {args.train_per_scale} R-MAT graphs per scale 14..18, "
                               f"judge on R-MAT scale {scale} (exact Brandes over all sources)",
                   "eval_nodes": V, "eval_arcs": E, "train_rows": train_rows, "frame": int(len(frame[0])),
                   "folded": int(frame[1].sum()), "nonsink": int(len(nonsink)),
                   "epochs": args.epochs, "reps": args.reps},
        "learned": {"sources_needed": need, "ratio_uniform_over": best,
                    "overlap": "share of the top-1% estimates whose exact BC reaches the exact top-1% threshold",
                    "curves": curves, "spearman_pred_vs_exact_totals": rho, "final_train_loss": losses[-1],
                    "label_s": round(t_label, 1), "train_s": round(t_train, 1), "exact_s": round(t_exact, 1)},
        "roofline": {"bound": "mfma", "achieved": round(flops / mlp_s / 1e12, 3), "peak": 2500.0,
                     "unit": "TFLOP/s", "frac": round(flops / mlp_s / 1e12 / 2500.0, 5), "traffic": None,
                     "kernel": "mlp_forward_kernel", "avg_call_us": round(mlp_s * 1e6, 1),
                     "note": "wall time per zg_mlp_forward call (launch and synchronisation included)"},
    }
    rg.destroy(g)
    sc.close()
    return line


def bench_weighted(args, steps, warmup):
    """Weighted Brandes (zg_bc_f64) on rmat20's graph and 4,096 sources with
    seeded arc weights uniform in (0, 1]."""
    import numpy as np
    import torch

    from zog_amd import _lib
    from zog_amd import resource_graph as rg
    from zog_amd import workloads as W
    from zog_amd.soa import SoA
    torch.cuda.set_device(0)
    t0 = time.time()
    n, src, dst = W.rmat20()
    w = W.rmat20w_weights(len(src))
    sources = W.rmat20_sources(n, src, dst, 4096)
    log(f"[bench] rmat20w generated in {time.time() - t0:.1f}s: n={n} arcs={len(src)} sources={len(sources)}")
    g = rg.new(SoA.from_arrays("directed", n, src, dst, w))
    V, n_live, E = g.resource.info()
    L = _lib.lib()
    h = g.resource.handle
    out = np.zeros(V, dtype=np.float64)
    info = _lib.ZgBcInfo()

    def step(srcs=sources):
        rc = L.zg_bc_f64(h, ctypes.c_void_p(srcs.ctypes.data), len(srcs), ctypes.c_void_p(out.ctypes.data),
                         ctypes.byref(info))
        _lib.check(rc, "zg_bc_f64")

    def set_delta(d):
        _lib.check(L.zg_set_tuning(h, _lib.ZG_TUNE_WEIGHTED_DELTA, float(d)), "zg_set_tuning")

    tuned = None
    if args.wbatch is not None:
        _lib.check(L.zg_set_tuning(h, _lib.ZG_TUNE_WEIGHTED_BATCH, float(args.wbatch)), "zg_set_tuning")
    if args.wdelta is not None:
        set_delta(args.wdelta)
    if not args.no_autotune and args.wdelta is None:
        # bound step of the near frontier (0 = the library's default), on a
        # quarter of the sources
        sub = np.ascontiguousarray(sources[:1024])
        t_best = None
        for d in (0.0, 0.01, 0.04, float("inf")):
            set_delta(d)
            step(sub)
            log(f"[bench] tune: delta {d}: {info.kernel_ms:.1f} ms for {len(sub)} sources")
            if t_best is None or info.kernel_ms < t_best[0]:
                t_best = (info.kernel_ms, d)
        set_delta(t_best[1])
        tuned = {"delta": t_best[1] if t_best[1] != float("inf") else "inf"}

    for i in range(warmup):
        step()
        log(f"[bench] rmat20w warmup {i} done ({info.kernel_ms:.1f} ms)")
    torch.cuda.synchronize()
    w0 = time.perf_counter()
    arcs = relaxed = reached = 0
    for i in range(steps):
        step()
        arcs += info.arcs_traversed
        relaxed += info.arcs_relaxed
        reached += info.nodes_reached
        log(f"[bench] rmat20w step {i} done ({info.kernel_ms:.1f} ms)")
    torch.cuda.synchronize()
    wall = time.perf_counter() - w0
    K = steps
    arcs_step, relaxed_step, nodes_step = arcs // K, relaxed // K, reached // K
    rg.set_profiling(g, True)
    step()
    prof = rg.get_profile(g)
    rg.set_profiling(g, False)
    log(f"[bench] profile {prof}")
    # algorithmic bytes per pass: an arc is its head id (4 B), its weight
    # (8 B) and the head's distance (8 B); a vertex its row bounds (8 B), its
    # distance (8 B) and, in the path-count and dependency passes, its sigma
    # or c value (8 B)
    kb = {"wbc_relax_kernel+wbc_hub_kernel+wbc_split_kernel": (20 * relaxed_step, prof["bc_forward"]),
          "wbc_count_kernel+wbc_peel_kernel": (40 * arcs_step + 48 * nodes_step, prof["bc_dag"]),
          "wbc_backward_kernel": (20 * arcs_step + 24 * nodes_step, prof["bc_backward"])}
    dom = max(kb, key=lambda k: kb[k][1][0])
    dom_bytes, (dom_ms, dom_n) = kb[dom]
    roofline = roofline_entry(dom_bytes, dom_ms, dom_n, dom, traffic_of("rmat20w", dom))
    roofline["kernel_ms_per_step"] = {k: round(v[0], 3) for k, v in prof.items() if v[1]}
    cpu = None
    parity = {}
    if not args.no_cpu:
        from oracle import oracle as O
        og = O.Graph(n, src, dst)
        og.set_weights(w)
        sample = sources[:6]
        t0 = time.perf_counter()
        ref = og.betweenness_f64(sources=sample)
        cpu_s = time.perf_counter() - t0
        got = rg.betweenness_f64_array(g, sources=sample, info=(ginf := {}))
        cpu_gteps = ginf["arcs_traversed"] / cpu_s / 1e9
        scale = max(1.0, float(np.abs(ref).max()))
        parity[f"bc_max_rel_err_{len(sample)}_sources"] = float(
            (np.abs(got - ref) / np.maximum(np.abs(ref), 1e-9 * scale)).max())
        log(f"[bench] oracle weighted BC {len(sample)} sources in {cpu_s:.1f}s = {cpu_gteps:.4f} GTEPS")
        cpu = {"value": round(cpu_gteps, 5), "unit": "GTEPS", "cores": 1, "kind": "port",
               "sample": f"oracle/oracle.c weighted Brandes (restated centrality.zig:529-667, binary heap), "
                         f"1 thread, {len(sample)} of the {len(sources)} sources, host cores beside the GPU"}
    rg.destroy(g)
    return {
        "metric": "weighted Brandes GTEPS (exact betweennessF64)",
        "value": round(arcs / wall / 1e9, 3), "unit": "GTEPS", "n_gpus": 1, "steps": K, "warmup": warmup,
        "ms_per_step": round(wall * 1000.0 / K, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "rmat20w: R-MAT scale 20, 2^24 distinct arcs, seeded f64 weights uniform in (0, 1], "
                               "exact weighted Brandes from 4096 seeded sources",
                   "nodes": n, "arcs": int(E), "sources": len(sources), "arcs_per_step": arcs_step,
                   "arcs_relaxed_per_step": relaxed_step, "relax_rounds": int(info.push_levels),
                   "batches": int(info.batches), "autotuned": tuned},
        "roofline": roofline, "cpu_baseline": cpu, "parity": parity,
    }


def bench_power(args, steps):
    """Eigenvector, Katz and alpha centrality on rmat20's graph: 100
    iterations each (tolerance 0 never stops early), FP64."""
    import numpy as np
    import torch

    from zog_amd import _lib
    from zog_amd import resource_graph as rg
    from zog_amd import workloads as W
    from zog_amd.soa import SoA
    torch.cuda.set_device(0)
    n, src, dst = W.rmat20()
    g = rg.new(SoA.from_arrays("directed", n, src, dst))
    apply_tuning(g, args)
    V, n_live, E = g.resource.info()
    iters = 100
    # alpha below 1 / spectral radius keeps Katz and alpha finite: the
    # largest in-degree bounds the radius from below, the arc count above
    alpha = 1.0 / float(E)
    runs = {"eigenvector": lambda inf: rg.eigenvector(g, iters, 0.0, raw=True, info=inf),
            "katz": lambda inf: rg.katz(g, alpha, 1.0, iters, 0.0, raw=True, info=inf),
            "alpha_centrality": lambda inf: rg.alpha_centrality(g, alpha, 1.0, iters, 0.0, raw=True, info=inf)}
    # per iteration: predecessor id (4 B) and gathered value (8 B) per arc;
    # row descriptor (16 B), old and new value (16 B) per row; the
    # eigenvector's normalising pass reads three iterates and writes one
    bytes_iter = {"eigenvector": 12 * E + 64 * n_live, "katz": 12 * E + 32 * n_live,
                  "alpha_centrality": 12 * E + 32 * n_live}
    res = {}
    for name, fn in runs.items():
        inf = {}
        fn(inf)  # warm-up
        times = []
        for _ in range(steps):
            fn(inf)
            times.append(inf["kernel_ms"])
            log(f"[bench] {name}: {inf['iterations']} iterations in {inf['kernel_ms']:.1f} ms")
        ms = float(np.median(times))
        rg.set_profiling(g, True)
        fn(inf)
        prof = rg.get_profile(g)
        rg.set_profiling(g, False)
        it_ms, it_n = prof["pr_iter"]
        gbs = bytes_iter[name] * inf["iterations"] / (ms / 1000.0) / 1e9
        rl = roofline_entry(bytes_iter[name] * it_n, it_ms, it_n, "pr_iter_kernel" +
                            ("+pr_finish_kernel+eig_norm_kernel+eig_decide_kernel" if name == "eigenvector" else ""),
                            traffic_of("power", "pr_iter_kernel<double, 1>" if name != "eigenvector"
                                       else "pr_iter_kernel<double, 2>"))
        res[name] = {"value": round(gbs, 1), "unit": "GB/s", "iterations": int(inf["iterations"]),
                     "ms": round(ms, 2), "bytes_per_iteration": bytes_iter[name], "roofline": rl}
    floor_us = ctypes.c_double()
    _lib.check(_lib.lib().zg_gather_floor(g.resource.handle, 64, 20, ctypes.byref(floor_us)), "zg_gather_floor")
    floor_lds_us = ctypes.c_double()
    _lib.check(_lib.lib().zg_gather_floor(g.resource.handle, 64 | _lib.ZG_FLOOR_LDS_COPY, 20,
                                          ctypes.byref(floor_lds_us)), "zg_gather_floor")
    for name in res:
        it_us = res[name]["roofline"]["avg_launch_us"]
        res[name]["roofline"]["gather_floor_us"] = round(floor_us.value, 2)
        res[name]["roofline"]["frac_of_gather_floor"] = round(floor_us.value / it_us, 3) if it_us > 0 else None
        res[name]["roofline"]["gather_floor_lds_us"] = round(floor_lds_us.value, 2)
        res[name]["roofline"]["frac_of_gather_floor_lds"] = (round(floor_lds_us.value / it_us, 3)
                                                             if it_us > 0 else None)
    cpu = None
    parity = {}
    if not args.no_cpu:
        from oracle import oracle as O
        og = O.Graph(n, src, dst)
        k = 10
        t0 = time.perf_counter()
        ref, _, _ = og.katz(alpha, 1.0, k, 0.0)
        cpu_s = time.perf_counter() - t0
        got = np.array(rg.katz(g, alpha, 1.0, k, 0.0, raw=True))
        parity["katz_max_rel_err_10_iterations"] = float((np.abs(got - ref) / np.abs(ref).max()).max())
        ref, _, _ = og.eigenvector(k, 0.0)
        got = np.array(rg.eigenvector(g, k, 0.0, raw=True))
        parity["eigenvector_max_abs_err_10_iterations"] = float(np.abs(got - ref).max())
        cpu = {"value": round(bytes_iter["katz"] * k / cpu_s / 1e9, 3), "unit": "GB/s", "cores": 1, "kind": "port",
               "sample": f"oracle/oracle.c Katz (restated centrality.zig:1212-1267), {k} iterations on the same "
                         f"graph, 1 thread, host cores beside the GPU"}
        log(f"[bench] oracle Katz {k} iterations in {cpu_s:.1f}s")
    rg.destroy(g)
    return {
        "metric": "power-iteration GB/s (Katz; eigenvector and alpha beside it)",
        "value": res["katz"]["value"], "unit": "GB/s", "n_gpus": 1, "steps": steps, "warmup": 1,
        "ms_per_step": res["katz"]["ms"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": "power: R-MAT scale 20, 2^24 distinct arcs; eigenvector, Katz (alpha = 1/arcs, "
                               "beta 1) and alpha centrality (alpha = 1/arcs, initial 1), 100 iterations each",
                   "nodes": n, "arcs": int(E)},
        "roofline": res["katz"]["roofline"], "methods": res, "cpu_baseline": cpu, "parity": parity,
    }


def workload_spec(name):
    from zog_amd import workloads as W
    if name == "rmat20":
        return dict(gen=W.rmat20, sources=lambda n, s, d: W.rmat20_sources(n, s, d, 4096),
                    pr=dict(tol=1e-6, max_iter=100, precision=32), cpu_sources=16,
                    desc="rmat20: R-MAT scale 20, 2^24 distinct arcs, exact Brandes from 4096 seeded sources, "
                         "256 per launch; PageRank FP32-Kahan tol 1e-6 max 100")
    if name == "er10k":
        return dict(gen=W.er10k, sources=None, pr=dict(tol=1e-6, max_iter=100, precision=32), cpu_sources=None,
                    desc="er10k: directed Erdos-Renyi n=10,000 p=8/(n-1), exact Brandes over all sources; "
                         "PageRank FP32-Kahan tol 1e-6 max 100")
    if name == "ba512":
        return dict(gen=None, sources=None, pr=None, cpu_sources=8,
                    desc="ba512: 512 undirected Barabasi-Albert graphs (m=4), node counts log-uniform in "
                         "[1,000, 100,000], seeded; exact all-sources Brandes for every graph in one batched resource")
    if name == "grid2048":
        return dict(gen=W.grid2048, sources=lambda n, s, d: W.grid2048_sources(n, s, d, 1024),
                    pr=dict(tol=1e-8, max_iter=1000, precision=96), cpu_sources=2, scaled=True,
                    desc="grid2048: 2048x2048 grid, 419,226 of 8,384,512 edges removed, exact Brandes from 1024 "
                         "seeded sources; PageRank mixed FP32/FP64 to 1e-8 L1, max 1000")
    raise SystemExit(f"unknown workload {name}")


def apply_tuning(g, args):
    from zog_amd import _lib
    from zog_amd import resource_graph as rg
    if args.no_graph:
        _lib.check(_lib.lib().zg_set_tuning(g.resource.handle, _lib.ZG_TUNE_GRAPH, 0.0), "zg_set_tuning")
    if args.pr_cache is not None:
        _lib.check(_lib.lib().zg_set_tuning(g.resource.handle, _lib.ZG_TUNE_PR_CACHE, float(args.pr_cache)),
                   "zg_set_tuning")
    if args.pr_grid is not None:
        _lib.check(_lib.lib().zg_set_tuning(g.resource.handle, _lib.ZG_TUNE_PR_GRID, float(args.pr_grid)),
                   "zg_set_tuning")
    if args.pr_block is not None:
        _lib.check(_lib.lib().zg_set_tuning(g.resource.handle, _lib.ZG_TUNE_PR_BLOCK, float(args.pr_block)),
                   "zg_set_tuning")
    if args.no_lane_order:
        rg.set_lane_order(g, False)
    if args.pull_grid is not None:
        rg.set_pull_grid(g, args.pull_grid)
    if args.alpha is not None:
        rg.set_push_alpha(g, args.alpha)
    if args.order is not None:
        rg.set_source_order(g, args.order)


def set_count_tree(h, on):
    from zog_amd import _lib
    _lib.check(_lib.lib().zg_set_tuning(h, _lib.ZG_TUNE_COUNT_TREE_ARCS, 1.0 if on else 0.0), "zg_set_tuning")


def brandes_roofline(name, prof, model, queue):
    """Roofline of the dominant Brandes kernel (backward, or forward = pull
    and push) under the engine's byte model."""
    fwd_ms = prof["bc_forward"][0] + prof["bc_push"][0]
    fwd_n = prof["bc_forward"][1] + prof["bc_push"][1]
    bwd_ms, bwd_n = prof["bc_backward"]
    fwd_bytes, bwd_bytes = model
    if queue:
        fk, bk = "bcq_forward_kernel", "bcq_backward_kernel"
    else:
        fk, bk = "bc_forward_kernel+bc_push_kernel", "bc_backward_kernel"
    if fwd_ms >= bwd_ms:
        r = roofline_entry(fwd_bytes, fwd_ms, fwd_n, fk, traffic_of(name, fk))
        other = roofline_entry(bwd_bytes, bwd_ms, bwd_n, bk, traffic_of(name, bk))
    else:
        r = roofline_entry(bwd_bytes, bwd_ms, bwd_n, bk, traffic_of(name, bk))
        other = roofline_entry(fwd_bytes, fwd_ms, fwd_n, fk, traffic_of(name, fk))
    r["byte_model"] = "queue engine: per source 16 m_s + 24 n_s forward, 24 m_s + 40 n_s backward" if queue else \
        "64-lane engine: backward 36 R + 12 A_R + 16 (N - S) + 8 T, forward 28 R + 20 S + 12 (A_R + A_0) + 8 N + 8 (T + A_0)"
    r["other_kernel"] = {k: other[k] for k in ("kernel", "achieved", "frac", "traffic", "avg_launch_us")}
    r["kernel_ms_per_step"] = {k: round(v[0], 3) for k, v in prof.items() if v[1]}
    return r


def bench_graph(args, name, steps, warmup):
    """One-graph Brandes workload (rmat20, er10k, grid2048) + PageRank."""
    import numpy as np
    import torch

    from zog_amd import _lib
    from zog_amd import resource_graph as rg
    from zog_amd.soa import SoA

    spec = workload_spec(name)
    torch.cuda.set_device(0)
    dev = torch.device("cuda:0")
    t0 = time.time()
    n, src, dst = spec["gen"]()
    sources = spec["sources"](n, src, dst) if spec["sources"] else np.arange(n, dtype=np.uint32)
    log(f"[bench] {name} generated in {time.time() - t0:.1f}s: n={n} arcs={len(src)} sources={len(sources)}")
    t0 = time.time()
    g = rg.new(SoA.from_arrays("directed", n, src, dst))
    rg.set_batch_groups(g, args.groups)
    apply_tuning(g, args)
    log(f"[bench] resource loaded in {time.time() - t0:.1f}s")
    if name != "grid2048" and not args.no_lane_order:
        # the 64-lane engine's degree-ordered view is built at its first use
        t1 = time.time()
        rg.betweenness_unweighted_array(g, sources=sources[:1])
        log(f"[bench] first sweep (builds the degree-ordered view) {time.time() - t1:.2f}s")
    tuned = None
    if not args.no_autotune and args.pull_grid is None and args.alpha is None and name != "grid2048":
        # rmat20 fixes the batch width (256 sources per launch); er10k sweeps
        # every source, so the tuner also picks the width there, on all of
        # them; otherwise it picks the pull grid, the push/pull switch and the
        # source grouping on two batches' worth of sources
        from zog_amd import autotune
        if name == "er10k":
            res_t = autotune.tune(g.resource, sources, groups=(4, 16, 64), alphas=(8.0,), pull_grids=(4096, 0),
                                  orders=(1, 2), repeats=1, log=log)
        else:
            res_t = autotune.tune(g.resource, sources[: 128 * args.groups], groups=(args.groups,), orders=(1, 2),
                                  log=log)
        tuned = dict(res_t["best"], seconds=round(time.time() - t0, 1))
    V, n_live, E = g.resource.info()
    log(f"[bench] resource built in {time.time() - t0:.1f}s")
    L = _lib.lib()
    h = g.resource.handle
    d_src = torch.from_numpy(sources.astype(np.int32)).to(dev)
    d_bc = torch.empty(V, dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream()
    info = _lib.ZgBcInfo()

    def bc_step():
        rc = L.zg_bc_unweighted_device(h, ctypes.c_void_p(d_src.data_ptr()), len(sources),
                                       ctypes.c_void_p(d_bc.data_ptr()), ctypes.c_void_p(stream.cuda_stream),
                                       ctypes.byref(info))
        _lib.check(rc, "zg_bc_unweighted_device")

    for i in range(warmup):
        bc_step()
        log(f"[bench] {name} warmup {i} done ({info.kernel_ms:.1f} ms)")
    torch.cuda.synchronize()
    w0 = time.perf_counter()
    arcs = 0
    step_kernel_ms = []  # device time of each timed sweep (the library's events)
    for i in range(steps):
        bc_step()
        arcs += info.arcs_traversed
        step_kernel_ms.append(info.kernel_ms)
        log(f"[bench] {name} step {i} done")
    torch.cuda.synchronize()
    wall = time.perf_counter() - w0
    log(f"[bench] {name}: {steps} steps in {wall:.3f}s")
    ms_per_step = wall * 1000.0 / steps
    step_info = info_dict(info)
    queue = step_info["row_visits"] == 0 and step_info["arcs_traversed"] > 0  # the per-source engine ran

    # per-kernel device time of one more step (HIP events around every launch
    # on the launch stream)
    rg.set_profiling(g, True)
    bc_step()
    prof = rg.get_profile(g)
    rg.set_profiling(g, False)
    log(f"[bench] profile {prof}")
    if queue:
        m, ns = step_info["arcs_traversed"], step_info["nodes_reached"]
        model = (16 * m + 24 * ns, 24 * m + 40 * ns)
        counts = {"arcs": m, "lane_vertices": ns}
    else:
        # the tree-arc count: one more (untimed) step with the counter on
        set_count_tree(h, True)
        bc_step()
        set_count_tree(h, False)
        cinfo = info_dict(info)
        assert cinfo["row_visits"] == step_info["row_visits"]
        outdeg = np.bincount(src, minlength=n)
        arcs0 = int(outdeg[sources].sum())
        model = lane_model(cinfo, len(sources), arcs0)
        counts = {"R": cinfo["row_visits"], "A_R": cinfo["row_arcs"], "N": cinfo["nodes_reached"],
                  "T": cinfo["tree_arcs"], "S": len(sources), "A_0": arcs0,
                  "per_source_model_bytes": 40 * cinfo["arcs_traversed"] + 64 * cinfo["nodes_reached"]}
    roofline = brandes_roofline(name, prof, model, queue)
    roofline["work_counts"] = counts
    achieved_step = (model[0] + model[1]) / (ms_per_step / 1000.0) / 1e9
    roofline["whole_step_algorithmic_GBs"] = round(achieved_step, 1)

    # PageRank on the same resident graph
    prs = spec["pr"]
    d_pr = torch.empty(V, dtype=torch.float64, device=dev)
    pinfo = _lib.ZgPrInfo()

    def pr_run(prec):
        rc = L.zg_pagerank_device(h, 0.85, prs["max_iter"], prs["tol"], prec, ctypes.c_void_p(d_pr.data_ptr()),
                                  ctypes.c_void_p(stream.cuda_stream), ctypes.byref(pinfo))
        _lib.check(rc, "zg_pagerank_device")

    fast = prs["precision"] | _lib.ZG_PR_SKIP_RESIDUAL
    pr_run(fast)
    torch.cuda.synchronize()
    pr_reps = 5
    w0 = time.perf_counter()
    iters = 0
    for _ in range(pr_reps):
        pr_run(fast)
        iters += pinfo.iterations
    torch.cuda.synchronize()
    pr_wall = time.perf_counter() - w0
    pr_bytes_iter = 8 * E + 20 * n_live
    pr_gbs = iters * pr_bytes_iter / pr_wall / 1e9
    rg.set_profiling(g, True)
    pr_run(prs["precision"])
    pprof = rg.get_profile(g)
    rg.set_profiling(g, False)
    pr_iters = int(pinfo.iterations)
    it_ms, it_n = pprof["pr_iter"]
    pr_result = d_pr.cpu().numpy()
    floor_us = ctypes.c_double()
    _lib.check(L.zg_gather_floor(h, 32, 20, ctypes.byref(floor_us)), "zg_gather_floor")
    floor_lds_us = ctypes.c_double()
    _lib.check(L.zg_gather_floor(h, 32 | _lib.ZG_FLOOR_LDS_COPY, 20, ctypes.byref(floor_lds_us)), "zg_gather_floor")
    pr_roof = roofline_entry(pr_bytes_iter * it_n, it_ms, it_n, "pr_iter_kernel",
                             traffic_of(name, "pr_iter_kernel<float, 0>" if prs["precision"] == 32
                                        else "pr_iter_kernel<float, 0>+pr_iter_kernel<double, 0>"))
    it_us = pr_roof["avg_launch_us"]
    pr_roof["gather_floor_us"] = round(floor_us.value, 2)
    pr_roof["frac_of_gather_floor"] = round(floor_us.value / it_us, 3) if it_us > 0 else None
    pr_roof["gather_floor_lds_us"] = round(floor_lds_us.value, 2)
    pr_roof["frac_of_gather_floor_lds"] = round(floor_lds_us.value / it_us, 3) if it_us > 0 else None
    log(f"[bench] {name} pagerank {pr_iters} iterations ({int(pinfo.fp64_iterations)} FP64), {pr_gbs:.1f} GB/s wall, "
        f"kernel {pr_roof['achieved']:.1f} GB/s, residual {pinfo.residual_l1:.3e}")

    cpu = None
    parity = {"pagerank_fp64_residual_l1": float(pinfo.residual_l1),
              "pagerank_error_bound_l1": float(pinfo.error_bound_l1)}
    if not args.no_cpu:
        from oracle import oracle as O
        og = O.Graph(n, src, dst)
        sample = sources if spec["cpu_sources"] is None else sources[: spec["cpu_sources"]]
        mode = O.SIGMA_EXP if spec.get("scaled") else O.SIGMA_U64
        t0 = time.perf_counter()
        ref_bc, _, arcs_cpu = og.betweenness(sources=sample, threads=1, return_arcs=True, mode=mode)
        cpu_s = time.perf_counter() - t0
        cpu_gteps = float(arcs_cpu.sum()) / cpu_s / 1e9
        log(f"[bench] {name} oracle BC {len(sample)} sources in {cpu_s:.1f}s = {cpu_gteps:.4f} GTEPS")
        gpu_bc = rg.betweenness_unweighted_array(g, sources=sample)
        scale = max(1.0, float(np.abs(ref_bc).max()))
        rel = np.abs(gpu_bc - ref_bc) / np.maximum(np.abs(ref_bc), 1e-9 * scale)
        parity[f"bc_max_rel_err_{len(sample)}_sources"] = float(rel.max())
        t0 = time.perf_counter()
        ref_pr, ref_it, _ = og.pagerank(0.85, prs["max_iter"], prs["tol"])
        cpu_pr_s = time.perf_counter() - t0
        parity["pagerank_l1_vs_fp64_oracle"] = float(np.abs(pr_result - ref_pr).sum())
        parity["pagerank_iterations_gpu_oracle"] = [pr_iters, ref_it]
        cpu = {"value": round(cpu_gteps, 5), "unit": "GTEPS", "cores": 1, "kind": "port",
               "sample": f"oracle/oracle.c Brandes (restated centrality.zig:694-798), 1 thread, {len(sample)} of "
                         f"the {len(sources)} sources of {name}, host cores beside the GPU",
               "pagerank_gbs": round(ref_it * pr_bytes_iter / cpu_pr_s / 1e9, 3),
               "pagerank_sample": f"full FP64 oracle PageRank on {name} ({ref_it} iterations, 1 thread)"}
        # the parallel mode: the same oracle with the sources split into
        # contiguous chunks over the host's CPU share (the reference's own
        # threaded Brandes, centrality.zig:216-256), on 4x the sample
        pthreads = max(1, min(16, os.cpu_count() or 1))
        psample = sources if spec["cpu_sources"] is None else sources[: 4 * spec["cpu_sources"]]
        t0 = time.perf_counter()
        _, _, parcs = og.betweenness(sources=psample, threads=pthreads, return_arcs=True, mode=mode)
        par_s = time.perf_counter() - t0
        cpu["parallel"] = {"value": round(float(parcs.sum()) / par_s / 1e9, 5), "unit": "GTEPS",
                           "cores": pthreads, "sample": f"{len(psample)} of the sources, source-chunked threads"}
        log(f"[bench] {name} oracle BC {len(psample)} sources on {pthreads} threads in {par_s:.1f}s = "
            f"{cpu['parallel']['value']:.4f} GTEPS")
        og.close()

    line = {
        "metric": "Brandes GTEPS (exact BC) + PageRank GB/s",
        "value": round(arcs / wall / 1e9, 3),
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": round(ms_per_step, 3),
        "step_kernel_ms": ({"min": round(min(step_kernel_ms), 3),
                            "mean": round(sum(step_kernel_ms) / len(step_kernel_ms), 3), "n": len(step_kernel_ms)}
                           if step_kernel_ms else None),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": spec["desc"], "nodes": n, "arcs": int(E), "sources": len(sources), "autotuned": tuned,
                   "engine": "queue" if queue else "64-lane",
                   "sources_per_launch": 64 * (tuned["groups"] if tuned else args.groups),
                   "levels": int(step_info["max_levels"]), "arcs_per_step": int(arcs // steps),
                   "push_levels": int(step_info["push_levels"]), "pull_levels": int(step_info["pull_levels"])},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "pagerank": {"metric": "PageRank GB/s", "value": round(pr_gbs, 2), "unit": "GB/s",
                     "iterations": pr_iters, "fp64_iterations": int(pinfo.fp64_iterations),
                     "dtype": {32: "f32 (Kahan rows, f64 reductions)", 64: "f64",
                               96: "f32 Kahan then f64"}[prs["precision"]],
                     "roofline": pr_roof},
        "parity": parity,
    }
    rg.destroy(g)
    return line


def bench_batch(args, steps, warmup):
    """Batched uneven graphs: one resource, all sources of every graph."""
    import numpy as np
    import torch

    from zog_amd import _lib
    from zog_amd import generators as gen
    from zog_amd import resource_graph as rg
    from zog_amd import workloads as W

    spec = workload_spec("ba512")
    torch.cuda.set_device(0)
    dev = torch.device("cuda:0")
    t0 = time.time()
    graphs = gen.ba_batch(512, 4, 1000, 100000, W.SEEDS["ba512"])
    n_tot = sum(g[0] for g in graphs)
    arcs_tot = sum(len(g[1]) for g in graphs)
    log(f"[bench] ba512 generated in {time.time() - t0:.1f}s: {len(graphs)} graphs, {n_tot} nodes, {arcs_tot} arcs")
    t0 = time.time()
    res = rg.BatchResource(graphs)
    L = _lib.lib()
    h = res.handle
    rc = L.zg_set_batch_groups(h, args.groups)
    if rc < 0:
        _lib.check(rc, "zg_set_batch_groups")
    gview = rg.ResourceGraph(resource=res, builder=None)
    apply_tuning(gview, args)
    log(f"[bench] batch resource built in {time.time() - t0:.1f}s")
    tuned = None
    if not args.no_autotune and args.pull_grid is None and args.alpha is None:
        from zog_amd import autotune
        t0 = time.time()
        free, _ = torch.cuda.mem_get_info()
        res_t = autotune.tune(res, autotune.batch_sample(res.offsets), groups=(32, 64), alphas=(8.0, 16.0),
                              pull_grids=(4096,), orders=(2,), repeats=1, max_state_bytes=0.5 * free,
                              span_vertices=int(np.diff(res.offsets).max()), log=log)
        tuned = dict(res_t["best"], seconds=round(time.time() - t0, 1))
    d_bc = torch.empty(n_tot, dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream()
    info = _lib.ZgBcInfo()

    def bc_step():
        rc = L.zg_bc_unweighted_device(h, None, 0, ctypes.c_void_p(d_bc.data_ptr()),
                                       ctypes.c_void_p(stream.cuda_stream), ctypes.byref(info))
        _lib.check(rc, "zg_bc_unweighted_device")

    for i in range(warmup):
        bc_step()
        log(f"[bench] ba512 warmup {i} done")
    torch.cuda.synchronize()
    w0 = time.perf_counter()
    arcs = 0
    for i in range(steps):
        bc_step()
        arcs += info.arcs_traversed
        log(f"[bench] ba512 step {i} done ({info.batches} batches, {info.max_levels} levels max)")
    torch.cuda.synchronize()
    wall = time.perf_counter() - w0
    step_info = info_dict(info)
    # the profiled step also counts tree arcs (their counter adds one atomic
    # per backward work entry, so the backward time is, if anything, high)
    rg.set_profiling(gview, True)
    set_count_tree(h, True)
    bc_step()
    set_count_tree(h, False)
    prof = rg.get_profile(gview)
    rg.set_profiling(gview, False)
    cinfo = info_dict(info)
    log(f"[bench] profile {prof}")
    outdeg = np.concatenate([np.bincount(g[1], minlength=g[0]) for g in graphs])
    model = lane_model(cinfo, n_tot, int(outdeg.sum()))
    roofline = brandes_roofline("ba512", prof, model, False)
    roofline["work_counts"] = {"R": cinfo["row_visits"], "A_R": cinfo["row_arcs"], "N": cinfo["nodes_reached"],
                               "T": cinfo["tree_arcs"], "S": n_tot, "A_0": int(outdeg.sum()),
                               "per_source_model_bytes": 40 * cinfo["arcs_traversed"] + 64 * cinfo["nodes_reached"]}
    out = d_bc.cpu().numpy()
    parity = {}
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as O
        offs = res.offsets
        small = sorted(range(len(graphs)), key=lambda i: graphs[i][0])[:3]
        worst = 0.0
        for i in small:
            n, s, dd = graphs[i]
            ref, _ = O.Graph(n, s, dd).betweenness()
            got = out[offs[i]: offs[i + 1]]
            scale = max(1.0, float(np.abs(ref).max()))
            worst = max(worst, float((np.abs(got - ref) / np.maximum(np.abs(ref), 1e-9 * scale)).max()))
        parity["bc_max_rel_err_3_smallest_graphs_all_sources"] = worst
        rng = np.random.default_rng(5)
        pick = rng.choice(len(graphs), 16, replace=False)
        t_cpu = 0.0
        arcs_cpu = 0
        for i in pick:
            n, s, dd = graphs[i]
            og = O.Graph(n, s, dd)
            srcs = np.arange(0, n, max(1, n // spec["cpu_sources"]), dtype=np.uint32)[: spec["cpu_sources"]]
            t1 = time.perf_counter()
            _, _, a_s = og.betweenness(sources=srcs, return_arcs=True)
            t_cpu += time.perf_counter() - t1
            arcs_cpu += int(a_s.sum())
        cpu = {"value": round(arcs_cpu / t_cpu / 1e9, 5), "unit": "GTEPS", "cores": 1, "kind": "port",
               "sample": "oracle/oracle.c Brandes, 1 thread, 8 sources in each of 16 random graphs of the batch"}
    res.release()
    return {
        "metric": "Brandes GTEPS (exact BC, batched graphs)",
        "value": round(arcs / wall / 1e9, 3), "unit": "GTEPS", "n_gpus": 1, "steps": steps, "warmup": warmup,
        "ms_per_step": round(wall * 1000.0 / steps, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": spec["desc"], "graphs": len(graphs), "nodes": n_tot, "arcs": arcs_tot,
                   "sources": n_tot, "sources_per_launch": 64 * (tuned["groups"] if tuned else args.groups),
                   "autotuned": tuned, "batches": int(step_info["batches"]),
                   "span_balance": round(step_info["span_balance"], 4),
                   "idle_lane_share": round(1.0 - n_tot / max(step_info["group_slots"], 1), 5),
                   "arcs_per_step": int(arcs // steps), "levels": int(step_info["max_levels"]),
                   "push_levels": int(step_info["push_levels"]), "pull_levels": int(step_info["pull_levels"])},
        "roofline": roofline, "cpu_baseline": cpu, "parity": parity,
    }


def compact(line):
    """A sub-workload's entry in the default run's line."""
    keep = ("metric", "value", "unit", "steps", "warmup", "ms_per_step", "dtype", "parity")
    out = {k: line[k] for k in keep if k in line}
    out["workload"] = line["config"]["workload"]
    r = dict(line["roofline"])
    r.pop("kernel_ms_per_step", None)
    out["roofline"] = r
    cb = line.get("cpu_baseline")
    if cb:
        out["cpu_baseline"] = {k: cb[k] for k in ("value", "unit", "cores", "kind", "sample") if k in cb}
    if "pagerank" in line:
        p = line["pagerank"]
        out["pagerank"] = {"value": p["value"], "unit": p["unit"], "iterations": p["iterations"],
                           "fp64_iterations": p["fp64_iterations"],
                           "roofline": {k: p["roofline"].get(k) for k in ("achieved", "frac", "traffic", "avg_launch_us",
                                                                          "frac_of_gather_floor",
                                                                          "frac_of_gather_floor_lds")}}
        if cb and "pagerank_gbs" in cb:
            out["pagerank"]["cpu_baseline_gbs"] = cb["pagerank_gbs"]
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="all",
                    choices=["all", "rmat20", "er10k", "grid2048", "ba512", "rmat20w", "power", "pivots", "copy"])
    ap.add_argument("--sub-steps", type=int, default=2, help="all: timed steps of each non-headline workload")
    ap.add_argument("--skip", default="", help="all: comma-separated workloads to leave out")
    ap.add_argument("--train-per-scale", type=int, default=40, help="pivots: R-MAT training graphs per scale 14..18")
    ap.add_argument("--epochs", type=int, default=30, help="pivots: training epochs")
    ap.add_argument("--eval-scale", type=int, default=20, help="pivots: R-MAT scale of the judged graph")
    ap.add_argument("--reps", type=int, default=3, help="pivots: samples per pivot budget")
    ap.add_argument("--ks", default=None, help="pivots: comma-separated pivot budgets (default 2048, 4096, ... < n)")
    ap.add_argument("--exact-cache", default=None, help="pivots: .npz of the exact scores (read if present, else written)")
    ap.add_argument("--model-cache", default=None, help="pivots: .npz of the trained network (read if present, else written)")
    ap.add_argument("--exact-only", action="store_true", help="pivots: stop after the exact scores")
    ap.add_argument("--k-max", type=int, default=1 << 30, help="pivots: largest pivot budget")
    ap.add_argument("--samplers", default="uniform,uniform_nonsink,uniform_folded,learned_folded",
                    help="pivots: comma-separated samplers")
    ap.add_argument("--groups", type=int, default=4, help="64-source groups per launch")
    ap.add_argument("--pull-grid", type=int, default=None, help="pull workgroups per group (0 = one wave per chunk)")
    ap.add_argument("--alpha", type=float, default=None, help="push/pull switch factor")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default launch shape")
    ap.add_argument("--order", type=int, default=None, choices=[0, 1, 2],
                    help="source grouping (zg_set_source_order mode)")
    ap.add_argument("--no-graph", action="store_true", help="launch every level kernel from the host")
    ap.add_argument("--no-lane-order", action="store_true",
                    help="run the 64-lane engine on the resource's own ids (no degree-ordered view)")
    ap.add_argument("--pr-cache", type=int, default=None, help="PageRank values staged in LDS (0 = none)")
    ap.add_argument("--pr-grid", type=int, default=None, help="PageRank iteration blocks")
    ap.add_argument("--pr-block", type=int, default=None, choices=[0, 256, 512, 1024],
                    help="PageRank iteration block threads (0 = default)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--wbatch", type=int, default=None, help="rmat20w: sources per batch (multiple of 64)")
    ap.add_argument("--wdelta", type=float, default=None, help="rmat20w: near-frontier bound step (0 = default)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(the path would shard by source: a graph replica per GPU, the sources split, one "
              "all-reduce of the scores)", file=sys.stderr)
        sys.exit(2)

    import torch  # noqa: F401  (initialises the GPU runtime before the library)

    w = args.workload
    if w == "pivots":
        line = bench_pivots(args)
    elif w == "copy":
        line = bench_copy(args)
    elif w == "rmat20w":
        line = bench_weighted(args, args.steps, args.warmup)
    elif w == "power":
        line = bench_power(args, args.steps)
    elif w == "ba512":
        line = bench_batch(args, args.steps, args.warmup)
    elif w in ("rmat20", "er10k", "grid2048"):
        line = bench_graph(args, w, args.steps, args.warmup)
    else:
        # headline workload, then a shorter run of every other one
        line = bench_graph(args, "rmat20", args.steps, args.warmup)
        cleanup()
        skip = set(x for x in args.skip.split(",") if x)
        ss = max(1, args.sub_steps)
        subs = {}
        runs = [("er10k", lambda: bench_graph(args, "er10k", ss, 1)),
                ("grid2048", lambda: bench_graph(args, "grid2048", ss, 1)),
                ("rmat20w", lambda: bench_weighted(args, ss, 1)),
                ("power", lambda: bench_power(args, ss)),
                ("ba512", lambda: bench_batch(args, 1, 0))]
        for name, fn in runs:
            if name in skip:
                continue
            t0 = time.time()
            subs[name] = compact(fn())
            subs[name]["seconds"] = round(time.time() - t0, 1)
            log(f"[bench] {name} done in {subs[name]['seconds']}s")
            cleanup()
        line["workloads"] = subs
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
