"""Benchmark: ms per SG-t-SNE gradient iteration on one MI355X.

Workload (default): powerlaw40_50d -- N=1,300,000 points in 50-D (40 Gaussians,
power-law sizes, anisotropic covariance), exact kNN k=90, perplexity 30, d=2,
1000 iterations (exaggeration 12 for the first 250).  P is built on the GPU
(exact kNN, FP64 bisection, P + P^T); `warmup` iterations run untimed, then
`steps` iterations are timed with inputs resident in HBM.

Prints ONE JSON line on stdout; progress goes to stderr.
"""
import argparse
import hashlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "gpu-sg-sne_amd"))
sys.path.insert(0, ROOT)

METRIC = "ms/iter at N=1.3M d=2; served embeds/s + p99 latency vs one-at-a-time"
PUBLISHED_MS_PER_ITER = 225.0  # SG-tSNE-CUDA, N=1.3M, d=2, GTX 1060 (README plot)
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def workload(name, n_override=None):
    from sgtsne_amd import synth

    if name == "powerlaw40_50d":
        n = n_override or 1_300_000
        X, _ = synth.powerlaw_mixture(n=n)
    elif name == "mixture10_50d":
        n = n_override or 10_000
        X, _ = synth.mixture_isotropic(n=n)
    elif name == "sparse784":
        n = n_override or 60_000
        X, _ = synth.sparse_clusters(n=n)
    else:
        raise SystemExit(f"unknown workload {name}")
    return X


def host_cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(P, Y, d, it0, target_s, max_iters):
    """FP64 oracle (OpenMP, all host cores, then one core) on the same P and Y
    snapshot: ms per gradient iteration."""
    from oracle import oracle as O

    O.lib()
    nthreads = O.get_threads()
    So = O.CSR(P.n, P.rowptr.astype(np.int64), P.col, P.val.astype(np.float64))

    def timed(iters_cap, seconds):
        Yd = Y.astype(np.float64).copy()
        uY = np.zeros_like(Yd)
        g = np.ones_like(Yd)
        t0 = time.perf_counter()
        O.run(So, Yd, uY, g, it0, it0 + 1)
        per = time.perf_counter() - t0
        iters = int(max(1, min(iters_cap, seconds / max(per, 1e-6)) - 1))
        t0 = time.perf_counter()
        O.run(So, Yd, uY, g, it0 + 1, it0 + 1 + iters)
        return (time.perf_counter() - t0) / iters * 1e3, iters + 1

    ms, iters = timed(max_iters, target_s)
    O.set_threads(1)
    ms1, iters1 = timed(3, 4.0)
    O.set_threads(nthreads)
    return ms, nthreads, iters, ms1, iters1


def plain_job(sg, X, d, k, u, y0, max_iter=1000):
    """The plain method on one job: exact kNN, perplexity, sgtsneCUDA-style run."""
    I, D = sg.knn_exact(X, k)
    P = sg.perplexityEqualization(I, D, X.shape[0], k, u)
    Y = sg.sgtsne(P, sg.tsneparams(d=d, maxIter=max_iter, layout=sg.LAYOUT_AOS), y_in=y0)
    return P, Y


def job_kl(sg, P, Y, d):
    with sg.Context(sg.tsneparams(d=d, reorder=0)) as ctx:
        ctx.upload_P(P, symmetrize=True)
        ctx.set_Y(Y)
        return ctx.kl(1.0)


def oracle_kl_ratios(P, y0, d, Ys, iters=1000, self_spread=False):
    """Final KL of each embedding in Ys over the FP64 oracle loop's final KL,
    run from the same P (row-stochastic kNN CSR) and start y0.  With
    self_spread, the last ratio is that of the oracle itself started one ulp
    away (every coordinate moved to the next double): how far the descent's
    own sensitivity alone moves the final KL."""
    from oracle import oracle as O

    S = O.symmetrize(O.CSR(P.n, P.rowptr.astype(np.int64), P.col, P.val.astype(np.float64)))
    h = H_DEFAULT[d]

    def kl_of(Y):
        Y = np.ascontiguousarray(Y, dtype=np.float64)
        _, Z, _ = O.repulsive_interp(Y.copy(), h, O.ZETA_DIRECT)
        return O.kl(S, Y, Z)

    Yo = np.ascontiguousarray(y0, dtype=np.float64).copy()
    O.run(S, Yo, np.zeros_like(Yo), np.ones_like(Yo), 0, iters, h=h, zmode=O.ZETA_DIRECT)
    klo = kl_of(Yo)
    out = [kl_of(Y) / klo for Y in Ys]
    if self_spread:
        Yp = np.nextafter(np.ascontiguousarray(y0, dtype=np.float64), np.inf)
        O.run(S, Yp, np.zeros_like(Yp), np.ones_like(Yp), 0, iters, h=h, zmode=O.ZETA_DIRECT)
        out.append(kl_of(Yp) / klo)
    return out


def _ratio_stats(v):
    return {"min": round(float(np.min(v)), 5), "mean": round(float(np.mean(v)), 5), "max": round(float(np.max(v)), 5),
            "within_1pct": int(np.sum(np.abs(np.asarray(v) - 1.0) < 0.01)), "count": len(v)}


def batched_leg(sg, jobs, sample, k=90, u=30.0, graph=True, front_depth=8):
    """256 jobs (N log-uniform in [500, 50,000], 50-D 5..20-cluster mixtures)
    in one batched engine against the plain method one job at a time.  The
    jobs' front ends run with up to `front_depth` in flight, then every job's
    1000 iterations.  (Iterating the admitted jobs while later front ends run
    measured slower: 4.17 against 4.02 s for the 256 jobs, both device-bound.)"""
    from sgtsne_amd import synth

    sizes = synth.batch_sizes(jobs=jobs)
    seeds = [synth.BASE_SEED + 3000 + i for i in range(jobs)]
    total = int(sizes.sum())
    log(f"[bench] batched: {jobs} jobs, {total} points")
    B = sg.Batch(sg.tsneparams(d=2), jobs, total, 2 * k * total)
    B.set_graph(graph)
    data = [(synth.gaussian_job(int(sizes[i]), seed=seeds[i])[0], synth.random_init(int(sizes[i]), 2, seed=seeds[i]))
            for i in range(jobs)]
    # front ends in flight (up to `front_depth` tickets, polled in turn): the
    # device runs one front end's part while the host takes up the next
    ids = [None] * jobs
    pending = {}
    nxt = 0
    t0 = time.perf_counter()
    while nxt < jobs or pending:
        while nxt < jobs and len(pending) < front_depth:
            X, y0 = data[nxt]
            pending[B.add_points_async(X, k=k, u=u, y_in=y0, max_iter=1000)] = nxt
            nxt += 1
        for t in list(pending):
            j = B.poll_points(t)
            if j is not None:
                ids[pending.pop(t)] = j
    t_front = time.perf_counter() - t0
    del data
    t0 = time.perf_counter()
    B.step(1001)
    B.sync()
    t_iter = time.perf_counter() - t0
    done = sum(1 for j in ids if B.job_status(j)[0] == sg.JOB_DONE)
    info, tune = B.info(), B.tuning()
    # plain method on a size-stratified sample; KL of the same jobs both ways
    order = np.argsort(sizes)
    pick = order[np.linspace(0, jobs - 1, sample).round().astype(int)]
    t_plain = 0.0
    ratios, vs_oracle_b, vs_oracle_p, vs_oracle_self = [], [], [], []
    for i in pick:
        X, _ = synth.gaussian_job(int(sizes[i]), seed=seeds[i])
        y0 = synth.random_init(int(sizes[i]), 2, seed=seeds[i])
        t0 = time.perf_counter()
        P, Yp = plain_job(sg, X, 2, k, u, y0)
        t_plain += time.perf_counter() - t0
        Yb = B.get_Y(int(ids[i]))
        ratios.append(job_kl(sg, P, Yb, 2) / job_kl(sg, P, Yp, 2))
        rb, rp, ro = oracle_kl_ratios(P, y0, 2, [Yb, Yp], self_spread=True)
        vs_oracle_b.append(rb)
        vs_oracle_p.append(rp)
        vs_oracle_self.append(ro)
    B.close()
    t_plain_all = t_plain * jobs / sample
    out = {"jobs": jobs, "points": total, "nnz_sym": info["nnz"], "jobs_done": done,
           "front_end_s": round(t_front, 3), "front_ends_in_flight": front_depth, "iterations_s": round(t_iter, 3),
           "ms_per_batch_iteration": round(t_iter * 1e3 / 1001, 4),
           "embeds_per_s": round(jobs / (t_front + t_iter), 3),
           "plain_embeds_per_s": round(jobs / t_plain_all, 3),
           "speedup_vs_plain": round(t_plain_all / (t_front + t_iter), 2),
           "plain_sample": f"{sample} jobs stratified by N, extrapolated to {jobs}",
           "kl_batched_over_plain": {"min": round(float(np.min(ratios)), 5), "mean": round(float(np.mean(ratios)), 5),
                                     "max": round(float(np.max(ratios)), 5)},
           "kl_over_fp64_oracle": {"batched": _ratio_stats(vs_oracle_b), "plain": _ratio_stats(vs_oracle_p),
                                   "oracle_from_start_one_ulp_away": _ratio_stats(vs_oracle_self),
                                   "per_job": [{"n": int(sizes[i]), "batched": round(b, 5), "plain": round(p_, 5),
                                                "oracle_one_ulp": round(o, 5)}
                                               for i, b, p_, o in zip(pick, vs_oracle_b, vs_oracle_p, vs_oracle_self)],
                                   "sample": "the same sampled jobs, FP64 oracle loop from the same P and start"},
           "graph_replays": info["graph_replays"], "attractive_tuning": tune}
    log(f"[bench] batched: {out}")
    return out


def served_leg(sg, requests, rate, tick, k=90, u=30.0, graph=True, oracle_checks=2, fronts=("async",)):
    """Poisson arrivals (N log-uniform in [1,000, 200,000], d in {2, 3}) through
    the continuous-batching server, then one at a time through the plain
    method.  The server runs once per front-end mode in `fronts`; the first
    is the one reported as `served` (the others as `served_<mode>`)."""
    from sgtsne_amd import serving, synth

    arrivals = synth.serving_requests(requests, rate)
    data = {seed: synth.gaussian_job(n, seed=seed)[0] for _, n, _, seed in arrivals}  # the requests' payloads

    def make_X(n, seed):
        return data[seed]

    log(f"[bench] served: {requests} requests at {rate}/s, N total {sum(a[1] for a in arrivals)}")
    runs = {}
    for mode in fronts:
        with serving.Server(tick=tick, k=k, u=u, front=mode) as srv:
            for e in srv.engines.values():
                e.set_graph(graph)
            t0 = time.perf_counter()
            runs[mode] = (srv.run(arrivals, make_X), time.perf_counter() - t0)
        log(f"[bench] served ({mode} front ends): {serving.latency_stats(runs[mode][0])}")
    got, t_srv = runs[fronts[0]]
    ok = sum(1 for r in got if r.Y is not None and np.all(np.isfinite(r.Y)))
    s_stats = serving.latency_stats(got)
    t0 = time.perf_counter()
    base = serving.one_at_a_time(arrivals, make_X, k=k, u=u)
    t_base = time.perf_counter() - t0
    b_stats = serving.latency_stats(base)
    # quality: the smallest requests of each d against the FP64 oracle loop
    # from the same P and start (served and one-at-a-time results alike)
    ordered = sorted(arrivals)
    check = []
    for dd in (2, 3):
        cand = sorted((a[1], idx) for idx, a in enumerate(ordered) if a[2] == dd)[:oracle_checks]
        check += [idx for _, idx in cand]
    vs_s, vs_b = [], []
    for idx in check:
        _, n, dd, seed = ordered[idx]
        I, D = sg.knn_exact(data[seed], k)
        P = sg.perplexityEqualization(I, D, n, k, u)
        rs, rb = oracle_kl_ratios(P, synth.random_init(n, dd, seed=seed), dd, [got[idx].Y, base[idx].Y])
        vs_s.append(rs)
        vs_b.append(rb)
    out = {"requests": requests, "rate_per_s": rate, "tick": tick, "front_ends": fronts[0], "completed_finite": ok,
           "served": s_stats, "one_at_a_time": b_stats,
           "embeds_per_s_ratio": round(s_stats["embeds_per_s"] / b_stats["embeds_per_s"], 2),
           "kl_over_fp64_oracle": ({"served": _ratio_stats(vs_s), "one_at_a_time": _ratio_stats(vs_b),
                                    "sample": f"the {oracle_checks} smallest requests of each d"} if check else None),
           "wall_s": {"served": round(t_srv, 2), "one_at_a_time": round(t_base, 2)}}
    for mode in fronts[1:]:
        other = runs[mode][0]
        out[f"served_{mode}"] = serving.latency_stats(other)
        out[f"served_{mode}"]["same_Y_as_served"] = all(
            (a.Y is None and b.Y is None) or (a.Y is not None and b.Y is not None and np.array_equal(a.Y, b.Y))
            for a, b in zip(got, other))
    log(f"[bench] served: {out}")
    return out


def engine_leg(sg, d, k=90, u=30.0, graph=True, warm=801, timed=190):
    """One batched engine holding every d-dimensional job of the served
    request stream at once: front ends, `warm` batch iterations, then `timed`
    ones; ms per batch iteration and the jobs' grid sides at the end."""
    from sgtsne_amd import synth

    arr = [a for a in synth.serving_requests(32, 40.0) if a[2] == d]
    with sg.Batch(sg.tsneparams(d=d), 32, 2_000_000, 300_000_000, 24 << 30) as B:
        B.set_graph(graph)
        ids = []
        for _, n, _, seed in arr:
            X, _ = synth.gaussian_job(n, seed=seed)
            ids.append(B.add_points(X, k=k, u=u, y_in=synth.random_init(n, d, seed=seed), max_iter=1000))
        B.step(warm)
        B.sync()
        t0 = time.perf_counter()
        B.step(timed)
        B.sync()
        ms = (time.perf_counter() - t0) * 1e3 / timed
        ngs = sorted(B.terms(j)[3] for j in ids)
        # fingerprint of every job's Y: equal across builds that compute the same bits
        h = hashlib.sha1()
        for j in ids:
            h.update(B.get_Y(j).tobytes())
        tuning = B.tuning()
    res = {"d": d, "jobs": len(ids), "points": sum(a[1] for a in arr), "ms_per_batch_iteration": round(ms, 4),
           "iterations": [warm, timed], "ng": ngs, "graph": graph, "y_sha1": h.hexdigest()[:16],
           "attractive_tuning": tuning}
    log(f"[bench] engine: {res}")
    return res


def knn_leg(sg, n=200_000, k=90):
    """Front-end kNN of one 50-D Gaussian-mixture job: the filtered method
    against the exact FP64 one (rows must be identical).  Times include the
    host-side conversion and copies of the C-ABI call."""
    from sgtsne_amd import synth

    X, _ = synth.gaussian_job(n, seed=5)
    sg.knn(X[:2000], k, sg.KNN_FILTERED)  # module load, pools
    res = {}
    rows = {}
    for name in ("FILTERED", "EXACT_FP64"):
        t0 = time.perf_counter()
        I, D, fb = sg.knn(X, k, getattr(sg, "KNN_" + name))
        res[name.lower() + "_s"] = round(time.perf_counter() - t0, 4)
        rows[name] = (I, D)
        if name == "FILTERED":
            res["fallback_rows"] = fb
    res["identical"] = bool(np.array_equal(rows["FILTERED"][0], rows["EXACT_FP64"][0])
                            and np.array_equal(rows["FILTERED"][1], rows["EXACT_FP64"][1]))
    res.update({"n": n, "dim": X.shape[1], "k": k})
    log(f"[bench] knn: {res}")
    return res


H_DEFAULT = {1: 0.5, 2: 0.7, 3: 1.2}


def front_end(sg, X, k=90, u=30.0):
    """The P front end on the GPU: exact kNN (certified filter where it
    applies), FP64 perplexity bisection.  Returns P and the timings."""
    n = X.shape[0]
    t0 = time.perf_counter()
    I, D, fb = sg.knn(X, k, sg.KNN_AUTO)
    t_knn = time.perf_counter() - t0
    t0 = time.perf_counter()
    P = sg.perplexityEqualization(I, D, n, k, u)
    t_perp = time.perf_counter() - t0
    return P, {"knn": round(t_knn, 2), "knn_rows_recomputed_exactly": fb, "perplexity": round(t_perp, 2)}


def kernel_name(d, tune, fmt="CSR_SEG"):
    if fmt == "COO":
        return f"attr_coo_kernel<{d}>"
    if fmt == "HYBRID":
        return f"attr_ell_kernel<{d}> + attr_coo_kernel<{d}>"
    if fmt == "CSR":
        return f"attr_csr_row_kernel<{d}>"
    return (f"attr_csr_cc_kernel<{d},{tune['group']},{tune['unroll']}>" if tune["window"] <= -2
            else f"attr_csr_loc_kernel<{d},{tune['group']}>" if tune["window"] < 0
            else f"attr_csr_win_kernel<{d},{tune['group']},{tune['unroll']}>" if tune["window"]
            else f"attr_csr_blk_kernel<{d},{tune['group']},{tune['unroll']}>" if tune["rows_per_block"]
            else f"attr_csr_seg2_kernel<{d},{tune['group']},{tune['unroll']}>" if tune["unroll"]
            else f"attr_csr_seg_kernel<{d},{tune['group']}>")


def profile_traffic(profile, kernel):
    """HBM bytes per launch of `kernel` from a committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary (scaled by a plain copy), or None."""
    path = os.path.join(ROOT, "profiles", profile)
    if not os.path.exists(path):
        return None
    pj = json.load(open(path))
    kern = {k.replace(" ", ""): v for k, v in pj["kernels"].items()}
    hit = kern.get(kernel.replace(" ", ""))
    return None if hit is None else hit["traffic_bytes_per_launch"]


def run_config(sg, P, d, steps, warmup, fmt="CSR_SEG", reorder=2, cpu_seconds=15.0, cpu_max_iters=30,
               want_cpu=True, want_terms=True, pin=None, kl_check=False, profile=None):
    """One embedding job from P: `warmup` untimed iterations, then `steps`
    timed ones (inputs resident); per-phase device times, the rooflines of
    the attractive term (b) and of the repulsion (c), the FP32-vs-FP64 term
    errors at the final Y, the CPU baseline on the same P_sym and Y."""
    n = P.n
    prm = sg.tsneparams(d=d, format=getattr(sg, "FORMAT_" + fmt), reorder=reorder)
    ctx = sg.Context(prm)
    t0 = time.perf_counter()
    ctx.upload_P(P, symmetrize=True)
    t_sym = time.perf_counter() - t0
    if pin:
        ctx.set_attractive(*pin)
    nnz = ctx.info()["nnz"]
    y0 = sg.synth.random_init(n, d, seed=1)
    ctx.set_Y(y0)
    y0c = ctx.get_Y() if kl_check else None  # the centred start both runs use
    ctx.set_timing(True)
    t0 = time.perf_counter()
    if warmup > 0:
        ctx.step(warmup)
    t_warm = time.perf_counter() - t0
    tune = ctx.tuning()
    Y_snap = ctx.get_Y()
    it0 = ctx.info()["iter"]
    t0 = time.perf_counter()
    ctx.step(steps)  # synchronises the stream at the end
    t_run = time.perf_counter() - t0
    ms_per_step = t_run * 1e3 / steps
    ph = ctx.timing()  # device time per phase summed over the timed steps
    iters = max(ph[6], 1)
    phases = {"attractive": ph[0] / iters, "s2g": ph[1] / iters, "conv": ph[2] / iters,
              "g2s_forces": ph[3] / iters, "update_center": ph[4] / iters, "device_total": ph[5] / iters}
    log(f"[bench] d={d} {fmt} n={n}: timed {steps} iterations: {ms_per_step:.3f} ms/iter; phases {phases}")
    Y_end = ctx.get_Y()
    reorders = ctx.reorders()
    kl = ctx.kl(1.0)
    Fa_g, Fr_g, Z_g, grid = ctx.terms()

    # (b) attractive SpMV: 8 nnz + 4 (N + 1) + 8 d N bytes per launch over its
    # average event-timed duration
    kern = kernel_name(d, tune, fmt)
    bytes_b = 8.0 * nnz + 4.0 * (n + 1) + 8.0 * d * n
    t_b = phases["attractive"] * 1e-3
    ach_b = bytes_b / t_b / 1e9 if t_b > 0 else 0.0
    roof_b = {"bound": "hbm", "achieved": round(ach_b, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
              "frac": round(ach_b / HBM_PEAK_GBS, 4), "traffic": None, "kernel": kern,
              "bytes_per_launch": bytes_b, "avg_launch_ms": round(phases["attractive"], 5)}
    if profile:
        tr = profile_traffic(profile, kern)
        if tr is not None:
            roof_b["traffic"] = tr
            roof_b["traffic_source"] = f"profiles/{profile} (rocprofv3 --pmc)"
    # (c) repulsion: 12 d N + 16 (d + 1) ng^d bytes per iteration (Y read by
    # the spread and the gather, Frep written; the grid written by the spread,
    # read and written by the convolution, read by the gather) over the
    # summed device time of spread, convolution and gather + forces; ng is the
    # grid side at the end of the timed window
    bytes_c = 12.0 * d * n + 16.0 * (d + 1) * grid ** d
    t_c = (phases["s2g"] + phases["conv"] + phases["g2s_forces"]) * 1e-3
    ach_c = bytes_c / t_c / 1e9 if t_c > 0 else 0.0
    roof_c = {"bound": "hbm", "achieved": round(ach_c, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
              "frac": round(ach_c / HBM_PEAK_GBS, 5), "bytes_per_iteration": bytes_c, "grid_side": grid,
              "ms_per_iteration": round(t_c * 1e3, 5),
              "kernels": "spread + grid reduce, FFT convolution passes, gather + forces + Z partials"}

    out = {"n": n, "d": d, "format": fmt.lower(), "nnz_sym": nnz, "ms_per_iter": round(ms_per_step, 4),
           "timed_iterations": [it0, it0 + steps - 1], "phases_ms": {k2: round(v, 5) for k2, v in phases.items()},
           "roofline": roof_b, "roofline_c": roof_c, "final_grid_side": grid, "final_kl": round(kl, 5),
           "attractive_tuning": tune, "morton_reorders": reorders, "symmetrize_upload_s": round(t_sym, 2),
           "full_run_ms": round((t_warm + t_run) * 1e3, 1)}
    if want_terms or want_cpu or kl_check:
        from oracle import oracle as O

        P_sym = ctx.get_P()
        So = O.CSR(P_sym.n, P_sym.rowptr.astype(np.int64), P_sym.col, P_sym.val.astype(np.float64))
        if want_terms:
            Yd = Y_end.astype(np.float64)
            Fa_o = O.attractive(So, Yd)
            Fr_o, Z_o, ng_o = O.repulsive_interp(Yd.copy(), H_DEFAULT[d], O.ZETA_DIRECT)
            rel = lambda a, b: float(np.linalg.norm(a - b) / np.linalg.norm(b))
            out["term_error_vs_fp64"] = {"attractive_rel_l2": rel(Fa_g, Fa_o), "repulsive_rel_l2": rel(Fr_g, Fr_o),
                                         "Z_rel": abs(Z_g - Z_o) / Z_o, "grid_side_gpu": grid,
                                         "grid_side_oracle": ng_o}
            log(f"[bench] FP32 vs FP64 oracle terms: {out['term_error_vs_fp64']}")
        if want_cpu:
            log(f"[bench] CPU baseline: FP64 oracle from the iteration-{it0} Y ...")
            ms_cpu, cores, cpu_iters, ms1, it1 = cpu_baseline(P_sym, Y_snap, d, it0, cpu_seconds, cpu_max_iters)
            out["cpu_baseline"] = {
                "value": round(ms_cpu, 2), "unit": "ms/iter", "cores": cores, "kind": "port",
                "sample": f"{cpu_iters} iterations (from iteration {it0}) of the FP64 C/OpenMP oracle on the "
                          f"same P_sym (nnz={nnz}) and the GPU's iteration-{it0} Y, own FFT (no FFTW on host)",
                "host_cpu": host_cpu_model(), "one_thread_ms": round(ms1, 1), "one_thread_iters": it1}
            out["gpu_speedup_vs_cpu"] = round(ms_cpu / ms_per_step, 1)
            log(f"[bench] CPU {ms_cpu:.1f} ms/iter on {cores} threads -> GPU speed-up {ms_cpu / ms_per_step:.1f}x")
        if kl_check:
            iters_all = it0 + steps
            log(f"[bench] KL check: FP64 oracle, {iters_all} iterations on the host ...")
            Yo = y0c.astype(np.float64)
            t0 = time.perf_counter()
            O.run(So, Yo, np.zeros_like(Yo), np.ones_like(Yo), 0, iters_all, h=H_DEFAULT[d], zmode=O.ZETA_DIRECT)
            t_o = time.perf_counter() - t0
            _, Zo, _ = O.repulsive_interp(Yo.copy(), H_DEFAULT[d], O.ZETA_DIRECT)
            _, Zg, _ = O.repulsive_interp(Y_end.astype(np.float64).copy(), H_DEFAULT[d], O.ZETA_DIRECT)
            klo = O.kl(So, Yo, Zo)
            klg = O.kl(So, Y_end.astype(np.float64), Zg)
            out["kl_check"] = {"iterations": iters_all, "kl_gpu_fp32": round(klg, 6), "kl_oracle_fp64": round(klo, 6),
                               "rel_diff": round(abs(klg - klo) / klo, 6), "oracle_s": round(t_o, 1),
                               "oracle_threads": O.get_threads()}
            log(f"[bench] KL check: {out['kl_check']}")
        del So, P_sym
    ctx.close()
    return out


def main_leg(sg, args):
    """C3: N=1.3M power-law mixture, d=2 (the headline) and, from the same P,
    d=3 with the 3-D grid FFT."""
    k, u = 90, 30.0
    t_setup = time.perf_counter()
    log(f"[bench] generating {args.workload} ...")
    X = workload(args.workload, args.n)
    n = X.shape[0]
    log(f"[bench] n={n} dim={X.shape[1]}: kNN k={k} on the GPU ...")
    P, setup = front_end(sg, X, k, u)
    del X
    log(f"[bench] front end {setup}; setup {time.perf_counter() - t_setup:.1f} s")
    pin = [int(x) for x in args.pin_attractive.split(",")] if args.pin_attractive else None
    full = args.workload == "powerlaw40_50d" and n == 1_300_000
    r = run_config(sg, P, args.d, args.steps, args.warmup, fmt=args.format, reorder=args.reorder,
                   cpu_seconds=args.cpu_seconds, cpu_max_iters=args.cpu_max_iters, want_cpu=not args.no_cpu,
                   pin=pin, kl_check=args.kl_check,
                   profile="attractive_spmv_powerlaw1.3M_d2.json" if full and args.d == 2 else None)
    setup["symmetrize_upload"] = r.pop("symmetrize_upload_s")
    d3 = None
    if "d3" in args.configs and args.d == 2:
        log("[bench] C3 at d=3 from the same P ...")
        d3 = run_config(sg, P, 3, args.steps, args.warmup, fmt=args.format, reorder=args.reorder,
                        cpu_seconds=min(args.cpu_seconds, 8.0), cpu_max_iters=10, want_cpu=not args.no_cpu,
                        profile="attractive_spmv_powerlaw1.3M_d2.json" if full else None)
    return {
        "value": r["ms_per_iter"],
        "ms_per_step": r["ms_per_iter"],
        "config": {"workload": f"{args.workload} N={n} d={args.d} k={k} u={u}", "n": n, "d": args.d,
                   "nnz_sym": r["nnz_sym"], "format": r["format"], "reorder": args.reorder,
                   "morton_reorders": r["morton_reorders"], "timed_iterations": r["timed_iterations"],
                   "final_grid_side": r["final_grid_side"], "final_kl": r["final_kl"],
                   "attractive_tuning": r["attractive_tuning"]},
        "phases_ms": r["phases_ms"],
        "full_run_ms": r["full_run_ms"],
        "setup_s": setup,
        "roofline": r["roofline"],
        "roofline_c": r["roofline_c"],
        "cpu_baseline": r.get("cpu_baseline"),
        "term_error_vs_fp64": r.get("term_error_vs_fp64"),
        "kl_check": r.get("kl_check"),
        "c3_d3": d3,
    }


def c2_leg(sg, args):
    """C2: N=60,000 sparse 784-D clusters: the hybrid ELL + COO format against
    COO (and this engine's CSR kernel), GPU against the CPU oracle."""
    X = workload("sparse784")
    P, setup = front_end(sg, X)
    del X
    out = {"workload": f"sparse784 N={P.n} dim=784 d=2 k=90 u=30", "setup_s": setup, "formats": {}}
    for fmt in ("HYBRID", "COO", "CSR_SEG"):
        r = run_config(sg, P, 2, args.steps, args.warmup, fmt=fmt, cpu_seconds=5.0, cpu_max_iters=20,
                       want_cpu=(fmt == "CSR_SEG" and not args.no_cpu), want_terms=True)
        out["formats"][fmt.lower()] = {k2: r.get(k2) for k2 in ("ms_per_iter", "phases_ms", "roofline",
                                                                "term_error_vs_fp64", "final_grid_side")}
        if "cpu_baseline" in r:
            out["cpu_baseline"] = r["cpu_baseline"]
    f = out["formats"]
    out["hybrid_over_coo_speedup"] = round(f["coo"]["ms_per_iter"] / f["hybrid"]["ms_per_iter"], 3)
    if out.get("cpu_baseline"):
        out["gpu_speedup_vs_cpu"] = {k2: round(out["cpu_baseline"]["value"] / v["ms_per_iter"], 1)
                                     for k2, v in f.items()}
    log(f"[bench] C2: {out}")
    return out


def c1_leg(sg, args):
    """C1: N=10,000 isotropic 10-cluster mixture in 50-D, GPU against CPU."""
    X = workload("mixture10_50d")
    P, setup = front_end(sg, X)
    r = run_config(sg, P, 2, args.steps, args.warmup, cpu_seconds=5.0, cpu_max_iters=50,
                   want_cpu=not args.no_cpu)
    r["workload"] = f"mixture10_50d N={P.n} dim=50 d=2 k=90 u=30"
    r["setup_s"] = setup
    log(f"[bench] C1: {r}")
    return r


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=500)
    ap.add_argument("--workload", default="powerlaw40_50d")
    ap.add_argument("--n", type=int, default=None)
    ap.add_argument("--d", type=int, default=2)
    ap.add_argument("--format", default="CSR_SEG")
    ap.add_argument("--reorder", type=int, default=2, help="device point order: 0 given, 1 reverse Cuthill-McKee of P, 2 that plus Morton re-orders of Y kept when faster")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--cpu-max-iters", type=int, default=30)
    ap.add_argument("--calibrate", action="store_true", help="also time a plain 256 MiB device copy")
    ap.add_argument("--pin-attractive", default=None,
                    help="G,U,RPB,WIN: fix the attractive kernel shape instead of autotuning (profiling)")
    ap.add_argument("--kl-check", action="store_true",
                    help="run the FP64 oracle for all iterations from the same P_sym and start; compare final KL")
    ap.add_argument("--legs", default="main,batched,served",
                    help="comma list of main (N=1.3M ms/iter), batched (256 jobs), served (request stream), "
                         "knn (front-end kNN methods at N=200k), engine (one batched engine holding the "
                         "served stream's jobs of --engine-d); knn and engine are not in the default list")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--batch-jobs", type=int, default=256)
    ap.add_argument("--batch-sample", type=int, default=16)
    ap.add_argument("--no-served", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=32)
    ap.add_argument("--serve-rate", type=float, default=40.0)
    ap.add_argument("--serve-tick", type=int, default=10)
    ap.add_argument("--batch-front-depth", type=int, default=8,
                    help="front ends in flight while the batched leg admits its jobs (1..8)")
    ap.add_argument("--serve-fronts", default="async",
                    help="comma-separated front-end modes of the server (async, sync); the first is reported")
    ap.add_argument("--configs", default="d3,c2,c1",
                    help="extra configurations besides the headline: d3 (C3 at d=3 from the same P), "
                         "c2 (sparse784 formats), c1 (mixture10_50d)")
    ap.add_argument("--engine-d", type=int, default=3,
                    help="engine leg: the served stream's jobs of this d in one batch")
    ap.add_argument("--no-graph", action="store_true",
                    help="batched and served engines launch each kernel instead of replaying a graph")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the multi-GPU "
              "sharding of this path is a design note, not a measured configuration.", file=sys.stderr)
        sys.exit(2)

    import sgtsne_amd as sg

    legs = set(args.legs.split(","))
    args.configs = set(c for c in args.configs.split(",") if c)
    if args.no_batched:
        legs.discard("batched")
    if args.no_served:
        legs.discard("served")
    m = main_leg(sg, args) if "main" in legs else {}

    copy = None
    if args.calibrate:
        nb = 256 << 20
        ms = sg.copy_probe(nb, 20)
        copy = {"bytes": nb, "ms": round(ms, 5), "GB_per_s": round(2 * nb / ms / 1e6, 1)}
        log(f"[bench] plain copy: {copy}")
    g = not args.no_graph
    batched = (batched_leg(sg, args.batch_jobs, args.batch_sample, graph=g, front_depth=args.batch_front_depth)
               if "batched" in legs else None)
    served = (served_leg(sg, args.serve_requests, args.serve_rate, args.serve_tick, graph=g,
                         fronts=tuple(args.serve_fronts.split(","))) if "served" in legs else None)
    knn = knn_leg(sg) if "knn" in legs else None
    c2 = c2_leg(sg, args) if "c2" in args.configs and "main" in legs else None
    c1 = c1_leg(sg, args) if "c1" in args.configs and "main" in legs else None
    engine = engine_leg(sg, args.engine_d, graph=g) if "engine" in legs else None

    value = m.get("value")
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "ms/iter",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": m.get("ms_per_step"),
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": round(value / PUBLISHED_MS_PER_ITER, 6) if value else None,
        "dtype": "f32",
        "data": "synthetic",
        "config": m.get("config", {"workload": "none (main leg skipped)"}),
        "phases_ms": m.get("phases_ms"),
        "full_run_ms": m.get("full_run_ms"),
        "setup_s": m.get("setup_s"),
        "roofline": m.get("roofline"),
        "roofline_c": m.get("roofline_c"),
        "cpu_baseline": m.get("cpu_baseline"),
        "term_error_vs_fp64": m.get("term_error_vs_fp64"),
        "copy_probe": copy,
        "kl_check": m.get("kl_check"),
        "c3_d3": m.get("c3_d3"),
        "c2_sparse784": c2,
        "c1_mixture10": c1,
        "batched": batched,
        "served": served,
        "knn": knn,
        "engine": engine,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
