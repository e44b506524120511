"""Benchmark of the DTO permutation null on one MI355X.

Headline (the metric's configuration): the genome-scale job -- lists of
22,000 features with a 20,000-feature background (N = 20,000 after the
drop), 600 planted shared features, 10% ties -- run as the reference CLI
runs it: the unpermuted task plus 10,000 permutations, here with the
10,000 explicit permutations already resident in HBM.  One step = one whole
job: one launch of the 10,001 tasks (task 0 is the identity permutation);
value = permutations per second.

Extra lines on the same JSON object: the plain FP64 kernel (every grid
cell exact), on-device Philox shuffles, the 512-job batched persistent
kernel against one-launch-per-job, the serving loop's p50/p99 latency under
Poisson arrivals, the live roofline of the dominant kernel, and the CPU
baseline (the reference's algorithm restated, timed on a bounded sample).

    python bench.py [--gpus 1] [--steps 5] [--warmup 1]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

RECORD_BYTES = 72
HBM_PEAK_GBS = 8000.0
METRIC = "permutations/sec at N=20k (batched vs plain, 1 GPU); p50/p99 DTO job latency"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=5)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--perms", type=int, default=10000)
    p.add_argument("--plain-perms", type=int, default=256)
    p.add_argument("--dense-perms", type=int, default=10000)
    p.add_argument("--dense-plain-perms", type=int, default=64)
    p.add_argument("--batch-jobs", type=int, default=512)
    p.add_argument("--batch-perms", type=int, default=1000)
    p.add_argument("--serving-requests", type=int, default=5000)
    p.add_argument("--serving-rate", type=float, default=500.0, help="offered load, requests/s")
    p.add_argument("--serving-list-pairs", type=int, default=500, help="distinct generated list pairs")
    p.add_argument("--cpu-seconds", type=float, default=15.0, help="target CPU baseline sample length")
    p.add_argument("--cli-runs", type=int, default=10, help="timed runs of the CLI on the README lists")
    p.add_argument("--only", choices=["all", "headline", "calibrate-copy"], default="all")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--traffic-file", default=os.path.join(ROOT, "profiles", "k_optimize_genome_20k_pmc.txt"))
    p.add_argument("--sq-file", default=os.path.join(ROOT, "profiles", "k_optimize_genome_20k_sq_counters.txt"))
    return p.parse_args()


def read_profile_value(path, key):
    """A `key = value` figure from a committed rocprofv3 summary (None if absent)."""
    try:
        with open(path) as f:
            for line in f:
                if line.startswith(key) and "=" in line:
                    return float(line.split("=", 1)[1].split()[0])
    except (OSError, ValueError):
        pass
    return None


def read_traffic(path):
    """HBM bytes per launch from the committed rocprofv3 --pmc summary
    (FETCH_SIZE and WRITE_SIZE of the permutation launch, each scaled by the
    copy-kernel calibration factor recorded in the same file)."""
    try:
        vals = {}
        with open(path) as f:
            for line in f:
                if "=" in line and not line.startswith("#"):
                    k, v = line.split("=", 1)
                    vals[k.strip()] = v.strip()
        return float(vals["traffic_bytes_per_launch"])
    except (OSError, KeyError, ValueError):
        return None


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path (permutations and jobs "
              "would shard across GPUs, with one reduction of the per-job counts).", file=sys.stderr)
        return 2
    import numpy as np
    import torch

    import dual_threshold_optimization_amd as dto
    from dual_threshold_optimization_amd import workloads as W

    dto._lib.require_device()
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)

    if args.only == "calibrate-copy":
        # a plain copy of known size, for scaling the FETCH_SIZE / WRITE_SIZE counters
        a = torch.ones(256 << 20, dtype=torch.float32, device=dev)  # 1 GiB
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": a.numel() * 4}))
        return 0

    # ---- headline job: genome scale, explicit permutations resident in HBM ----
    t = time.perf_counter()
    ids1, r1, ids2, r2, bg = W.genome_20k()
    job = dto.Job.from_lists(ids1, r1, ids2, r2, background=bg, drop_outside_background=True)
    N1, N2, P = job.len1, job.len2, args.perms
    log("job: N=%d len1=%d len2=%d T1=%d T2=%d (%.1fs)" % (job.population_size, N1, N2, job.n_thresholds1,
                                                           job.n_thresholds2, time.perf_counter() - t))
    gen = torch.Generator(device=dev)
    gen.manual_seed(2)
    # row 0: the identity (task 0, the unpermuted run); rows 1..P: the permutations.
    # One launch runs all P + 1 tasks, as run_single_node runs them on one pool.
    idx1 = torch.argsort(torch.rand(P + 1, N1, device=dev, generator=gen), dim=1).to(torch.int32)
    idx2 = torch.argsort(torch.rand(P + 1, N2, device=dev, generator=gen), dim=1).to(torch.int32)
    idx1[0] = torch.arange(N1, dtype=torch.int32, device=dev)
    idx2[0] = torch.arange(N2, dtype=torch.int32, device=dev)
    idx1, idx2 = idx1.contiguous(), idx2.contiguous()
    out_all = torch.empty((P + 1) * RECORD_BYTES, dtype=torch.uint8, device=dev)
    stream = torch.cuda.current_stream()
    sptr = stream.cuda_stream
    torch.cuda.synchronize()

    def launch(mode, n, i1, i2, o):
        job.run_device(n, o.data_ptr(), d_idx1=i1.data_ptr(), d_idx2=i2.data_ptr(), mode=mode, stream=sptr)

    def step(mode="screened", events=None):
        if events is not None:
            events[0].record(stream)
        launch(mode, P + 1, idx1, idx2, out_all)
        if events is not None:
            events[1].record(stream)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        step(events=evs[k])
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    wall = t1 - t0
    kern_ms = float(np.mean([a.elapsed_time(b) for a, b in evs]))
    value = P * args.steps / wall
    allrec = np.frombuffer(out_all.cpu().numpy().tobytes(), dtype=dto._lib.RECORD_DTYPE)
    unperm, rec = allrec[0], allrec[1:]
    emp = float(np.mean(rec["pvalue"] <= unperm["pvalue"]))
    log("screened explicit: %.1f ms/step, kernel %.2f ms, %.0f perms/s, unpermuted p=%r empirical=%r"
        % (wall / args.steps * 1e3, kern_ms, value, float(unperm["pvalue"]), emp))

    bytes_per_launch = (P + 1) * (4 * N1 + 4 * N2 + RECORD_BYTES)
    achieved = bytes_per_launch / (kern_ms * 1e-3) / 1e9
    # the kernel is bound by VALU issue and latency, not by HBM: VALU instructions
    # per launch (committed SQ counter summary of the same launch) over the
    # launch's live duration and the chip's wave64 VALU issue rate
    valu_insts = read_profile_value(args.sq_file, "valu_insts_per_launch")
    valu_rate = 256 * 4 * 2.4e9 / 4  # CUs x SIMDs x clock / 4 cycles per wave64 instruction
    valu_frac = None if valu_insts is None else round(valu_insts / (kern_ms * 1e-3 * valu_rate), 4)
    roofline = {"bound": "hbm", "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": read_traffic(args.traffic_file),
                "valu_issue_frac": valu_frac,
                "valu_issue_basis": "SQ_INSTS_VALU per launch (profiles/k_optimize_genome_20k_sq_counters.txt) / "
                                    "(live kernel time x 256 CUs x 4 SIMDs x 2.4 GHz / 4)",
                "kernel": "k_optimize<512 threads, explicit permutations, H table in LDS, screened, pairs in LDS, "
                          "column constants in registers, bins in LDS>",
                "kernel_ms": round(kern_ms, 4),
                "algorithmic_bytes_per_launch": bytes_per_launch}

    result = {"metric": METRIC, "value": round(value, 2), "unit": "permutations/s", "n_gpus": 1,
              "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(wall / args.steps * 1e3, 3),
              "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
              "data": "synthetic (seeded planted-overlap lists; torch-seeded explicit permutations)",
              "config": {"workload": "genome_20k: lists 22,000 features, background 20,000 (N=20,000 after drop), "
                                     "600 planted, 10%% ties, reference thresholds (T=%dx%d), unpermuted + %d explicit "
                                     "permutations per step" % (job.n_thresholds1, job.n_thresholds2, P),
                         "mode": "screened (FP32 log-space screen + exact FP64 recheck)"},
              "roofline": roofline}
    if args.only == "headline":
        print(json.dumps(result))
        return 0

    # ---- same job: plain FP64 kernel (every cell exact) and Philox shuffles ----
    Pp = min(args.plain_perms, P)
    i1p, i2p = idx1[1:], idx2[1:]
    out = torch.empty(P * RECORD_BYTES, dtype=torch.uint8, device=dev)
    launch("plain", Pp, i1p, i2p, out)
    torch.cuda.synchronize()
    t = time.perf_counter()
    launch("plain", Pp, i1p, i2p, out)
    torch.cuda.synchronize()
    plain_rate = Pp / (time.perf_counter() - t)
    rec_plain = np.frombuffer(out[:Pp * RECORD_BYTES].cpu().numpy().tobytes(), dtype=dto._lib.RECORD_DTYPE)
    same = bool(np.array_equal(rec_plain["pvalue"].view(np.uint64), rec["pvalue"][:Pp].view(np.uint64)) and
                np.array_equal(rec_plain["idx1"], rec["idx1"][:Pp]) and np.array_equal(rec_plain["idx2"], rec["idx2"][:Pp]))
    log("plain FP64: %.0f perms/s (screened identical on %d perms: %s)" % (plain_rate, Pp, same))
    job.run_device(P, out.data_ptr(), seed=9, stream=sptr)
    torch.cuda.synchronize()
    t = time.perf_counter()
    job.run_device(P, out.data_ptr(), seed=9, stream=sptr)
    torch.cuda.synchronize()
    philox_rate = P / (time.perf_counter() - t)
    log("screened Philox: %.0f perms/s" % philox_rate)
    result["genome_20k"] = {"screened_explicit_perms_per_s": round(value, 2),
                            "plain_fp64_perms_per_s": round(plain_rate, 2),
                            "screened_philox_perms_per_s": round(philox_rate, 2),
                            "screened_vs_plain": round(value / plain_rate, 2),
                            "screened_equals_plain_on_%d_perms" % Pp: same}
    del idx1, idx2, i1p, i2p, out_all
    torch.cuda.empty_cache()

    # ---- dense thresholds: 2,000 x 2,000 custom thresholds on the same job (4M cells) ----
    if args.dense_perms > 0:
        t1, t2 = W.dense_2k_thresholds()
        dj = dto.Job.from_lists(ids1, r1, ids2, r2, background=bg, drop_outside_background=True)
        dj.set_thresholds(t1, t2)
        DP = args.dense_perms
        dout = torch.empty(DP * RECORD_BYTES, dtype=torch.uint8, device=dev)
        dj.run_device(DP, dout.data_ptr(), seed=4, stream=sptr)
        torch.cuda.synchronize()
        t = time.perf_counter()
        dj.run_device(DP, dout.data_ptr(), seed=4, stream=sptr)
        torch.cuda.synchronize()
        dense_rate = DP / (time.perf_counter() - t)
        DS = min(args.dense_plain_perms, DP)
        t = time.perf_counter()
        dj.run_device(DS, dout.data_ptr(), seed=4, mode="plain", stream=sptr)
        torch.cuda.synchronize()
        dense_plain_rate = DS / (time.perf_counter() - t)
        # accuracy against the all-FP64 oracle: its records for the unpermuted
        # grid and the explicit permutations of tests/golden/dense_2k_genome_20k.json
        with open(os.path.join(ROOT, "tests", "golden", "dense_2k_genome_20k.json")) as f:
            fx = json.load(f)
        FP = len(fx["permuted"])
        e1, e2 = W.explicit_permutations(FP, dj.len1, dj.len2, 21)
        res, perm = dj.run(FP, idx1=e1, idx2=e2, mode="screened")
        got = [(int(res.unpermuted.rank1), int(res.unpermuted.rank2), float(res.unpermuted.pvalue))]
        got += [(int(r["rank1"]), int(r["rank2"]), float(r["pvalue"])) for r in perm]
        want = [(d["rank1"], d["rank2"], float.fromhex(d["pvalue"])) for d in [fx["unpermuted"]] + fx["permuted"]]
        rel = [abs(g[2] - w[2]) / w[2] if w[2] > 0 else abs(g[2] - w[2]) for g, w in zip(got, want)]
        agree = float(np.mean([g[:2] == w[:2] for g, w in zip(got, want)]))
        log("dense 2k x 2k: screened %.0f perms/s (Philox), plain FP64 %.1f perms/s; vs the FP64 oracle on %d grids: "
            "max rel err of min p %.3g, argmin agreement %.4f" % (dense_rate, dense_plain_rate, len(want),
                                                                   max(rel), agree))
        result["dense_2k_thresholds"] = {
            "workload": "genome_20k job with thresholds 10, 20, ..., 20,000 on both lists (4,000,000 cells), "
                        "%d Philox permutations per launch" % DP,
            "screened_perms_per_s": round(dense_rate, 2), "plain_fp64_perms_per_s": round(dense_plain_rate, 3),
            "screened_vs_plain": round(dense_rate / dense_plain_rate, 2),
            "max_rel_err_min_p_vs_fp64_oracle": float(max(rel)), "argmin_agreement_vs_fp64_oracle": agree,
            "oracle_grids_compared": len(want),
            "oracle": "committed all-FP64 oracle records (tests/golden/make_dense_fixture.py): the unpermuted grid "
                      "and %d explicit permutations" % FP}
        del dout
        dj.close()

    # ---- batched: 512 uneven jobs in one persistent launch vs one launch per job ----
    if args.batch_jobs > 0:
        t = time.perf_counter()
        specs = W.batch_jobs(n_jobs=args.batch_jobs, seed=3)
        jobs = [dto.Job.from_lists(a, b, c, d) for a, b, c, d, _ in specs]
        del specs
        log("batch: %d jobs built (%.1fs)" % (len(jobs), time.perf_counter() - t))
        BP = args.batch_perms
        b = dto.Batch(jobs, [BP] * len(jobs), seed=3)
        waves = b.autotune()
        b.launch()
        torch.cuda.synchronize()
        t = time.perf_counter()
        reps = 2
        for _ in range(reps):
            b.launch()
        torch.cuda.synchronize()
        tb = (time.perf_counter() - t) / reps
        batched_rate = len(jobs) * BP / tb
        # one launch per job (screened), same work
        outs = torch.empty((BP + 1) * RECORD_BYTES, dtype=torch.uint8, device=dev)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for j in jobs:
            j.run_device(BP, outs.data_ptr(), seed=3, stream=sptr)
        torch.cuda.synchronize()
        per_job_rate = len(jobs) * BP / (time.perf_counter() - t)
        # one launch per job, plain FP64 kernel, on a 32-job sample of 100 permutations each
        sample = jobs[:: max(1, len(jobs) // 32)]
        t = time.perf_counter()
        for j in sample:
            j.run_device(100, outs.data_ptr(), seed=3, mode="plain", stream=sptr)
        torch.cuda.synchronize()
        per_job_plain_rate = len(sample) * 100 / (time.perf_counter() - t)
        log("batched: %.0f perms/s (waves/WG %d); per-job screened %.0f; per-job plain FP64 %.0f (sample)"
            % (batched_rate, waves, per_job_rate, per_job_plain_rate))
        result["batched_512"] = {"perms_per_s": round(batched_rate, 2), "launch_ms": round(tb * 1e3, 3),
                                 "autotuned_waves_per_workgroup": waves,
                                 "one_launch_per_job_screened_perms_per_s": round(per_job_rate, 2),
                                 "one_launch_per_job_plain_fp64_perms_per_s": round(per_job_plain_rate, 2),
                                 "batched_vs_plain_fp64": round(batched_rate / per_job_plain_rate, 2),
                                 "jobs": len(jobs), "perms_per_job": BP, "hipgraph": True}
        b.close()
        del jobs

    # ---- serving: Poisson arrivals, continuous batching vs one launch per job ----
    if args.serving_requests > 0:
        result["serving"] = serving(dto, W, np, args)

    # ---- the reference's own published run: README lists, -p 1000, CLI wall time ----
    if args.cli_runs > 0:
        result["cli_readme_p1000"] = cli_wall(args.cli_runs)

    # ---- CPU baseline: the reference algorithm restated, bounded sample ----
    if not args.no_cpu:
        result["cpu_baseline"], result["cpu_optimized"] = cpu_baseline(ids1, r1, ids2, r2, bg, args.cpu_seconds)
    print(json.dumps(result))
    return 0


def _make_jobs(dto, sets, which, threads):
    """Deferred jobs (tables built on the host, uploaded when first needed)
    for requests whose list pairs are sets[which[i]]; built on a thread pool
    (the library call releases the GIL; the id arrays are built once per
    list pair)."""
    from concurrent.futures import ThreadPoolExecutor

    def one(w):
        c1, r1, c2, r2, cb, nb = sets[w]
        return dto.Job.from_c_arrays(c1, r1, c2, r2, cb, nb, drop_outside_background=True, defer_upload=True)
    with ThreadPoolExecutor(threads) as ex:
        return list(ex.map(one, which))


def serving(dto, W, np, args):
    """The serving workload: R requests with Poisson arrivals at the offered
    rate, N log-uniform in [1,000, 30,000], background 80-100% of the union,
    P in {100, 1,000, 10,000} w.p. {.5, .4, .1}, planted 0-5% of N.  The same
    arrivals go (a) to the continuous-batching server and (b) to the plain
    path, one job per launch in arrival order (Job.run: upload, unpermuted
    task, P permutations, results); latency = completion - arrival."""
    R = args.serving_requests
    arr, Ns, Ps, planted, bgf, seeds = W.serving_requests(n_req=R, seed=5)
    arr = arr / args.serving_rate
    threads, _ = host_threads()
    # the first G requests' list pairs are generated; request i uses pair i % G
    # (its own job, request id and Philox permutations)
    G = min(R, args.serving_list_pairs)
    t = time.perf_counter()
    sets = []
    for g in range(G):
        a, b, c, d, bg = W.serving_job_lists(int(Ns[g]), int(planted[g]), float(bgf[g]), int(seeds[g]))
        sets.append((dto._lib.cstr_array(a), np.ascontiguousarray(b, dtype=np.uint32), dto._lib.cstr_array(c),
                     np.ascontiguousarray(d, dtype=np.uint32), dto._lib.cstr_array(bg), len(bg)))
    which = [i % G for i in range(R)]
    Nreq = np.array([Ns[w] for w in which])
    log("serving: %d list pairs generated for %d requests (%.1fs)" % (G, R, time.perf_counter() - t))
    out = {"requests": R, "distinct_list_pairs": G, "offered_rate_per_s": args.serving_rate,
           "total_permutations": int(np.sum(Ps[:R])), "mean_N": round(float(np.mean(Nreq)), 1),
           "workload": "Poisson arrivals; N log-uniform [1k, 30k]; background 80-100%; P in {100, 1k, 10k} "
                       "w.p. {.5, .4, .1}; planted 0-5%"}

    # (a) continuous batching
    t = time.perf_counter()
    jobs = _make_jobs(dto, sets, which, threads)
    log("serving: %d deferred jobs built (%.1fs)" % (R, time.perf_counter() - t))
    sv = dto.Server(hbm_budget_bytes=8 << 30, perms_per_slice=1024, seed=5)
    lat, done, i = [], 0, 0
    t0 = time.perf_counter()
    while done < R:
        now = time.perf_counter() - t0
        while i < R and arr[i] <= now:
            sv.submit(jobs[i], int(Ps[i]))
            jobs[i] = None
            i += 1
        for rid, res, l in sv.poll(timeout_ms=1):
            lat.append(l)
            done += 1
        if i < R and done < R:
            time.sleep(max(0.0, min(0.001, arr[i] - (time.perf_counter() - t0))))
    span = time.perf_counter() - t0
    sv.close()
    lat = np.array(lat)
    out.update({"p50_latency_s": round(float(np.percentile(lat, 50)), 5),
                "p99_latency_s": round(float(np.percentile(lat, 99)), 5),
                "perms_per_s": round(float(np.sum(Ps[:R])) / span, 2), "span_s": round(span, 3)})
    log("serving (continuous batching): p50 %.4fs p99 %.4fs, %.0f perms/s over %.2fs"
        % (out["p50_latency_s"], out["p99_latency_s"], out["perms_per_s"], span))

    # (b) one job per launch, same arrivals, first come first served
    jobs = _make_jobs(dto, sets, which, threads)
    lat = []
    t0 = time.perf_counter()
    for i in range(R):
        now = time.perf_counter() - t0
        if arr[i] > now:
            time.sleep(arr[i] - now)
        jobs[i].run(int(Ps[i]), seed=5 + i)
        lat.append(time.perf_counter() - t0 - arr[i])
        jobs[i].close()
        jobs[i] = None
    span = time.perf_counter() - t0
    lat = np.array(lat)
    out["one_launch_per_job"] = {"p50_latency_s": round(float(np.percentile(lat, 50)), 5),
                                 "p99_latency_s": round(float(np.percentile(lat, 99)), 5),
                                 "perms_per_s": round(float(np.sum(Ps[:R])) / span, 2), "span_s": round(span, 3)}
    log("serving (one job per launch): p50 %.4fs p99 %.4fs, %.0f perms/s over %.2fs"
        % (out["one_launch_per_job"]["p50_latency_s"], out["one_launch_per_job"]["p99_latency_s"],
           out["one_launch_per_job"]["perms_per_s"], span))
    return out


def cli_wall(runs):
    """Wall time of the drop-in CLI (a fresh process per run: HIP start-up,
    reading the files, the unpermuted task and 1,000 Philox permutations on the
    GPU, JSON out) on the reference's README lists, the run the reference
    times with hyperfine (profiling/hyperfine_results.json:4-11: mean 0.0705 s,
    median 0.0685 s, min 0.0671 s at --threads 28 on an unstated CPU)."""
    import subprocess
    import numpy as np
    exe = os.path.join(ROOT, "dual_threshold_optimization_amd", "dto_amd")
    g = os.path.join(ROOT, "tests", "golden")
    cmd = [exe, "-1", os.path.join(g, "ranklist1.csv"), "-2", os.path.join(g, "ranklist2.csv"), "-p", "1000",
           "-t", "1", "--seed", "1"]
    subprocess.run(cmd, check=True, capture_output=True, timeout=120)  # warm-up (page cache, code objects)
    ts = []
    for _ in range(runs):
        t = time.perf_counter()
        subprocess.run(cmd, check=True, capture_output=True, timeout=120)
        ts.append(time.perf_counter() - t)
    ts = np.array(ts)
    out = {"runs": runs, "mean_s": round(float(ts.mean()), 4), "median_s": round(float(np.median(ts)), 4),
           "min_s": round(float(ts.min()), 4),
           "command": "dto_amd -1 tests/golden/ranklist1.csv -2 tests/golden/ranklist2.csv -p 1000 -t 1 --seed 1",
           "reference_hyperfine": {"mean_s": 0.0705, "median_s": 0.0685, "min_s": 0.0671, "threads": 28,
                                   "hardware": "unstated CPU (context, not a target)",
                                   "source": "profiling/hyperfine_results.json:4-11"}}
    log("CLI README run, -p 1000: mean %.4fs median %.4fs (reference hyperfine mean 0.0705s, 28 threads)"
        % (out["mean_s"], out["median_s"]))
    return out


def host_threads():
    """Threads for the CPU baseline: every CPU this process may run on, up to
    the 16 host CPUs a one-GPU box allots per GPU, and what the host is."""
    try:
        allowed = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        allowed = os.cpu_count() or 1
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    threads = max(1, min(16, allowed))
    return threads, {"nproc": os.cpu_count(), "affinity_cpus": allowed, "cpu_model": model,
                     "threads_used": threads,
                     "note": "threads = the CPUs this process may use, capped at the 16 host CPUs allotted per GPU"}


def cpu_baseline(ids1, r1, ids2, r2, bg, target_s):
    """Times oracle/ (the checker's CPU restatement) on the host cores: the
    reference's per-cell hash-set intersection + statrs tail recomputed per
    call, on every stride-th threshold1 row of one permutation per thread, and
    the optimized restatement (histogram overlaps + memoised ln n!)."""
    import numpy as np

    from oracle import oracle as O
    keep = set(bg)
    f1 = [(a, b) for a, b in zip(ids1, r1) if a in keep]
    f2 = [(a, b) for a, b in zip(ids2, r2) if a in keep]
    l1 = O.OracleList([a for a, _ in f1], [b for _, b in f1])
    l2 = O.OracleList([a for a, _ in f2], [b for _, b in f2])
    threads, host = host_threads()
    rng = np.random.default_rng(7)
    a = np.stack([rng.permutation(len(f1)) for _ in range(threads)])
    b = np.stack([rng.permutation(len(f2)) for _ in range(threads)])
    cells_per_perm = len(l1.thresholds) * len(l2.thresholds)
    # calibrate the row stride on a small sample, then run ~target_s
    t = time.perf_counter()
    c = O.reference_sample(l1, l2, 20000, a[:1], b[:1], 64, 1)
    per_cell = (time.perf_counter() - t) / max(c, 1)
    stride = max(1, int(round(cells_per_perm * per_cell / target_s)))
    t = time.perf_counter()
    cells = O.reference_sample(l1, l2, 20000, a, b, stride, threads)
    dt = time.perf_counter() - t
    rate = cells / dt / cells_per_perm
    log("cpu reference-algorithm port: %.4f perms/s on %d threads (stride %d, %.1fs)" % (rate, threads, stride, dt))
    base = {"value": round(rate, 6), "unit": "permutations/s", "cores": threads, "kind": "port", "host": host,
            "sample": "genome_20k, %d permutations (one per thread), every %d-th of %d threshold1 rows x all %d "
                      "threshold2 columns (%d cells, %.1fs); rate = cells/s / cells per permutation"
                      % (threads, stride, len(l1.thresholds), len(l2.thresholds), cells, dt)}
    t = time.perf_counter()
    O.optimize_many(l1, l2, 20000, a, b, fast=True, threads=threads)
    dt = time.perf_counter() - t
    opt = {"value": round(threads / dt, 4), "unit": "permutations/s", "cores": threads, "kind": "port",
           "sample": "genome_20k, %d permutations, histogram overlaps + memoised ln n!, full statrs tails (%.1fs)"
                     % (threads, dt)}
    log("cpu optimized port: %.3f perms/s" % opt["value"])
    return base, opt


if __name__ == "__main__":
    sys.exit(main())
