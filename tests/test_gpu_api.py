"""The reference's callers on the GPU path: the CLI (native and Python), the
crate-shaped Python API (optimize with and without debug, run_single_node,
empirical_pvalue), job inspection, and the per-permutation records of
dto_run against dto_run_device."""
import ctypes
import json
import os
import subprocess
import sys

import numpy as np
import pytest

from conftest import ROOT
from helpers import GOLDEN, readme_lists, rec_tuple

pytestmark = pytest.mark.gpu

README = {"population_size": 30, "rank1": 24, "rank2": 13, "set1_len": 24, "set2_len": 13,
          "unpermuted_intersection_size": 12, "unpermuted_pvalue": 0.15632183908046102, "fdr": 0.0}


def _cli(args, python=False):
    if python:
        cmd = [sys.executable, "-m", "dual_threshold_optimization_amd"] + args
    else:
        cmd = [os.path.join(ROOT, "dual_threshold_optimization_amd", "dto_amd")] + args
    return subprocess.run(cmd, capture_output=True, text=True, timeout=120, cwd=ROOT)


@pytest.mark.parametrize("python", [False, True])
def test_cli_readme_run(dto, gpu, python):
    """README.md:165-186: the README run, its stderr lines and its JSON."""
    r1, r2 = os.path.join(GOLDEN, "ranklist1.csv"), os.path.join(GOLDEN, "ranklist2.csv")
    r = _cli(["-1", r1, "-2", r2, "-p", "5", "-t", "1", "--seed", "11"], python)
    assert r.returncode == 0, r.stderr
    assert r.stderr.splitlines()[:6] == [
        "Ranked list 1: %s" % r1, "Ranked list 2: %s" % r2, "Permutations: 5", "Threads: 1",
        "Multi-node mode: disabled",
        "The product of the lengths of the threshold lists (this describes the asymptotic runtime of a single "
        "job): 900"]
    out = json.loads(r.stdout)
    assert list(out) == sorted(out)  # serde_json's sorted map
    for k, v in README.items():
        assert out[k] == v, k
    assert out["empirical_pvalue"] * 5 == round(out["empirical_pvalue"] * 5)
    # the same seed gives the same permutations and result on every call, in either CLI
    again = _cli(["-1", r1, "-2", r2, "-p", "5", "--seed", "11"], not python)
    assert json.loads(again.stdout) == out
    # background: the background file's 30 ids make N = 30
    b = _cli(["-1", r1, "-2", r2, "-b", os.path.join(GOLDEN, "background.txt"), "-p", "3", "--seed", "2"], python)
    assert b.returncode == 0 and json.loads(b.stdout)["population_size"] == 30
    # zero permutations: empirical p-value 1.0 (empirical_pvalue.rs:146-158)
    z = _cli(["-1", r1, "-2", r2, "-p", "0"], python)
    assert z.returncode == 0 and json.loads(z.stdout)["empirical_pvalue"] == 1.0


def test_cli_modes_agree(dto, gpu):
    r1, r2 = os.path.join(GOLDEN, "ranklist1.csv"), os.path.join(GOLDEN, "ranklist2.csv")
    a = _cli(["-1", r1, "-2", r2, "-p", "64", "--seed", "5", "--mode", "screened"])
    b = _cli(["-1", r1, "-2", r2, "-p", "64", "--seed", "5", "--mode", "plain"])
    assert a.returncode == 0 and b.returncode == 0
    assert a.stdout == b.stdout


def test_python_api_mirror(dto, O, gpu):
    """optimize / process_threshold_pairs / run_single_node / empirical_pvalue
    with the crate's argument order, against the oracle."""
    i1, r1, i2, r2 = readme_lists()
    l1 = dto.RankedFeatureList.from_(dto.FeatureList(i1), r1)
    l2 = dto.RankedFeatureList.from_(dto.FeatureList(i2), r2)
    N = dto.compute_population_size(l1, l2)
    best = dto.optimize(l1, l2, False, N).best
    assert (best.rank1, best.rank2, best.set1_len, best.set2_len, best.intersection_size, best.pvalue) == \
        (24, 13, 24, 13, 12, 0.15632183908046102)
    assert best.permuted is False
    # debug: every record, with the feature sets of both thresholds
    dbg = dto.optimize(l1, l2, False, N, debug=True).debug
    ol1, ol2 = O.OracleList(i1, r1), O.OracleList(i2, r2)
    want = O.process_threshold_pairs(ol1, ol2, N)
    assert len(dbg) == len(want) == 900
    for d, w in zip(dbg, want):
        assert (d.rank1, d.rank2, d.set1_len, d.set2_len, d.population_size, d.intersection_size, d.pvalue) == \
            w.as_tuple()
        s1, s2 = d.feature_sets
        assert len(s1) == d.set1_len and len(s2) == d.set2_len
        assert len({g.id() for g in s1} & {g.id() for g in s2}) == d.intersection_size
    # permuted lists with explicit indices
    rng = np.random.default_rng(4)
    p1, p2 = rng.permutation(30), rng.permutation(30)
    pl1 = dto.PermutedRankedFeatureList(l1, indices=p1)
    pl2 = dto.PermutedRankedFeatureList(l2, indices=p2)
    got = dto.optimize(pl1, pl2, True, N).best
    assert rec_tuple_obj(got) == O.optimize(ol1, ol2, N, p1, p2).as_tuple()
    assert got.permuted is True
    # run_single_node + empirical_pvalue (main.rs:122-165)
    tasks = [dto.Task(0, False)] + [dto.Task(t, True) for t in range(1, 33)]
    results = dto.run_single_node(tasks, l1, l2, N, 4, seed=21)
    assert len(results) == 33 and sum(1 for r in results if not r.best.permuted) == 1
    out = dto.empirical_pvalue(results)
    perm_p = [r.best.pvalue for r in results if r.best.permuted]
    assert out["empirical_pvalue"] == sum(p <= out["unpermuted_pvalue"] for p in perm_p) / 32
    i_1 = np.stack([dto.philox_permutation(21, 0, p, 1, 30) for p in range(32)])
    i_2 = np.stack([dto.philox_permutation(21, 0, p, 2, 30) for p in range(32)])
    ob = O.optimize_many(ol1, ol2, N, i_1, i_2, fast=False)
    assert [r.best.pvalue for r in results if r.best.permuted] == [b.pvalue for b in ob]


def rec_tuple_obj(r):
    return (r.rank1, r.rank2, r.set1_len, r.set2_len, r.population_size, r.intersection_size, r.pvalue)


def test_job_info_thresholds_and_drop(dto, gpu):
    from dual_threshold_optimization_amd import workloads as W
    ids1, r1, ids2, r2, bg = W.genome_20k()
    job = dto.Job.from_lists(ids1, r1, ids2, r2, background=bg, drop_outside_background=True)
    assert (job.population_size, job.len1, job.len2, job.dropped1, job.dropped2) == (20000, 20000, 20000, 2000, 2000)
    L = dto._lib.lib()
    info = dto._lib.JobInfo()
    dto._lib.check(L.dto_job_info_get(job.handle, ctypes.byref(info)))
    assert info.device_bytes > 0 and info.n_thresholds1 == job.n_thresholds1
    t = np.zeros(job.n_thresholds1, dtype=np.uint32)
    assert L.dto_job_thresholds(job.handle, 1, dto._lib.u32p(t), len(t)) == job.n_thresholds1
    assert L.dto_job_thresholds(job.handle, 3, None, 0) == -dto._lib.DTO_ERR_INVALID_ARGUMENT
    max_rank = max(int(r) for a, r in zip(ids1, r1) if a in set(bg))
    assert list(t) == list(dto.generate_thresholds(max_rank))
    job.close()


def test_run_records_equal_device_records(dto, gpu):
    """dto_run's per-permutation records are the device launch's records."""
    import torch
    from dual_threshold_optimization_amd import workloads as W
    ids1, r1, ids2, r2, _ = W.plain_6k()
    job = dto.Job.from_lists(ids1, r1, ids2, r2)
    P = 40
    res, perm = job.run(P, seed=77)
    out = torch.empty(P * 72, dtype=torch.uint8, device="cuda:0")
    job.run_device(P, out.data_ptr(), seed=77)
    torch.cuda.synchronize()
    dev = np.frombuffer(out.cpu().numpy().tobytes(), dtype=dto._lib.RECORD_DTYPE)
    assert [rec_tuple(x) for x in perm] == [rec_tuple(x) for x in dev]
    assert res.n_permutations == P
    assert res.n_le == sum(1 for x in perm if x["pvalue"] <= res.unpermuted.pvalue)
    # perm_base offsets the permutation ids: the second half alone equals the tail
    out2 = torch.empty((P // 2) * 72, dtype=torch.uint8, device="cuda:0")
    job.run_device(P // 2, out2.data_ptr(), perm_base=P // 2, seed=77)
    torch.cuda.synchronize()
    tail = np.frombuffer(out2.cpu().numpy().tobytes(), dtype=dto._lib.RECORD_DTYPE)
    assert [rec_tuple(x) for x in tail] == [rec_tuple(x) for x in dev[P // 2:]]
    job.close()


def test_batch_autotune_keeps_results(dto, gpu):
    from dual_threshold_optimization_amd import workloads as W
    specs = W.batch_jobs(n_jobs=24, seed=8, n_min=300, n_max=3000)
    jobs = [dto.Job.from_lists(a, b, c, d) for a, b, c, d, _ in specs]
    b = dto.Batch(jobs, [50] * len(jobs), seed=8)
    b.launch()
    before = b.results()
    waves = b.autotune()
    assert waves in (4, 8)
    b.launch()
    after = b.results()
    for x, y in zip(before, after):
        assert rec_tuple_obj(x.unpermuted) == rec_tuple_obj(y.unpermuted)
        assert (x.empirical_pvalue, x.fdr, x.n_le) == (y.empirical_pvalue, y.fdr, y.n_le)
    b.close()


@pytest.mark.parametrize("python", [False, True])
def test_cli_tie_warnings(dto, O, gpu, tmp_path, python):
    """optimize_main.rs:86-108: a task whose minimum p is shared prints the
    first warning (with {:.15} of the p-value), and the second when the
    largest intersection is shared too; one pair of lines per such task, in
    task order here.  The counts come from the kernel and equal the oracle's."""
    # list-2 duplicates of one list-1 id give several p == 0 cells (k > K)
    ids1 = ["a", "b", "c", "d", "e", "f"]
    ids2 = ["a", "a", "a", "b", "c", "x"]
    bg = ids1 + ["x", "y"]
    p1 = tmp_path / "l1.csv"
    p2 = tmp_path / "l2.csv"
    pb = tmp_path / "bg.txt"
    p1.write_text("".join("%s,%d\n" % (g, r + 1) for r, g in enumerate(ids1)))
    p2.write_text("".join("%s,%d\n" % (g, r + 1) for r, g in enumerate(ids2)))
    pb.write_text("\n".join(bg) + "\n")
    P, seed = 6, 123
    r = _cli(["-1", str(p1), "-2", str(p2), "-b", str(pb), "-p", str(P), "--seed", str(seed)], python)
    assert r.returncode == 0, r.stderr
    l1 = O.OracleList(ids1, list(range(1, 7)))
    l2 = O.OracleList(ids2, list(range(1, 7)))
    N = len(bg)
    recs = [O.optimize(l1, l2, N)]
    i1 = np.stack([dto.philox_permutation(seed, 0, p, 1, 6) for p in range(P)])
    i2 = np.stack([dto.philox_permutation(seed, 0, p, 2, 6) for p in range(P)])
    recs += O.optimize_many(l1, l2, N, i1, i2, fast=False)
    want = []
    for o in recs:
        n_p, n_pk = o.tie_counts()
        if n_p > 1:
            want.append("Multiple results with the same minimum p-value (%.15f). Choosing the result with the "
                        "largest intersection size." % o.pvalue)
            if n_pk > 1:
                want.append("Multiple results with the same maximum intersection size (%d). Choosing an arbitrary "
                            "result based on the order of thresholds." % o.intersection_size)
    assert recs[0].tie_counts()[0] > 1 and any("(0.000000000000000)" in w for w in want)
    got = [ln for ln in r.stderr.splitlines() if ln.startswith("Multiple results")]
    assert got == want


def test_optimize_prints_tie_warnings(dto, O, gpu, capsys):
    """The crate's optimize prints the warnings itself; so does the mirror."""
    l1 = dto.RankedFeatureList.from_(dto.FeatureList(["a", "b", "c"]), [1, 2, 3])
    l2 = dto.RankedFeatureList.from_(dto.FeatureList(["a", "a", "a", "b"]), [1, 2, 3, 4])
    best = dto.optimize(l1, l2, False, 5).best
    o = O.optimize(O.OracleList(["a", "b", "c"], [1, 2, 3]), O.OracleList(["a", "a", "a", "b"], [1, 2, 3, 4]), 5)
    assert (best.rank1, best.rank2, best.intersection_size, best.pvalue) == (o.rank1, o.rank2, o.intersection_size,
                                                                             o.pvalue)
    err = capsys.readouterr().err.splitlines()
    n_p, n_pk = o.tie_counts()
    assert n_p > 1
    assert err[0] == ("Multiple results with the same minimum p-value (0.000000000000000). Choosing the result with "
                      "the largest intersection size.")
    assert len(err) == (2 if n_pk > 1 else 1)
