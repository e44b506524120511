"""GPU parity: the HIP path (through the C ABI) against the CPU oracle on the
same inputs and the same explicit permutations.  Bar: bit-exact records
(thresholds, set sizes, overlap, p-value) and empirical p-value / FDR."""
import ctypes

import numpy as np
import pytest

from helpers import readme_lists, rec_tuple, random_lists, ten_gene_lists

pytestmark = pytest.mark.gpu


def _oracle_pair(O, ids1, r1, ids2, r2):
    return O.OracleList(ids1, r1), O.OracleList(ids2, r2)


def test_device_exp_bit_exact(dto, O, gpu):
    rng = np.random.default_rng(1)
    x = np.concatenate([
        -750.0 * rng.random(400000), -rng.random(200000), -745.2 + 40 * rng.random(200000),
        rng.random(100000) * 1400 - 700, -1e-17 * rng.random(1000),
        np.array([0.0, -0.0, -745.1332191019411, -745.1332191019412, -708.39641853226408, -708.4, -1e-300,
                  -512.0, -511.9999999999999, -1024.0, -1e308, 1.0, 700.0, -np.inf]),
    ])
    want = O.exp_many(x)
    got = np.zeros_like(x)
    L = dto._lib.lib()
    dto._lib.check(L.dto_internal_exp(dto._lib.dblp(x), len(x), dto._lib.dblp(got)))
    bad = np.nonzero(got.view(np.uint64) != want.view(np.uint64))[0]
    assert len(bad) == 0, [(x[i], got[i], want[i]) for i in bad[:5]]


def test_hypergeometric_known_answers(dto, gpu):
    # src/stat_operations/hypergeometric_pvalue.rs:27-31 (exact) and :57-86 (1e-10)
    assert dto.hypergeometric_pvalue(1000, 50, 60, 10) == 0.00044068070222441115
    assert dto.hypergeometric_pvalue(6060, 5808, 154, 153) == 0.010413637619010246
    for args, want in [((3, 1, 1, 1), 0.3333333333333333), ((3, 1, 2, 1), 0.6666666666666666), ((3, 1, 3, 1), 1.0),
                       ((3, 2, 1, 1), 0.6666666666666666), ((3, 2, 2, 2), 0.3333333333333333),
                       ((10, 1, 1, 0), 1.0)]:
        assert abs(dto.hypergeometric_pvalue(*args) - want) < 1e-10
    with pytest.raises(dto._lib.DtoError):
        dto.hypergeometric_pvalue(10, 11, 1, 1)


def test_hypergeometric_random_vs_oracle(dto, O, gpu):
    rng = np.random.default_rng(2)
    Ns = rng.integers(1, 30000, size=3000)
    Ks = np.array([rng.integers(0, n + 1) for n in Ns])
    ns = np.array([rng.integers(0, n + 1) for n in Ns])
    ks = np.array([rng.integers(0, min(K, n) + 2) for K, n in zip(Ks, ns)])
    got = dto.hypergeometric_pvalues(Ns, Ks, ns, ks)
    want = np.array([O.hypergeometric_pvalue(*q) for q in zip(Ns, Ks, ns, ks)])
    assert np.array_equal(got.view(np.uint64), want.view(np.uint64))


def test_philox_device_matches_host(dto, gpu):
    for n in (1, 2, 3, 30, 1000, 4096, 4097, 20000, 32768, 65535):
        for perm in (0, 7):
            host = dto.philox_permutation(99, 3, perm, 1, n)
            dev = np.zeros(n, dtype=np.uint32)
            dto._lib.check(dto._lib.lib().dto_internal_shuffle_device(99, 3, perm, 1, n, dto._lib.u32p(dev)))
            assert np.array_equal(host, dev)
            assert np.array_equal(np.sort(host), np.arange(n, dtype=np.uint32))


def _grid_case(dto, O, ids1, r1, ids2, r2, N, idx1=None, idx2=None):
    job = dto.Job.with_population(ids1, r1, ids2, r2, N)
    got = job.process_threshold_pairs(idx1, idx2)
    l1, l2 = _oracle_pair(O, ids1, r1, ids2, r2)
    want = O.process_threshold_pairs(l1, l2, N, idx1, idx2)
    assert len(got) == len(want)
    for g, w in zip(got, want):
        assert rec_tuple(g) == rec_tuple(w)


def test_grid_readme_lists(dto, O, gpu):
    ids1, r1, ids2, r2 = readme_lists()
    _grid_case(dto, O, ids1, r1, ids2, r2, 30)
    rng = np.random.default_rng(3)
    for _ in range(3):
        _grid_case(dto, O, ids1, r1, ids2, r2, 30, rng.permutation(30), rng.permutation(30))


def test_grid_random_edge_cases(dto, O, gpu):
    rng = np.random.default_rng(4)
    for case in range(24):
        n = int(rng.integers(2, 160))
        ids1, r1, ids2, r2 = random_lists(rng, n, tie_frac=0.2 * (case % 3), dup=(case % 4 == 1),
                                          extra=int(rng.integers(0, 5)) if case % 5 == 2 else 0)
        N = len(set(ids1) | set(ids2)) + int(rng.integers(0, 3))
        p1 = rng.permutation(len(ids1)) if case % 2 else None
        p2 = rng.permutation(len(ids2)) if case % 2 else None
        _grid_case(dto, O, ids1, r1, ids2, r2, N, p1, p2)


@pytest.mark.parametrize("mode", ["plain", "screened"])
def test_optimize_known_answers(dto, mode, gpu):
    g1, r1, g2, r2 = ten_gene_lists()
    job = dto.Job.with_population(g1, r1, g2, r2, 10)
    b = job.optimize(mode=mode)
    assert (int(b["rank1"]), int(b["rank2"]), float(b["pvalue"])) == (3, 4, 0.33333333333333337)
    ids1, r1, ids2, r2 = readme_lists()
    job = dto.Job.from_lists(ids1, r1, ids2, r2)
    b = job.optimize(mode=mode)
    # README.md:174-186: a unique minimum (one cell at its p, hence no tie warning)
    assert rec_tuple(b) == (24, 13, 24, 13, 30, 12, 0.15632183908046102, 0x10001)


@pytest.mark.parametrize("mode", ["plain", "screened"])
def test_optimize_random_vs_oracle(dto, O, mode, gpu):
    rng = np.random.default_rng(5)
    for case in range(40):
        n = int(rng.integers(2, 400))
        ids1, r1, ids2, r2 = random_lists(rng, n, tie_frac=0.3 * (case % 3), dup=(case % 7 == 3),
                                          extra=int(rng.integers(0, 9)) if case % 4 == 2 else 0)
        N = len(set(ids1) | set(ids2)) + int(rng.integers(0, 4))
        job = dto.Job.with_population(ids1, r1, ids2, r2, N)
        l1, l2 = _oracle_pair(O, ids1, r1, ids2, r2)
        for t in range(3):
            p1 = rng.permutation(len(ids1)) if t else None
            p2 = rng.permutation(len(ids2)) if t else None
            got = job.optimize(p1, p2, mode=mode)
            want = O.optimize(l1, l2, N, p1, p2)
            assert rec_tuple(got) == rec_tuple(want), (case, t, n)


def test_list2_duplicates_vs_oracle(dto, O, gpu):
    """Duplicate ids in list 2 only: two list2 entries pair with one list1
    entry, so overlaps can exceed K (empty tails, p == 0).  The screen keeps
    its k > K test for such jobs; results match the oracle."""
    rng = np.random.default_rng(23)
    for case in range(24):
        n = int(rng.integers(6, 300))
        ids1, r1, ids2, r2 = random_lists(rng, n, tie_frac=0.2 * (case % 3))
        for q in range(0, len(ids2) - 1, 3 + case % 4):
            ids2[q + 1] = ids2[q]
        N = len(set(ids1) | set(ids2)) + int(rng.integers(0, 3))
        job = dto.Job.with_population(ids1, r1, ids2, r2, N)
        l1, l2 = _oracle_pair(O, ids1, r1, ids2, r2)
        for t in range(3):
            p1 = rng.permutation(len(ids1)) if t else None
            p2 = rng.permutation(len(ids2)) if t else None
            want = O.optimize(l1, l2, N, p1, p2)
            for mode in ("screened", "plain"):
                assert rec_tuple(job.optimize(p1, p2, mode=mode)) == rec_tuple(want), (case, t, n, mode)


@pytest.mark.parametrize("mode", ["plain", "screened"])
def test_run_plain_6k_explicit(dto, O, mode, gpu):
    from dual_threshold_optimization_amd import workloads as W
    ids1, r1, ids2, r2, _ = W.plain_6k()
    job = dto.Job.from_lists(ids1, r1, ids2, r2)
    P = 6
    i1, i2 = W.explicit_permutations(P, job.len1, job.len2, 11)
    res, perm = job.run(P, idx1=i1, idx2=i2, mode=mode)
    l1, l2 = _oracle_pair(O, ids1, r1, ids2, r2)
    N = O.population_size(l1, l2)
    ub = O.optimize(l1, l2, N, fast=True)
    u = res.unpermuted
    assert rec_tuple(u) == rec_tuple(ub)
    ob = O.optimize_many(l1, l2, N, i1, i2, fast=True)
    for p in range(P):
        assert rec_tuple(perm[p]) == rec_tuple(ob[p]), p
    le = sum(1 for b in ob if b.pvalue <= ub.pvalue)
    assert res.empirical_pvalue == le / P
    assert res.fdr == O.fdr(ub.set1_len, ub.set2_len, ub.intersection_size, N)


def test_run_philox_matches_oracle(dto, O, gpu):
    ids1, r1, ids2, r2 = readme_lists()
    job = dto.Job.from_lists(ids1, r1, ids2, r2)
    P, seed = 64, 7
    res, perm = job.run(P, seed=seed)
    i1 = np.stack([dto.philox_permutation(seed, 0, p, 1, 30) for p in range(P)])
    i2 = np.stack([dto.philox_permutation(seed, 0, p, 2, 30) for p in range(P)])
    l1, l2 = _oracle_pair(O, ids1, r1, ids2, r2)
    ob = O.optimize_many(l1, l2, 30, i1, i2, fast=False)
    for p in range(P):
        assert rec_tuple(perm[p]) == rec_tuple(ob[p]), p


def test_genome_20k_vs_oracle(dto, O, gpu):
    from dual_threshold_optimization_amd import workloads as W
    ids1, r1, ids2, r2, bg = W.genome_20k()
    job = dto.Job.from_lists(ids1, r1, ids2, r2, background=bg, drop_outside_background=True)
    assert job.population_size == 20000 and job.len1 == 20000 and job.dropped1 == 2000
    P = 2
    i1, i2 = W.explicit_permutations(P, job.len1, job.len2, 12)
    res, perm = job.run(P, idx1=i1, idx2=i2)
    keep = set(bg)
    f1 = [(a, b) for a, b in zip(ids1, r1) if a in keep]
    f2 = [(a, b) for a, b in zip(ids2, r2) if a in keep]
    l1 = O.OracleList([a for a, _ in f1], [b for _, b in f1])
    l2 = O.OracleList([a for a, _ in f2], [b for _, b in f2])
    ub = O.optimize(l1, l2, 20000, fast=True)
    u = res.unpermuted
    assert rec_tuple(u) == rec_tuple(ub)
    ob = O.optimize_many(l1, l2, 20000, i1, i2, fast=True)
    for p in range(P):
        assert rec_tuple(perm[p]) == rec_tuple(ob[p]), p


def test_genome_20k_fixture_many_permutations(dto, gpu):
    """The genome-scale job on the 256 explicit permutations whose oracle
    records are committed (tests/golden/make_genome_fixture.py): every
    record, p-value bits and tie counts, in both modes."""
    import json
    import os
    from dual_threshold_optimization_amd import workloads as W
    from helpers import GOLDEN
    with open(os.path.join(GOLDEN, "genome_20k_perms.json")) as f:
        fx = json.load(f)
    ids1, r1, ids2, r2, bg = W.genome_20k()
    job = dto.Job.from_lists(ids1, r1, ids2, r2, background=bg, drop_outside_background=True)
    P = len(fx["permuted"])
    i1, i2 = W.explicit_permutations(P, job.len1, job.len2, 47)

    def dec(d):
        return tuple(d[:6]) + (float.fromhex(d[6]), d[7])
    for mode in ("screened", "plain"):
        res, perm = job.run(P, idx1=i1, idx2=i2, mode=mode)
        assert rec_tuple(res.unpermuted) == dec(fx["unpermuted"]), mode
        bad = [p for p in range(P) if rec_tuple(perm[p]) != dec(fx["permuted"][p])]
        assert not bad, (mode, bad[:5], [rec_tuple(perm[p]) for p in bad[:2]], [fx["permuted"][p] for p in bad[:2]])
        le = sum(1 for d in fx["permuted"] if float.fromhex(d[6]) <= float.fromhex(fx["unpermuted"][6]))
        assert res.n_le == le and res.empirical_pvalue == le / P


def test_screened_equals_plain_genome_20k(dto, gpu):
    """Full-size property: the screened grid returns exactly the plain (every
    cell exact) answer on every permutation."""
    from dual_threshold_optimization_amd import workloads as W
    ids1, r1, ids2, r2, bg = W.genome_20k()
    job = dto.Job.from_lists(ids1, r1, ids2, r2, background=bg, drop_outside_background=True)
    P = 96
    rs, ps = job.run(P, seed=5, mode="screened")
    rp, pp = job.run(P, seed=5, mode="plain")
    assert rs.unpermuted.pvalue == rp.unpermuted.pvalue
    assert (rs.unpermuted.rank1, rs.unpermuted.rank2) == (rp.unpermuted.rank1, rp.unpermuted.rank2)
    assert [rec_tuple(x) for x in ps] == [rec_tuple(x) for x in pp]
    assert rs.empirical_pvalue == rp.empirical_pvalue


def test_screened_equals_plain_many_permutations(dto, gpu):
    """The screened kernel's waves share bounds and run their sweeps and the
    exact tier without barriers between them; thousands of permutations, on
    both permutation paths, must still give exactly the plain records."""
    from dual_threshold_optimization_amd import workloads as W
    ids1, r1, ids2, r2, bg = W.genome_20k()
    job = dto.Job.from_lists(ids1, r1, ids2, r2, background=bg, drop_outside_background=True)
    P = 2048
    rs, ps = job.run(P, seed=17, mode="screened")
    rp, pp = job.run(P, seed=17, mode="plain")
    assert [rec_tuple(x) for x in ps] == [rec_tuple(x) for x in pp]
    assert (rs.n_le, rs.empirical_pvalue) == (rp.n_le, rp.empirical_pvalue)
    i1, i2 = W.explicit_permutations(1024, job.len1, job.len2, 29)
    _, es = job.run(1024, idx1=i1, idx2=i2, mode="screened")
    _, ep = job.run(1024, idx1=i1, idx2=i2, mode="plain")
    assert [rec_tuple(x) for x in es] == [rec_tuple(x) for x in ep]


def test_screened_equals_plain_weak_minima(dto, gpu):
    """Uncorrelated lists have weak minima, where cells below their means
    crowd the screen and the lower-tail bound decides most rescans.  Screened
    records must equal the plain FP64 ones, both for rows held in registers
    (reference thresholds) and for wide rows (over 640 custom thresholds)."""
    rng = np.random.default_rng(41)
    for case, n in enumerate((300, 900, 2500, 6000)):
        ids1, r1, ids2, r2 = random_lists(rng, n, tie_frac=0.1 * (case % 2), extra=n // 10)
        N = len(set(ids1) | set(ids2))
        job = dto.Job.with_population(ids1, r1, ids2, r2, N)
        for wide in (False, True):
            if wide:
                L = len(ids1)
                t = np.arange(1, L + 1, max(1, L // 1200), dtype=np.uint32)
                job.set_thresholds(t, t)
            P = 384
            _, ps = job.run(P, seed=100 + case, mode="screened")
            _, pp = job.run(P, seed=100 + case, mode="plain")
            assert [rec_tuple(x) for x in ps] == [rec_tuple(x) for x in pp], (n, wide)


def test_dense_thresholds_vs_oracle(dto, O, gpu):
    """Custom threshold grids wider than one register segment per lane
    (column constants read from LDS), against the oracle with the same
    thresholds, in both modes."""
    from dual_threshold_optimization_amd import workloads as W
    for n, step, planted in [(2500, 2, 120), (1400, 1, 0), (1800, 3, 60)]:
        ids1, r1, ids2, r2, _ = W.planted_lists(n, planted, 300, 250, 0.05, seed=n)
        job = dto.Job.from_lists(ids1, r1, ids2, r2)
        t1 = np.arange(step, n + 1, step, dtype=np.uint32)
        t2 = np.arange(1, n + 1, max(1, step - 1), dtype=np.uint32)
        job.set_thresholds(t1, t2)
        assert (job.n_thresholds1, job.n_thresholds2) == (len(t1), len(t2))
        l1, l2 = _oracle_pair(O, ids1, r1, ids2, r2)
        l1.set_thresholds(t1)
        l2.set_thresholds(t2)
        P = 3
        i1, i2 = W.explicit_permutations(P, job.len1, job.len2, 13)
        ub = O.optimize(l1, l2, n, fast=True)
        ob = O.optimize_many(l1, l2, n, i1, i2, fast=True)
        for mode in ("screened", "plain"):
            res, perm = job.run(P, idx1=i1, idx2=i2, mode=mode)
            u = res.unpermuted
            assert rec_tuple(u) == rec_tuple(ub), (n, mode)
            for p in range(P):
                assert rec_tuple(perm[p]) == rec_tuple(ob[p]), (n, mode, p)


def test_dense_2k_genome_20k_fixture(dto, gpu):
    """4M-cell grids (thresholds 10..20,000 step 10 on the genome-scale job):
    the committed oracle records (tests/golden/make_dense_fixture.py), then
    screened == plain on further permutations."""
    import json
    import os
    from dual_threshold_optimization_amd import workloads as W
    from helpers import GOLDEN
    with open(os.path.join(GOLDEN, "dense_2k_genome_20k.json")) as f:
        fx = json.load(f)
    ids1, r1, ids2, r2, bg = W.genome_20k()
    job = dto.Job.from_lists(ids1, r1, ids2, r2, background=bg, drop_outside_background=True)
    t1, t2 = W.dense_2k_thresholds()
    job.set_thresholds(t1, t2)
    P = len(fx["permuted"])
    i1, i2 = W.explicit_permutations(P, job.len1, job.len2, 21)

    def dec(d):
        return (d["rank1"], d["rank2"], d["set1_len"], d["set2_len"], d["population_size"],
                d["intersection_size"], float.fromhex(d["pvalue"]), d["ties"])
    for mode in ("screened", "plain"):
        res, perm = job.run(P, idx1=i1, idx2=i2, mode=mode)
        u = res.unpermuted
        assert rec_tuple(u) == dec(fx["unpermuted"]), mode
        for p in range(P):
            assert rec_tuple(perm[p]) == dec(fx["permuted"][p]), (mode, p)
    rs, ps = job.run(24, seed=31, mode="screened")
    rp, pp = job.run(24, seed=31, mode="plain")
    assert [rec_tuple(x) for x in ps] == [rec_tuple(x) for x in pp]


def test_batch_vs_oracle_and_graph_replay(dto, O, gpu):
    from dual_threshold_optimization_amd import workloads as W
    specs = W.batch_jobs(n_jobs=10, seed=9, n_min=60, n_max=1500)
    jobs, lists = [], []
    for ids1, r1, ids2, r2, _ in specs:
        jobs.append(dto.Job.from_lists(ids1, r1, ids2, r2))
        lists.append((ids1, r1, ids2, r2))
    n_perm = [5 + 7 * j for j in range(len(jobs))]
    seed = 31
    b = dto.Batch(jobs, n_perm, seed=seed)
    b.launch()
    first = [(rec_tuple(r.unpermuted), r.n_le) for r in b.results()]
    for _ in range(3):  # graph replays give the same answers
        b.launch()
        again = [(rec_tuple(r.unpermuted), r.n_le) for r in b.results()]
        assert again == first
    for j, (ids1, r1, ids2, r2) in enumerate(lists):
        l1, l2 = _oracle_pair(O, ids1, r1, ids2, r2)
        N = O.population_size(l1, l2)
        P = n_perm[j]
        i1 = np.stack([dto.philox_permutation(seed, j, p, 1, len(ids1)) for p in range(P)])
        i2 = np.stack([dto.philox_permutation(seed, j, p, 2, len(ids2)) for p in range(P)])
        ub = O.optimize(l1, l2, N, fast=True)
        ob = O.optimize_many(l1, l2, N, i1, i2, fast=True)
        le = sum(1 for x in ob if x.pvalue <= ub.pvalue)
        assert first[j] == (rec_tuple(ub), le), j


def test_server_mix_equals_alone(dto, O, gpu):
    from dual_threshold_optimization_amd import workloads as W
    specs = W.batch_jobs(n_jobs=12, seed=13, n_min=60, n_max=1200)
    sv = dto.Server(perms_per_slice=16, seed=77)
    want = {}
    for j, (ids1, r1, ids2, r2, _) in enumerate(specs):
        job = dto.Job.from_lists(ids1, r1, ids2, r2)
        P = 10 + 9 * j
        rid = sv.submit(job, P)
        want[rid] = (ids1, r1, ids2, r2, P)
    got = {}
    for _ in range(2000):
        for rid, res, lat in sv.poll(timeout_ms=50):
            got[rid] = res
            assert lat >= 0
        if len(got) == len(want):
            break
    assert len(got) == len(want)
    for rid, (ids1, r1, ids2, r2, P) in want.items():
        l1, l2 = _oracle_pair(O, ids1, r1, ids2, r2)
        N = O.population_size(l1, l2)
        i1 = np.stack([dto.philox_permutation(77, rid, p, 1, len(ids1)) for p in range(P)])
        i2 = np.stack([dto.philox_permutation(77, rid, p, 2, len(ids2)) for p in range(P)])
        ub = O.optimize(l1, l2, N, fast=True)
        ob = O.optimize_many(l1, l2, N, i1, i2, fast=True)
        le = sum(1 for x in ob if x.pvalue <= ub.pvalue)
        r = got[rid]
        assert (rec_tuple(r.unpermuted), r.n_le) == (rec_tuple(ub), le)
    sv.close()
