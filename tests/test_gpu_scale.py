"""GPU parity at the sizes of the batched and serving workloads and past the
point where a job's tables leave LDS.

From N ~ 20.5k the screen table no longer fits in LDS next to the pair
staging and the kernels read it from HBM; from N ~ 31k the pair staging
leaves LDS too.  These tests run every such launch shape against the oracle
on the same explicit or Philox permutations (bit-exact records)."""
import numpy as np
import pytest

from helpers import rec_tuple

pytestmark = pytest.mark.gpu


def _filtered_oracle_lists(O, ids1, r1, ids2, r2, bg):
    keep = set(bg)
    f1 = [(a, b) for a, b in zip(ids1, r1) if a in keep]
    f2 = [(a, b) for a, b in zip(ids2, r2) if a in keep]
    return (O.OracleList([a for a, _ in f1], [b for _, b in f1]),
            O.OracleList([a for a, _ in f2], [b for _, b in f2]))


def _genome_like(W, n, seed):
    """Genome-style job of n background features plus 2,000 list-specific
    extras per list (dropped on upload), planted overlap, 10% ties."""
    return W.planted_lists(n, int(0.03 * n), n // 10, (3 * n) // 40, 0.10, seed, extras=2000)


@pytest.mark.parametrize("n", [24000, 30000, 40000])
def test_large_job_vs_oracle(dto, O, gpu, n):
    from dual_threshold_optimization_amd import workloads as W
    ids1, r1, ids2, r2, bg = _genome_like(W, n, 40 + n)
    job = dto.Job.from_lists(ids1, r1, ids2, r2, background=bg, drop_outside_background=True)
    assert job.population_size == n and job.len1 == n and job.len2 == n
    l1, l2 = _filtered_oracle_lists(O, ids1, r1, ids2, r2, bg)
    P = 3
    i1, i2 = W.explicit_permutations(P, n, n, 7 + n)
    ub = O.optimize(l1, l2, n, fast=True)
    ob = O.optimize_many(l1, l2, n, i1, i2, fast=True)
    for mode in ("screened", "plain"):
        print("large job N=%d mode=%s" % (n, mode), flush=True)
        res, perm = job.run(P, idx1=i1, idx2=i2, mode=mode)
        u = res.unpermuted
        assert rec_tuple(u) == rec_tuple(ub), (n, mode)
        for p in range(P):
            assert rec_tuple(perm[p]) == rec_tuple(ob[p]), (n, mode, p)
    # Philox shuffles of the same job: screened against the oracle on the drawn permutations
    res, perm = job.run(4, seed=3 + n)
    j1 = np.stack([dto.philox_permutation(3 + n, 0, p, 1, n) for p in range(4)])
    j2 = np.stack([dto.philox_permutation(3 + n, 0, p, 2, n) for p in range(4)])
    oph = O.optimize_many(l1, l2, n, j1, j2, fast=True)
    for p in range(4):
        assert rec_tuple(perm[p]) == rec_tuple(oph[p]), (n, "philox", p)


def test_batch_uneven_jobs_vs_oracle(dto, O, gpu):
    """The batched workload's job mix (N log-uniform in [2,000, 20,000],
    planted 0-5% of N, 5% ties) in one persistent launch, autotuned, with
    graph replays; every job's records against the oracle on the Philox
    permutations the device drew."""
    from dual_threshold_optimization_amd import workloads as W
    specs = W.batch_jobs(n_jobs=16, seed=33)
    jobs = [dto.Job.from_lists(a, b, c, d) for a, b, c, d, _ in specs]
    n_perm = [4 + (j % 5) for j in range(len(jobs))]
    seed = 19
    b = dto.Batch(jobs, n_perm, seed=seed)
    b.autotune()
    b.launch()
    first = [(rec_tuple(r.unpermuted), r.n_le) for r in b.results()]
    b.launch()
    again = [(rec_tuple(r.unpermuted), r.n_le) for r in b.results()]
    assert again == first
    for j, (ids1, r1, ids2, r2, _) in enumerate(specs):
        l1, l2 = O.OracleList(ids1, r1), O.OracleList(ids2, r2)
        N = O.population_size(l1, l2)
        P = n_perm[j]
        i1 = np.stack([dto.philox_permutation(seed, j, p, 1, len(ids1)) for p in range(P)])
        i2 = np.stack([dto.philox_permutation(seed, j, p, 2, len(ids2)) for p in range(P)])
        ub = O.optimize(l1, l2, N, fast=True)
        ob = O.optimize_many(l1, l2, N, i1, i2, fast=True)
        le = sum(1 for x in ob if x.pvalue <= ub.pvalue)
        assert first[j] == (rec_tuple(ub), le), (j, len(ids1))


def test_server_mix_with_large_jobs(dto, O, gpu):
    """Serving requests of the serving workload's shape (N up to 30,000,
    background 80-100% of the union, dropped on upload) mixed with small
    ones: each result equals the oracle's on the permutations the device
    drew, whatever launch shape the other requests in its slices need."""
    from dual_threshold_optimization_amd import workloads as W
    sizes = [1200, 24000, 3000, 30000, 800, 9000]
    sv = dto.Server(perms_per_slice=4, seed=41)
    want = {}
    for j, n in enumerate(sizes):
        ids1, r1, ids2, r2, bg = W.serving_job_lists(n, int(0.02 * n), 0.8 + 0.04 * j, 500 + j)
        job = dto.Job.from_lists(ids1, r1, ids2, r2, background=bg, drop_outside_background=True)
        P = 3 + j
        rid = sv.submit(job, P)
        want[rid] = (ids1, r1, ids2, r2, bg, P)
    got = {}
    for _ in range(4000):
        for rid, res, lat in sv.poll(timeout_ms=50):
            got[rid] = res
        if len(got) == len(want):
            break
    assert len(got) == len(want)
    for rid, (ids1, r1, ids2, r2, bg, P) in want.items():
        l1, l2 = _filtered_oracle_lists(O, ids1, r1, ids2, r2, bg)
        N = len(bg)
        i1 = np.stack([dto.philox_permutation(41, rid, p, 1, l1.n) for p in range(P)])
        i2 = np.stack([dto.philox_permutation(41, rid, p, 2, l2.n) for p in range(P)])
        ub = O.optimize(l1, l2, N, fast=True)
        ob = O.optimize_many(l1, l2, N, i1, i2, fast=True)
        le = sum(1 for x in ob if x.pvalue <= ub.pvalue)
        r = got[rid]
        assert (rec_tuple(r.unpermuted), r.n_le) == (rec_tuple(ub), le), (rid, N)
    sv.close()


def test_concurrent_launches_of_one_large_job(dto, O, gpu):
    """A job whose pair staging lives in global memory (N = 36,000), launched
    on two streams at once with different permutations: the launches share
    the job's staging buffer, so the library orders them; both outputs equal
    the oracle's."""
    import torch
    from dual_threshold_optimization_amd import _lib
    from dual_threshold_optimization_amd import workloads as W
    n = 36000
    ids1, r1, ids2, r2, bg = _genome_like(W, n, 77)
    job = dto.Job.from_lists(ids1, r1, ids2, r2, background=bg, drop_outside_background=True)
    l1, l2 = _filtered_oracle_lists(O, ids1, r1, ids2, r2, bg)
    dev = torch.device("cuda:0")
    P = 3
    runs = []
    for seed in (5, 6):
        i1, i2 = W.explicit_permutations(P, n, n, seed)
        runs.append((i1, i2, torch.from_numpy(i1.astype(np.int32)).to(dev), torch.from_numpy(i2.astype(np.int32)).to(dev),
                     torch.zeros(P * _lib.RECORD_DTYPE.itemsize, dtype=torch.uint8, device=dev), torch.cuda.Stream()))
    torch.cuda.synchronize()
    for _ in range(2):
        for i1, i2, d1, d2, out, st in runs:
            job.run_device(P, out.data_ptr(), d_idx1=d1.data_ptr(), d_idx2=d2.data_ptr(), stream=st.cuda_stream)
    torch.cuda.synchronize()
    for i1, i2, d1, d2, out, st in runs:
        got = np.frombuffer(out.cpu().numpy().tobytes(), dtype=_lib.RECORD_DTYPE)
        want = O.optimize_many(l1, l2, n, i1, i2, fast=True)
        for p in range(P):
            assert rec_tuple(got[p]) == rec_tuple(want[p]), p


def test_server_pool_admission_deferred_jobs(dto, O, gpu):
    """Jobs built with defer_upload stay on the host until admitted; a small
    HBM budget admits a few at a time (the rest wait in the queue), and
    requests of different launch shapes run in their own slices.  Every
    result equals the oracle's."""
    from dual_threshold_optimization_amd import workloads as W
    sizes = [900, 2500, 700, 22000, 1500, 600, 4000, 1100]
    # 1 MiB: the 22,000-feature job's tables (~0.6 MB) fit only beside a few small ones
    sv = dto.Server(hbm_budget_bytes=1 << 20, perms_per_slice=3, seed=9)
    want = {}
    for j, n in enumerate(sizes):
        ids1, r1, ids2, r2, bg = W.serving_job_lists(n, int(0.03 * n), 0.9, 900 + j)
        job = dto.Job.from_lists(ids1, r1, ids2, r2, background=bg, drop_outside_background=True, defer_upload=True)
        P = 2 + (j % 4)
        rid = sv.submit(job, P)
        want[rid] = (ids1, r1, ids2, r2, bg, P)
    q, a = sv.pending()
    assert q == len(sizes) and a == 0
    got = {}
    for _ in range(4000):
        for rid, res, lat in sv.poll(timeout_ms=50):
            got[rid] = res
        if len(got) == len(want):
            break
    assert len(got) == len(want)
    assert sv.pending() == (0, 0)
    for rid, (ids1, r1, ids2, r2, bg, P) in want.items():
        l1, l2 = _filtered_oracle_lists(O, ids1, r1, ids2, r2, bg)
        N = len(bg)
        i1 = np.stack([dto.philox_permutation(9, rid, p, 1, l1.n) for p in range(P)])
        i2 = np.stack([dto.philox_permutation(9, rid, p, 2, l2.n) for p in range(P)])
        ub = O.optimize(l1, l2, N, fast=True)
        ob = O.optimize_many(l1, l2, N, i1, i2, fast=True)
        le = sum(1 for x in ob if x.pvalue <= ub.pvalue)
        r = got[rid]
        assert (rec_tuple(r.unpermuted), r.n_le) == (rec_tuple(ub), le), (rid, N)
    sv.close()


def test_deferred_job_runs_on_first_use(dto, O, gpu):
    """A deferred job used outside a server uploads itself on first use."""
    from dual_threshold_optimization_amd import workloads as W
    ids1, r1, ids2, r2, _ = W.planted_lists(800, 30, 100, 80, 0.05, seed=5)
    job = dto.Job.from_lists(ids1, r1, ids2, r2, defer_upload=True)
    l1, l2 = O.OracleList(ids1, r1), O.OracleList(ids2, r2)
    N = O.population_size(l1, l2)
    assert rec_tuple(job.optimize()) == rec_tuple(O.optimize(l1, l2, N, fast=True))


def test_large_job_with_list2_duplicates(dto, O, gpu):
    """N = 24,000 (screen table in LDS, bins global) with duplicate ids in
    list 2: the CSR join path, k > K cells (p == 0) and their tie counts,
    in both modes, against the oracle."""
    from dual_threshold_optimization_amd import workloads as W
    n = 24000
    ids1, r1, ids2, r2, _ = W.planted_lists(n, 700, 2400, 1800, 0.1, seed=91)
    rng = np.random.default_rng(92)
    for q in rng.choice(len(ids2) - 1, size=400, replace=False):
        ids2[q + 1] = ids2[q]
    N = n + 50
    job = dto.Job.with_population(ids1, r1, ids2, r2, N)
    l1, l2 = O.OracleList(ids1, r1), O.OracleList(ids2, r2)
    P = 2
    i1, i2 = W.explicit_permutations(P, n, n, 93)
    ub = O.optimize(l1, l2, N, fast=True)
    ob = O.optimize_many(l1, l2, N, i1, i2, fast=True)
    for mode in ("screened", "plain"):
        res, perm = job.run(P, idx1=i1, idx2=i2, mode=mode)
        assert rec_tuple(res.unpermuted) == rec_tuple(ub), mode
        for p in range(P):
            assert rec_tuple(perm[p]) == rec_tuple(ob[p]), (mode, p)
