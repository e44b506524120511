"""Seeded synthetic DTO workloads (two ranked lists with a planted overlap).

Used by bench.py and the tests.  Sizes follow the benchmark configurations:
  plain_6k    N 6,000 features g00000..g05999, no background; list1 ranks a
              seeded permutation of 1..N; 200 features from list1's top 500
              placed uniformly in list2's top 400, the rest shuffled into the
              remaining positions; 5% min-method ties; 1,000 permutations.
  genome_20k  background of 20,000 ids; each list = background + 2,000
              list-specific extras (ranks 1..22,000); 600 features from
              list1's top 2,000 planted in list2's top 1,500; 10% ties; the
              extras are dropped on upload (N = 20,000); 10,000 permutations.
  batch_512   512 jobs, N log-uniform in [2,000, 20,000], planted size
              uniform in [0, 5% of N], 5% ties, 1,000 Philox permutations.
  dense_2k    N 20,000 with thresholds 10, 20, ..., 20,000 on both lists.
  serving     Poisson arrivals, N log-uniform in [1,000, 30,000], background
              80-100% of the union, P in {100, 1,000, 10,000} w.p. {.5, .4, .1}.
"""
import numpy as np


def _min_ties(n, frac, rng):
    """Ranks of positions 1..n with frac*n positions (from 2..n, without
    replacement) merged into their predecessor; a tie run takes its min rank."""
    ranks = np.arange(1, n + 1, dtype=np.int64)
    k = int(round(frac * n))
    if k and n > 1:
        merged = np.zeros(n, dtype=bool)
        merged[rng.choice(np.arange(1, n), size=min(k, n - 1), replace=False)] = True
        # a merged position takes the rank of the last unmerged one before it
        ranks = np.maximum.accumulate(np.where(merged, 0, ranks))
    return ranks


def planted_lists(n, planted, top1, top2, tie_frac, seed, extras=0, prefix="g"):
    """Returns (ids1, ranks1, ids2, ranks2, background) with `planted` features
    of list1's top `top1` (background features only) placed uniformly in
    list2's top `top2`.  With extras > 0 each list also holds `extras`
    list-specific features and the background is the n shared ids; otherwise
    background is None and both lists hold the same n ids."""
    rng = np.random.default_rng(seed)
    width = max(5, len(str(n + 2 * extras)))
    shared = ["%s%0*d" % (prefix, width, i) for i in range(n)]
    ex1 = ["%sx1_%0*d" % (prefix, width, i) for i in range(extras)]
    ex2 = ["%sx2_%0*d" % (prefix, width, i) for i in range(extras)]
    L = n + extras
    # list1: position -> feature
    order1 = np.array(shared + ex1, dtype=object)[rng.permutation(L)]
    # features of list1's top top1 that exist in list2
    in_bg = {s for s in shared}
    cand = [f for f in order1[:top1] if f in in_bg]
    planted = min(planted, len(cand), top2)
    plant = list(rng.choice(np.array(cand, dtype=object), size=planted, replace=False)) if planted else []
    plant_set = set(plant)
    rest = [f for f in shared + ex2 if f not in plant_set]
    rest = list(np.array(rest, dtype=object)[rng.permutation(len(rest))])
    order2 = [None] * L
    pos = rng.choice(top2, size=planted, replace=False) if planted else []
    for f, p in zip(plant, pos):
        order2[p] = f
    it = iter(rest)
    for p in range(L):
        if order2[p] is None:
            order2[p] = next(it)
    r1 = _min_ties(L, tie_frac, rng)
    r2 = _min_ties(L, tie_frac, rng)
    # emit in a shuffled file order (the loader sorts stably by rank)
    o1 = rng.permutation(L)
    o2 = rng.permutation(L)
    ids1 = [str(order1[i]) for i in o1]
    ids2 = [str(order2[i]) for i in o2]
    ranks1 = r1[o1].astype(np.uint32)
    ranks2 = r2[o2].astype(np.uint32)
    background = shared if extras else None
    return ids1, ranks1, ids2, ranks2, background


def plain_6k(seed=1):
    return planted_lists(6000, 200, 500, 400, 0.05, seed)


def genome_20k(seed=2):
    return planted_lists(20000, 600, 2000, 1500, 0.10, seed, extras=2000)


def dense_2k_thresholds():
    t = np.arange(10, 20001, 10, dtype=np.uint32)
    return t, t.copy()


def batch_jobs(n_jobs=512, seed=3, n_min=2000, n_max=20000, tie_frac=0.05):
    rng = np.random.default_rng(seed)
    out = []
    for j in range(n_jobs):
        n = int(round(np.exp(rng.uniform(np.log(n_min), np.log(n_max)))))
        planted = int(rng.integers(0, int(0.05 * n) + 1))
        top2 = max(2 * planted, 1)
        top1 = max(int(2.5 * planted), 1)
        out.append(planted_lists(n, planted, top1, top2, tie_frac, int(rng.integers(2**31))))
    return out


def explicit_permutations(n_perm, len1, len2, seed):
    """n_perm seeded permutations of each list (rows are permutations of [0, len))."""
    rng = np.random.default_rng(seed)
    a = np.empty((n_perm, len1), dtype=np.uint32)
    b = np.empty((n_perm, len2), dtype=np.uint32)
    for p in range(n_perm):
        a[p] = rng.permutation(len1)
        b[p] = rng.permutation(len2)
    return a, b


def serving_requests(n_req=5000, seed=5, n_min=1000, n_max=30000):
    """(arrival-rank, N, P, planted, background fraction) per request; the
    caller scales arrival times by its offered rate."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0, size=n_req)
    Ns = np.round(np.exp(rng.uniform(np.log(n_min), np.log(n_max), size=n_req))).astype(int)
    Ps = rng.choice([100, 1000, 10000], p=[0.5, 0.4, 0.1], size=n_req)
    planted = [int(rng.integers(0, int(0.05 * n) + 1)) for n in Ns]
    bgf = rng.uniform(0.8, 1.0, size=n_req)
    return np.cumsum(gaps), Ns, Ps, planted, bgf, rng.integers(2**31, size=n_req)


def serving_job_lists(n, planted, bg_frac, seed):
    """Lists of n shared features; the background keeps a bg_frac share of
    them (the rest are dropped from the lists on upload)."""
    ids1, r1, ids2, r2, _ = planted_lists(n, planted, max(int(2.5 * planted), 1), max(2 * planted, 1), 0.05, seed)
    rng = np.random.default_rng(seed + 1)
    keep = sorted(set(ids1))
    m = max(2, int(round(bg_frac * len(keep))))
    bg = list(np.array(keep, dtype=object)[rng.choice(len(keep), size=m, replace=False)])
    return ids1, r1, ids2, r2, [str(x) for x in bg]
