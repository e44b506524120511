"""The oracle (oracle/dto_oracle.c) pinned before it is trusted (CPU only).

Each check is independent of the oracle's own formulas: the reference's
known-answer values, exact rational arithmetic, Python's lgamma, brute-force
set intersection on tiny inputs, and the reference's documented error
behaviour.  A dropped term, wrong index or transposed operand in the oracle
fails at least one of them.
"""
import math
import random
from fractions import Fraction

import numpy as np
import pytest

from helpers import readme_lists, ten_gene_lists, random_lists


def exact_sf(N, K, n, k):
    """P(X >= k) for X ~ Hypergeometric(N, K, n), exact."""
    lo, hi = max(0, n + K - N), min(K, n)
    if k <= lo:
        return Fraction(1)
    tot = math.comb(N, n)
    return Fraction(sum(math.comb(K, i) * math.comb(N - K, n - i) for i in range(max(k, lo), hi + 1)), tot)


# ---- A9 / A8: the statrs replay ------------------------------------------------

def test_hypergeometric_kats(O):
    # src/stat_operations/hypergeometric_pvalue.rs:26-31 (exact, Lanczos branch)
    assert O.hypergeometric_pvalue(1000, 50, 60, 10) == 0.00044068070222441115
    assert O.hypergeometric_pvalue(6060, 5808, 154, 153) == 0.010413637619010246


def test_hypergeometric_small_cases(O):
    # src/stat_operations/hypergeometric_pvalue.rs:59-86 (|p - expected| < 1e-10)
    for N, K, n, k, want in [(3, 1, 1, 1, 0.3333333333333333), (3, 1, 2, 1, 0.6666666666666666),
                             (3, 1, 3, 1, 1.0), (3, 2, 1, 1, 0.6666666666666666),
                             (3, 2, 2, 2, 0.3333333333333333), (10, 1, 1, 0, 1.0)]:
        assert abs(O.hypergeometric_pvalue(N, K, n, k) - want) < 1e-10


def test_hypergeometric_invalid(O):
    # Hypergeometric::new(N, K, n).expect(...) panics when K > N or n > N
    with pytest.raises(ValueError):
        O.hypergeometric_pvalue(10, 11, 3, 1)
    with pytest.raises(ValueError):
        O.hypergeometric_pvalue(10, 3, 11, 1)


def test_hypergeometric_vs_exact_rationals(O):
    """Across the table branch (x <= 170) and the Lanczos branch (x > 170) the
    replay stays within ~1e-11 relative of the exact tail; p == 1 and p == 0
    edges are exact."""
    rng = random.Random(11)
    cases = [(30, 10, 12, 5), (200, 40, 60, 20), (171, 85, 85, 50), (172, 86, 86, 50), (500, 250, 250, 140),
             (2000, 100, 300, 40), (6000, 400, 500, 100)]
    for _ in range(60):
        N = rng.randint(2, 700)
        K, n = rng.randint(0, N), rng.randint(0, N)
        lo, hi = max(0, n + K - N), min(K, n)
        cases.append((N, K, n, rng.randint(lo, hi)))
    for N, K, n, k in cases:
        got = O.hypergeometric_pvalue(N, K, n, k)
        want = exact_sf(N, K, n, k)
        if want == 1:
            assert got == 1.0
        elif want == 0:
            assert got == 0.0
        else:
            w = float(want)
            if w > 1e-290:
                assert abs(got - w) <= 2e-11 * w + 1e-300, (N, K, n, k, got, w)
    # k above min(K, n): sf returns exactly 0 (hypergeometric_pvalue.rs:48, x >= max)
    assert O.hypergeometric_pvalue(100, 10, 20, 11) == 0.0
    # k at or below max(0, n + K - N): exactly 1
    assert O.hypergeometric_pvalue(100, 90, 20, 10) == 1.0


def test_ln_factorial_and_gamma(O):
    for x in list(range(0, 400)) + [1000, 6060, 20000, 123456]:
        want = math.lgamma(x + 1)
        got = O.ln_factorial(x)
        assert abs(got - want) <= 1e-13 * max(1.0, abs(want)), x
    for x in [0.5, 1.0, 1.5, 7.25, 171.0, 1e4 + 0.5]:
        assert abs(O.ln_gamma(x) - math.lgamma(x)) <= 1e-13 * max(1.0, abs(math.lgamma(x)))
    # table branch: log of the running f64 product (x <= 170) exactly
    fc = 1.0
    for x in range(1, 171):
        fc *= x
        assert O.ln_factorial(x) == math.log(fc)


# ---- A2: thresholds -------------------------------------------------------------

def test_thresholds_properties(O):
    assert O.generate_thresholds(1) == [1]  # ranked.rs:258-259: thresholds[0] == 1
    assert O.generate_thresholds(0) == []
    for m in [2, 3, 10, 100, 6000, 20000, 30000]:
        t = O.generate_thresholds(m)
        assert t[0] == 1 and t[-1] <= m
        assert all(b > a for a, b in zip(t, t[1:]))
        # the next value would pass the maximum rank
        assert int(t[-1] * 1.01 + 1.0) > m
        # the step is floor(T * 1.01) + 1 >= 1, so small lists use every rank 1..m
        if m <= 100:
            assert t == list(range(1, m + 1))
    # counts at the measured sizes
    assert len(O.generate_thresholds(6000)) == 469
    t = O.generate_thresholds(20000)
    assert len(t) == 589 and t[-1] == 19850
    assert len(O.generate_thresholds(30000)) == 630


# ---- A6 / A7 / A10: the grid and its optimum ------------------------------------

def brute_grid(ids1, r1, ids2, r2, N, idx1=None, idx2=None, T1=None, T2=None):
    """Per-cell set intersection over stably sorted lists (exact rationals).
    T1/T2: the threshold values to evaluate (default: every integer from 1 to
    the list's largest rank); a set holds the features of rank <= threshold."""
    o1 = sorted(range(len(r1)), key=lambda i: r1[i])
    o2 = sorted(range(len(r2)), key=lambda i: r2[i])
    g1, s1 = [ids1[i] for i in o1], [int(r1[i]) for i in o1]
    g2, s2 = [ids2[i] for i in o2], [int(r2[i]) for i in o2]
    p1 = g1 if idx1 is None else [g1[int(i)] for i in idx1]
    p2 = g2 if idx2 is None else [g2[int(i)] for i in idx2]
    if T1 is None:
        T1 = [t for t in range(1, s1[-1] + 1)]
    if T2 is None:
        T2 = [t for t in range(1, s2[-1] + 1)]
    out = []
    for a in T1:
        A = [p1[q] for q in range(len(p1)) if s1[q] <= a]
        setA = set(A)
        for b in T2:
            B = [p2[q] for q in range(len(p2)) if s2[q] <= b]
            k = sum(1 for x in B if x in setA)  # duplicates in list 2 count each time
            out.append((a, b, len(A), len(B), k))
    return out


def test_readme_kat(O):
    # README.md:174-186: rank1 24, rank2 13, set lengths 24 / 13, N 30, k 12, p 0.15632183908046102
    i1, r1, i2, r2 = readme_lists()
    l1, l2 = O.OracleList(i1, r1), O.OracleList(i2, r2)
    N = O.population_size(l1, l2)
    assert N == 30
    best = O.optimize(l1, l2, N)
    assert best.as_tuple() == (24, 13, 24, 13, 30, 12, 0.15632183908046102)
    assert best.tie_counts() == (1, 1)  # README.md:174-186 shows no tie warning
    assert O.optimize(l1, l2, N, fast=True).as_tuple() == best.as_tuple()
    # the exact-rational optimum is the same cell
    ex = min(((exact_sf(30, A, B, k), -k, a, b) for a, b, A, B, k in brute_grid(i1, r1, i2, r2, 30)))
    assert (ex[2], ex[3], -ex[1]) == (24, 13, 12)


def test_ten_gene_kat(O):
    # src/dto/optimize_main.rs:126-175: best (3, 4), p 0.33333333333333337
    g1, r1, g2, r2 = ten_gene_lists()
    l1, l2 = O.OracleList(g1, r1), O.OracleList(g2, r2)
    N = O.population_size(l1, l2)
    best = O.optimize(l1, l2, N)
    assert (best.rank1, best.rank2, best.pvalue) == (3, 4, 0.33333333333333337)
    assert best.intersection_size == 2
    assert len(O.process_threshold_pairs(l1, l2, N)) == 100


@pytest.mark.parametrize("seed", range(6))
def test_grid_vs_brute_force(O, seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(3, 40))
    i1, r1, i2, r2 = random_lists(rng, n, dup=seed % 2 == 1, max_rank=int(rng.integers(n, 2 * n)))
    l1, l2 = O.OracleList(i1, r1), O.OracleList(i2, r2)
    N = 50
    idx1 = idx2 = None
    if seed >= 3:
        idx1, idx2 = rng.permutation(len(i1)), rng.permutation(len(i2))
    recs = O.process_threshold_pairs(l1, l2, N, idx1, idx2)
    t1, t2 = l1.thresholds, l2.thresholds
    brute = {(a, b): (A, B, k) for a, b, A, B, k in brute_grid(i1, r1, i2, r2, N, idx1, idx2)}
    assert len(recs) == len(t1) * len(t2)
    for c, r in enumerate(recs):
        a, b = t1[c // len(t2)], t2[c % len(t2)]
        assert (r.rank1, r.rank2) == (a, b)
        assert (r.set1_len, r.set2_len, r.intersection_size) == brute[(a, b)]
        assert r.population_size == N
        assert r.pvalue == O.hypergeometric_pvalue(N, r.set1_len, r.set2_len, r.intersection_size)
    # the optimum: min p, then max k, then min (rank1, rank2)
    best = O.optimize(l1, l2, N, idx1, idx2)
    key = min((r.pvalue, -r.intersection_size, r.rank1, r.rank2) for r in recs)
    assert (best.pvalue, -best.intersection_size, best.rank1, best.rank2) == key
    # tie counts behind optimize_main.rs:86-108's two warnings
    at_min = [r for r in recs if r.pvalue == key[0]]
    n_pk = sum(1 for r in at_min if r.intersection_size == -key[1])
    assert best.tie_counts() == (len(at_min), n_pk)
    assert O.optimize(l1, l2, N, idx1, idx2, fast=True).tie_counts() == (len(at_min), n_pk)


def test_tie_counts_p_one_and_p_zero_classes(O):
    """Grids whose minimum is shared: every cell at p == 1 (disjoint sets),
    and p == 0 cells (empty tails from list-2 duplicates); the counts follow
    the cells' (p, k) exactly."""
    # no shared feature anywhere within the thresholds: every k = 0, every p = 1
    l1 = O.OracleList(["a", "b", "c", "d"], [1, 2, 3, 4])
    l2 = O.OracleList(["x", "y", "z", "w"], [1, 2, 3, 4])
    recs = O.process_threshold_pairs(l1, l2, 12)
    b = O.optimize(l1, l2, 12)
    assert b.pvalue == 1.0 and all(r.pvalue == 1.0 for r in recs)
    kmax = max(r.intersection_size for r in recs)
    assert b.tie_counts() == (len(recs), sum(1 for r in recs if r.intersection_size == kmax))
    # list-2 duplicates of one list-1 id: k > K cells have p == 0 exactly
    l1 = O.OracleList(["a", "b", "c"], [1, 2, 3])
    l2 = O.OracleList(["a", "a", "a", "b"], [1, 2, 3, 4])
    recs = O.process_threshold_pairs(l1, l2, 5)
    b = O.optimize(l1, l2, 5)
    zero = [r for r in recs if r.pvalue == 0.0]
    assert b.pvalue == 0.0 and len(zero) >= 2
    kmax = max(r.intersection_size for r in zero)
    assert b.tie_counts() == (len(zero), sum(1 for r in zero if r.intersection_size == kmax))


@pytest.mark.parametrize("seed", range(8))
def test_custom_thresholds_vs_brute_force(O, seed):
    """or_list_set_thresholds (custom threshold lists, which replace
    generate_thresholds in ranked.rs:359-375 and feed the set rule of
    ranked.rs:561-568 / permuted.rs:90-101): the records on explicit threshold
    lists equal per-cell set intersection with sets {rank <= t}.  The lists
    cover thresholds between ranks, equal to a tied rank, below the smallest
    rank (an empty set), above the largest rank, and a single threshold."""
    rng = np.random.default_rng(200 + seed)
    n = int(rng.integers(3, 36))
    i1, r1, i2, r2 = random_lists(rng, n, tie_frac=0.3, dup=seed % 3 == 2, max_rank=int(rng.integers(n, 3 * n)))
    l1, l2 = O.OracleList(i1, r1), O.OracleList(i2, r2)
    m1, m2 = int(max(r1)), int(max(r2))

    def pick(m, case):
        if case == 0:
            return [m + 5]                                   # one threshold, above every rank
        if case == 1:
            return [int(min(r1 if m == m1 else r2))]         # one threshold, the smallest rank
        vals = set(int(x) for x in rng.choice(np.arange(0, 3 * m + 2), size=min(3 * m, 12), replace=False))
        vals |= {int(x) for x in rng.choice(r1 if m == m1 else r2, size=3)}  # tied / exact ranks
        vals.add(m + 1 + int(rng.integers(0, 5)))              # past the largest rank
        return sorted(vals)
    t1 = np.array(pick(m1, seed % 4), dtype=np.uint32)
    t2 = np.array(pick(m2, (seed + 1) % 4), dtype=np.uint32)
    l1.set_thresholds(t1)
    l2.set_thresholds(t2)
    assert l1.thresholds == list(t1) and l2.thresholds == list(t2)
    N = 2 * n + 10
    for idx1, idx2 in ((None, None), (rng.permutation(len(i1)), rng.permutation(len(i2)))):
        recs = O.process_threshold_pairs(l1, l2, N, idx1, idx2)
        brute = brute_grid(i1, r1, i2, r2, N, idx1, idx2, T1=list(t1), T2=list(t2))
        assert len(recs) == len(brute) == len(t1) * len(t2)
        for r, (a, b, A, B, k) in zip(recs, brute):
            assert (r.rank1, r.rank2, r.set1_len, r.set2_len, r.intersection_size) == (a, b, A, B, k)
            assert r.pvalue == O.hypergeometric_pvalue(N, A, B, k)
        k, p = O.fast_grid(l1, l2, N, idx1, idx2)
        assert np.array_equal(k.ravel(), [x[4] for x in brute])
        best = O.optimize(l1, l2, N, idx1, idx2, fast=True)
        key = min((r.pvalue, -r.intersection_size, r.rank1, r.rank2) for r in recs)
        assert (best.pvalue, -best.intersection_size, best.rank1, best.rank2) == key
    with pytest.raises(ValueError):
        l1.set_thresholds(np.array([3, 3], dtype=np.uint32))


def test_duplicate_id_asymmetry(O):
    """intersect_genes.rs:38-56: a duplicated id in list 2 counts twice, in
    list 1 once."""
    l1 = O.OracleList(["a", "b", "c", "d"], [1, 2, 3, 4])
    l2 = O.OracleList(["a", "a", "c", "x"], [1, 2, 3, 4])
    recs = O.process_threshold_pairs(l1, l2, 10)
    cell = {(r.rank1, r.rank2): r.intersection_size for r in recs}
    assert cell[(1, 2)] == 2 and cell[(4, 4)] == 3
    l1 = O.OracleList(["a", "a", "c", "x"], [1, 2, 3, 4])
    l2 = O.OracleList(["a", "b", "c", "d"], [1, 2, 3, 4])
    cell = {(r.rank1, r.rank2): r.intersection_size for r in O.process_threshold_pairs(l1, l2, 10)}
    assert cell[(2, 1)] == 1 and cell[(4, 4)] == 2


@pytest.mark.parametrize("seed", range(4))
def test_fast_oracle_equals_reference_oracle(O, seed):
    """The histogram + memoised restatement gives the reference-structured
    restatement's records bit for bit."""
    rng = np.random.default_rng(100 + seed)
    n = int(rng.integers(50, 400))
    i1, r1, i2, r2 = random_lists(rng, n, dup=seed == 3, extra=int(rng.integers(0, 20)))
    l1, l2 = O.OracleList(i1, r1), O.OracleList(i2, r2)
    N = n + 40
    idx1, idx2 = rng.permutation(len(i1)), rng.permutation(len(i2))
    for a, b in ((None, None), (idx1, idx2)):
        recs = O.process_threshold_pairs(l1, l2, N, a, b)
        k, p = O.fast_grid(l1, l2, N, a, b)
        assert np.array_equal(k.ravel(), [r.intersection_size for r in recs])
        assert np.array_equal(p.ravel().view(np.uint64), np.array([r.pvalue for r in recs]).view(np.uint64))
        assert O.optimize(l1, l2, N, a, b).as_tuple() == O.optimize(l1, l2, N, a, b, fast=True).as_tuple()
    many = O.optimize_many(l1, l2, N, np.stack([idx1, idx1[::-1]]), np.stack([idx2, idx2[::-1]]), threads=2)
    assert many[0].as_tuple() == O.optimize(l1, l2, N, idx1, idx2).as_tuple()
    assert many[1].as_tuple() == O.optimize(l1, l2, N, idx1[::-1], idx2[::-1]).as_tuple()


# ---- A12: population size --------------------------------------------------------

def test_population_size_rules(O):
    # src/dto/compute_population_size.rs:111-258
    g = ["gene1", "gene2", "gene3"]
    r = [1, 2, 3]
    l1, l2 = O.OracleList(g, r), O.OracleList(g, r)
    assert O.population_size(l1, l2) == 3
    assert O.population_size(l1, l2, g + ["gene4"]) == 4
    with pytest.raises(ValueError, match="identical genes"):
        O.population_size(l1, O.OracleList(["gene1", "gene2", "gene4"], r))
    with pytest.raises(ValueError, match="first ranked feature list"):
        O.population_size(l1, l2, ["gene1", "gene2"])
    with pytest.raises(ValueError, match="second ranked feature list"):
        O.population_size(l1, O.OracleList(["gene1", "gene2", "gene4"], r), g)
    # background lines count as they are, duplicates and blank lines included
    assert O.population_size(l1, l2, g + ["gene1", ""]) == 5


# ---- A14: empirical p-value and FDR -------------------------------------------------

def test_fdr_and_empirical(O):
    # src/stat_operations/fdr.rs:19-27
    assert abs(O.fdr(200, 150, 50, 1000, 0.8) - 0.240625) < 1e-4
    # closed form with exact rationals: max(0, l1 - k/s) max(0, l2 - k/s) / (N k)
    for l1, l2, k, N in [(200, 150, 50, 1000), (24, 13, 12, 30), (50, 40, 10, 100), (5, 5, 10, 100)]:
        df = Fraction(k) / Fraction(0.8)
        want = max(Fraction(0), l1 - df) * max(Fraction(0), l2 - df) / (N * k)
        assert abs(O.fdr(l1, l2, k, N) - float(want)) <= 1e-15 * max(1.0, float(want))
    assert O.fdr(10, 10, 0, 100) == 0.0
    assert O.fdr(50, 40, 10, 100) == 1.03125  # empirical_pvalue.rs:104
    # src/stat_operations/empirical_pvalue.rs:40-107: permuted p 0.10, 0.03, 0.07 vs 0.05
    assert O.empirical_pvalue(0.05, [0.10, 0.03, 0.07]) == 1 / 3
    assert O.empirical_pvalue(0.05, []) == 1.0
    assert O.empirical_pvalue(0.05, [0.05, 0.05]) == 1.0  # <= counts ties


def test_glibc_exp_fixture(O):
    """The oracle's exp is the host glibc; the device replica is checked
    against it elsewhere.  A few values exact in binary pin it to IEEE exp."""
    x = np.array([0.0, -0.0, 1e-300, -745.2, 710.0, -np.inf], dtype=np.float64)
    y = O.exp_many(x)
    assert y[0] == 1.0 and y[1] == 1.0 and y[2] == 1.0 and y[3] == 0.0 and np.isinf(y[4]) and y[5] == 0.0
    assert O.exp_many(np.array([1.0]))[0] == math.e
