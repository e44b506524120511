"""The screen's lower-tail bound, checked on the CPU against exact tails.

The screened kernels raise a cell's lower bound of log p to log(1/2) when an
integer test says the cell lies at least max(sigma, 1) - 1 below its mean
(lift_offset / lift_lower in csrc/dto_kernels.h); Cantelli's inequality then
gives P(X >= k) >= 1/2.  Rows held in registers use a scaled test (4 k N +
4 D <= K 4n) and a per-lane variance bound; wider rows the plain test (k N +
D <= K n) and the bound over all n.  Here the same test, in the same float32 and u32
arithmetic, is run over whole (N, K, n) grids and every cell it lifts is
checked against scipy's hypergeometric tail.  The check fails a test that
lifts cells at or above the median, e.g. one whose offset D4 moves the bound
to the wrong side of the mean (planted below to show it).
"""
import numpy as np
import pytest
from scipy.stats import hypergeom

LIFT_MAX_N = 32000  # kLiftMaxN (scaled test)
LIFT_MAX_N1 = 65000  # kLiftMaxN1 (plain test)


def lane_var_bound(N, ns):
    """The lane's bound of n (N - n) / (N - 1) over its columns (float32)."""
    Nf = np.float32(N)
    nf = np.asarray(ns, dtype=np.float32)
    v = np.max(nf * (Nf - nf))
    return np.float32(v / (Nf - np.float32(1))) if N > 1 else np.float32(0)


def row_var_bound(N):
    """The wide rows' bound of n (N - n) / (N - 1) over all n (float32)."""
    Nf = np.float32(N)
    return np.float32(np.float32(Nf * Nf / (np.float32(4) * (Nf - np.float32(1)))) * np.float32(1.001))


def lift_offset(K, N, V):
    """D of the kernel: ceil(N (sigma' - 1) + 1), sigma' >= max(sigma, 1)."""
    Nf = np.float32(N)
    invN2 = np.float32(1) / (Nf * Nf)
    s = np.float32(K) * np.float32(N - K) * invN2
    sg = np.float32(np.sqrt(np.float32(s * V)))
    sg = max(np.float32(sg * np.float32(1.001) + np.float32(1e-3)), np.float32(1))
    d = np.float32(min(np.float32((sg - np.float32(1)) * Nf + np.float32(1)), np.float32(1e9)))
    return int(np.ceil(d))


def lifted_cells(N, K, n, V, offset_side=1, scale=4):
    """Overlaps k in [0, min(K, n)] that the kernel's test lifts (scale 4:
    rows in registers, 1: wider rows; offset_side -1: the planted mistake,
    D on the other side)."""
    D = (lift_offset(K, N, V) * scale) & 0xFFFFFFFF
    k = np.arange(0, min(K, n) + 1, dtype=np.uint64)
    lhs = (k * np.uint64(scale * N) + np.uint64(D if offset_side > 0 else 0)) & np.uint64(0xFFFFFFFF)
    rhs = np.uint64(scale * n * K + (D if offset_side < 0 else 0)) & np.uint64(0xFFFFFFFF)
    return k[lhs <= rhs].astype(np.int64)


def violations(N, K, n, V, offset_side=1, scale=4):
    k = lifted_cells(N, K, n, V, offset_side, scale)
    if k.size == 0:
        return 0, 0
    p = hypergeom.sf(k - 1, N, K, n)  # P(X >= k)
    return int(np.sum(p < 0.5 - 1e-12)), int(k.size)


def grid_cases(seed, count):
    rng = np.random.default_rng(seed)
    out = []
    for N in (2, 3, 10, 57, 500, 2000, 20000, LIFT_MAX_N):
        for _ in range(count):
            K = int(rng.integers(1, N + 1))
            n = int(rng.integers(1, N + 1))
            out.append((N, K, n))
        out += [(N, 1, N), (N, N, 1), (N, N // 2 or 1, N // 2 or 1), (N, N - 1 or 1, N - 1 or 1)]
    return out


def test_lower_tail_lift_is_a_valid_bound():
    lifted = 0
    for N, K, n in grid_cases(7, 12):
        # a lane holding only this column (the tightest bound), and a lane
        # of ten geometric-ish columns around it (the kernel's usual case)
        ns = sorted({max(1, min(N, int(n * f))) for f in (0.95, 0.97, 0.99, 1.0, 1.01, 1.03)})
        for V in (lane_var_bound(N, [n]), lane_var_bound(N, ns)):
            bad, cnt = violations(N, K, n, V)
            assert bad == 0, (N, K, n)
            lifted += cnt
    assert lifted > 10000  # the test lifts cells; it is not vacuous


def test_lower_tail_lift_reaches_the_far_tail():
    # at N = 20,000 and K = n = 10,000 (mean 5,000, sigma ~35) every cell
    # 2 sigma below the mean is lifted, and no cell at or above the mean is
    N, K, n = 20000, 10000, 10000
    k = lifted_cells(N, K, n, lane_var_bound(N, [n]))
    sigma = np.sqrt(n * (K / N) * (1 - K / N) * (N - n) / (N - 1))
    assert k.max() >= 5000 - 2 * sigma
    assert k.max() < 5000
    # p == 0 cells (k > K) and the upper tail are never lifted
    assert lifted_cells(N, 100, 5000, lane_var_bound(N, [5000])).max() < 100 * 5000 / N


def test_lower_tail_lift_check_catches_a_wrong_side_offset():
    # planted mistake: k <= mu + sigma' - 1 lifts cells above the median
    bad = 0
    for N, K, n in grid_cases(7, 12):
        bad += violations(N, K, n, lane_var_bound(N, [n]), offset_side=-1)[0]
    assert bad > 0


def test_lower_tail_lift_wide_rows_is_a_valid_bound():
    lifted = 0
    cases = grid_cases(11, 6) + [(N, K, n) for N in (40000, LIFT_MAX_N1)
                                 for K, n in ((N // 2, N // 2), (N // 3, N // 7), (N - 5, 17), (N // 50, N // 2))]
    for N, K, n in cases:
        bad, cnt = violations(N, K, n, row_var_bound(N), scale=1)
        assert bad == 0, (N, K, n)
        lifted += cnt
    assert lifted > 10000


@pytest.mark.parametrize("N,scale", [(LIFT_MAX_N, 4), (LIFT_MAX_N1, 1)])
def test_lower_tail_lift_integer_range(N, scale):
    # scale (k N + D) stays below 2^32 for every k <= N at the largest lifted N
    V = max(lane_var_bound(N, [N // 2]), row_var_bound(N))
    D = max(lift_offset(K, N, V) for K in range(1, N + 1, 97))
    assert scale * (N * N + D) < 2 ** 32
