"""dto_process_threshold_pairs_noalloc (the allocation-free debug grid)
against dto_process_threshold_pairs on the same job and permutation, bit for
bit, on two calls in a row (the second on a smaller grid, so that stale
counts of the first call would show), plus its size limit and status word.
Inputs and outputs are device buffers the caller owns (torch tensors)."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _noalloc(dto, torch, job, idx1, idx2, stream=None):
    L = dto._lib.lib()
    dev = torch.device("cuda:0")
    cells = job.n_thresholds1 * job.n_thresholds2
    d1 = None if idx1 is None else torch.from_numpy(np.ascontiguousarray(idx1, dtype=np.int32)).to(dev)
    d2 = None if idx2 is None else torch.from_numpy(np.ascontiguousarray(idx2, dtype=np.int32)).to(dev)
    out = torch.full((cells * dto._lib.RECORD_DTYPE.itemsize,), 0xAB, dtype=torch.uint8, device=dev)
    status = torch.full((1,), 7, dtype=torch.int32, device=dev)
    torch.cuda.synchronize()
    rc = L.dto_process_threshold_pairs_noalloc(job.handle, None if d1 is None else d1.data_ptr(),
                                               None if d2 is None else d2.data_ptr(), out.data_ptr(),
                                               status.data_ptr(), stream)
    dto._lib.check(rc)
    torch.cuda.synchronize()
    rec = np.frombuffer(out.cpu().numpy().tobytes(), dtype=dto._lib.RECORD_DTYPE)
    return rec, int(status.cpu().item())


def test_noalloc_grid_equals_allocating_grid(dto, gpu):
    import torch
    from dual_threshold_optimization_amd import workloads as W
    torch.cuda.set_device(0)
    rng = np.random.default_rng(71)
    # first call: a 1,500-feature job with ties and a planted overlap (permuted)
    a1, ar1, a2, ar2, _ = W.planted_lists(1500, 60, 200, 150, 0.05, seed=72)
    big = dto.Job.from_lists(a1, ar1, a2, ar2)
    p1, p2 = rng.permutation(big.len1), rng.permutation(big.len2)
    want = big.process_threshold_pairs(p1, p2)
    got, st = _noalloc(dto, torch, big, p1, p2)
    assert st == 0
    assert got.tobytes() == want.tobytes()
    # second call, right after: a smaller grid (unpermuted)
    b1, br1, b2, br2, _ = W.planted_lists(300, 12, 40, 30, 0.1, seed=73)
    small = dto.Job.from_lists(b1, br1, b2, br2)
    assert small.n_thresholds1 * small.n_thresholds2 < big.n_thresholds1 * big.n_thresholds2
    want = small.process_threshold_pairs()
    got, st = _noalloc(dto, torch, small, None, None)
    assert st == 0
    assert got.tobytes() == want.tobytes()
    # and the first job once more, now on a user stream
    s = torch.cuda.Stream()
    q1 = rng.permutation(big.len1)
    want = big.process_threshold_pairs(q1, None)
    with torch.cuda.stream(s):
        got, st = _noalloc(dto, torch, big, q1, None, stream=s.cuda_stream)
    assert st == 0 and got.tobytes() == want.tobytes()


def test_noalloc_limits_and_status(dto, gpu):
    import torch
    from dual_threshold_optimization_amd import workloads as W
    torch.cuda.set_device(0)
    a1, ar1, a2, ar2, _ = W.planted_lists(2200, 50, 200, 150, 0.05, seed=74)
    job = dto.Job.from_lists(a1, ar1, a2, ar2)
    # a grid above the limit is refused before anything runs
    t = np.arange(1, 2101, dtype=np.uint32)
    job.set_thresholds(t, t)
    assert 2100 * 2100 > dto._lib.DTO_NOALLOC_MAX_CELLS
    L = dto._lib.lib()
    rc = L.dto_process_threshold_pairs_noalloc(job.handle, None, None, ctypes.c_void_p(1), None, None)
    assert rc == dto._lib.DTO_ERR_TOO_LARGE and b"DTO_NOALLOC_MAX_CELLS" in L.dto_last_error()
    # an out-of-range index sets the status word (the grid is then meaningless)
    job.set_thresholds(t[:40], t[:50])
    bad = np.arange(job.len1, dtype=np.int64)
    bad[5] = job.len1 + 3
    _, st = _noalloc(dto, torch, job, bad, None)
    assert st != 0
    # and a good call right after clears it again
    got, st = _noalloc(dto, torch, job, None, None)
    assert st == 0 and got.tobytes() == job.process_threshold_pairs().tobytes()
