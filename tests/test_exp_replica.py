"""The library's exp (the same source the kernels compile) against the host
glibc exp, bit for bit (CPU only).

The reference calls f64::exp, which lowers to glibc's exp on linux-gnu; the
device reproduces that routine.  The host build of the replica
(dto_internal_exp_host) is compared here with glibc through the oracle's
or_exp_many; the device build is compared in the GPU suite.
"""
import numpy as np


def _sweep(rng):
    parts = [
        rng.uniform(-745.2, 709.8, 600_000),            # whole finite range
        rng.uniform(-40.0, 0.0, 300_000),               # where tail terms live
        rng.uniform(-745.2, -708.0, 200_000),           # subnormal results
        rng.uniform(-1e-6, 1e-6, 50_000),               # tiny arguments
        rng.uniform(512.0, 709.8, 50_000),              # k > 0 special case
        rng.uniform(-1022.0 * np.log(2), -700.0, 50_000),
    ]
    edges = np.array([0.0, -0.0, 1.0, -1.0, 2.0 ** -60, -(2.0 ** -60), 709.782712893384, 709.7827128933841,
                      -745.1332191019411, -745.1332191019412, -708.3964185322641, -708.3964185322642,
                      512.0, -512.0, 1024.0, -1024.0, np.inf, -np.inf], dtype=np.float64)
    # neighbours of k/128 ln 2 table points, where the reduction is delicate
    k = np.arange(-200, 200, dtype=np.float64)
    grid = k * (np.log(2.0) / 128.0)
    near = np.concatenate([np.nextafter(grid, np.inf), np.nextafter(grid, -np.inf), grid])
    return np.concatenate(parts + [edges, near])


def test_host_exp_replica_bit_exact(dto, O):
    import ctypes
    rng = np.random.default_rng(1234)
    x = _sweep(rng)
    want = O.exp_many(x)
    got = np.empty_like(x)
    P = ctypes.POINTER(ctypes.c_double)
    rc = dto._lib.lib().dto_internal_exp_host(x.ctypes.data_as(P), len(x), got.ctypes.data_as(P))
    assert rc == 0
    bad = np.nonzero(got.view(np.uint64) != want.view(np.uint64))[0]
    assert bad.size == 0, [(float(x[i]).hex(), float(got[i]).hex(), float(want[i]).hex()) for i in bad[:5]]
