"""python -m dual_threshold_optimization_amd: the reference CLI (src/main.rs:20-166)
over the library; same flags, stderr lines and pretty JSON on stdout."""
import argparse
import os
import sys

from . import _lib
from .dto import Job, tie_warnings
from .read import read_ranked_feature_list_from_csv
from .stat_operations import result_json


def main(argv=None):
    p = argparse.ArgumentParser(prog="dual_threshold_optimization", description="Dual Threshold Optimization CLI")
    p.add_argument("-1", "--ranked-list1", required=True, metavar="FILE")
    p.add_argument("-2", "--ranked-list2", required=True, metavar="FILE")
    p.add_argument("-b", "--background", metavar="FILE")
    p.add_argument("-p", "--permutations", type=int, default=1000)
    p.add_argument("-t", "--threads", type=int, default=1)
    p.add_argument("-m", "--multi-node", action="store_true")
    p.add_argument("--seed", type=int)
    p.add_argument("--mode", choices=["screened", "plain"], default="screened")
    p.add_argument("--drop-outside-background", action="store_true")
    a = p.parse_args(argv)
    threads = a.threads
    if threads == 0:
        print("Warning: Number of threads cannot be 0. Setting threads to 1.", file=sys.stderr)
        threads = 1
    err = sys.stderr
    print("Ranked list 1: %s" % a.ranked_list1, file=err)
    print("Ranked list 2: %s" % a.ranked_list2, file=err)
    print("Permutations: %d" % a.permutations, file=err)
    print("Threads: %d" % threads, file=err)
    print("Multi-node mode: %s" % ("enabled" if a.multi_node else "disabled"), file=err)
    product = ("The product of the lengths of the threshold lists (this describes the asymptotic runtime of a "
               "single job): %d")
    try:
        try:
            job = Job.from_files(a.ranked_list1, a.ranked_list2, a.background, a.drop_outside_background)
        except _lib.DtoError as e:
            # main.rs prints the product after reading both lists, before the
            # background and population checks
            if e.code not in (_lib.DTO_ERR_IO, _lib.DTO_ERR_PARSE):
                l1 = read_ranked_feature_list_from_csv(a.ranked_list1)
                l2 = read_ranked_feature_list_from_csv(a.ranked_list2)
                print(product % (len(l1.thresholds()) * len(l2.thresholds())), file=err)
            raise
        print(product % (job.n_thresholds1 * job.n_thresholds2), file=err)
        if a.multi_node:
            raise _lib.DtoError(1, "Multi-node mode requires MPI. Rebuild with the 'mpi' feature enabled.")
        seed = a.seed if a.seed is not None else int.from_bytes(os.urandom(8), "little")
        res, perm = job.run(a.permutations, seed=seed, mode=a.mode)
        # each task's optimize warns about its own ties (optimize_main.rs:86-108)
        for line in tie_warnings(res.unpermuted):
            print(line, file=err)
        for r in perm:
            for line in tie_warnings(r):
                print(line, file=err)
    except _lib.DtoError as e:
        print("thread 'main' panicked:\n%s" % e.message, file=err)
        return 101
    print(result_json(res))
    return 0


if __name__ == "__main__":
    sys.exit(main())
