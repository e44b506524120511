"""MI355X (gfx950) Dual Threshold Optimization.

Mirrors the public API of the reference crate (src/lib.rs:2-10): collections,
dto, read, run and stat_operations.  The threshold-pair grid, its p-values and
the permutation null run as HIP kernels in libdto_amd.so (include/dto.h).
"""
from . import _lib
from .collections import (Feature, FeatureList, PermutedRankedFeatureList, RankedFeatureList, generate_thresholds,
                          philox_permutation)
from .dto import Job, OptimizationResult, OptimizationResultRecord, compute_population_size, optimize, \
    process_threshold_pairs
from .read import read_feature_list_from_file, read_ranked_feature_list_from_csv
from .run import Batch, Server, Task, run_single_node
from .stat_operations import empirical_pvalue, fdr, hypergeometric_pvalue, hypergeometric_pvalues, intersect_genes, \
    result_json

__version__ = "2.0.1"
