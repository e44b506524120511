"""Feature containers of the reference crate (src/collections/), host side.

These hold ids and ranks on the host; everything computed per threshold pair
runs on the GPU through ``Job`` (dual_threshold_optimization_amd/dto.py).
"""
import os

import numpy as np

from . import _lib


class Feature:
    """A feature id (src/collections/feature.rs:25-28)."""
    __slots__ = ("_id",)

    def __init__(self, id):
        self._id = str(id)

    @classmethod
    def from_(cls, id):
        return cls(id)

    def id(self):
        return self._id

    def __eq__(self, other):
        return isinstance(other, Feature) and other._id == self._id

    def __hash__(self):
        return hash(self._id)

    def __repr__(self):
        return "Feature(%r)" % self._id


def _ids(genes):
    if isinstance(genes, FeatureList):
        return [g.id() for g in genes.genes()]
    return [g.id() if isinstance(g, Feature) else str(g) for g in genes]


class FeatureList:
    """Ordered list of features (src/collections/feature_list.rs)."""

    def __init__(self, genes=()):
        self._genes = [g if isinstance(g, Feature) else Feature(g) for g in genes]

    @classmethod
    def from_(cls, genes):
        return cls(genes)

    def genes(self):
        return list(self._genes)

    def ids(self):
        return [g.id() for g in self._genes]

    def __len__(self):
        return len(self._genes)

    def __getitem__(self, i):
        return self._genes[i]

    def __iter__(self):
        return iter(self._genes)

    def is_empty(self):
        return not self._genes

    def intersect(self, other):
        """Features of self whose id is in both lists (feature_list.rs:278-286)."""
        a, b = set(self.ids()), set(_ids(other))
        return [g for g in self._genes if g.id() in a and g.id() in b]

    def difference(self, other, symmetric=False):
        """Set or symmetric difference by id (feature_list.rs:318-344)."""
        a, b = set(self.ids()), set(_ids(other))
        if symmetric:
            other_genes = other.genes() if isinstance(other, FeatureList) else [Feature(x) for x in _ids(other)]
            sd = a ^ b
            return [g for g in self._genes + other_genes if g.id() in sd]
        return [g for g in self._genes if g.id() in a - b]


class RankedFeatureList:
    """Features with integer ranks, stable-sorted by rank, with thresholds
    (src/collections/ranked.rs:176-191, 359-375, 527-542)."""

    def __init__(self):
        self._genes = FeatureList()
        self._ranks = np.zeros(0, dtype=np.uint32)
        self._thresholds = np.zeros(0, dtype=np.uint32)

    @classmethod
    def from_(cls, genes, ranks):
        genes = genes if isinstance(genes, FeatureList) else FeatureList(genes)
        ranks = [int(r) & 0xFFFFFFFF for r in ranks]
        if len(genes) != len(ranks):
            raise ValueError("Feature list length (%d) does not match ranks length (%d)" % (len(genes), len(ranks)))
        self = cls()
        order = sorted(range(len(ranks)), key=lambda i: ranks[i])  # stable
        self._genes = FeatureList([genes[i] for i in order])
        self._ranks = np.array([ranks[i] for i in order], dtype=np.uint32)
        self._thresholds = generate_thresholds(int(self._ranks[-1]) if len(self._ranks) else 0)
        return self

    def genes(self):
        return self._genes

    def ranks(self):
        return self._ranks

    def thresholds(self):
        return self._thresholds

    def ids(self):
        return self._genes.ids()

    def __len__(self):
        return len(self._genes)

    def len(self):
        return len(self._genes)

    def is_empty(self):
        return len(self._genes) == 0

    def get_feature_set_by_threshold(self, threshold):
        """Features with rank <= threshold (ranked.rs:561-568)."""
        return [g for g, r in zip(self._genes, self._ranks) if r <= threshold]


class PermutedRankedFeatureList:
    """Permuted view of a ranked list (src/collections/permuted.rs:56-101):
    position k holds the feature at sorted position indices[k]; ranks stay.

    The reference draws its shuffle from thread_rng; here ``indices`` may be
    given explicitly, or is the device's Philox shuffle for (seed, job, perm,
    list) -- seed defaults to OS entropy."""

    def __init__(self, original, indices=None, seed=None, job=0, perm=0, list_id=1):
        self.original = original
        n = len(original)
        if indices is None:
            if seed is None:
                seed = int.from_bytes(os.urandom(8), "little")
            indices = philox_permutation(seed, job, perm, list_id, n)
        self.indices = np.ascontiguousarray(indices, dtype=np.uint32)
        if self.indices.shape != (n,):
            raise ValueError("indices must have length %d" % n)

    @classmethod
    def new(cls, original, **kw):
        return cls(original, **kw)

    def thresholds(self):
        return self.original.thresholds()

    def ranks(self):
        return self.original.ranks()

    def get_feature_set_by_threshold(self, threshold):
        g = self.original.genes()
        return [g[int(i)] for r, i in zip(self.original.ranks(), self.indices) if r <= threshold]


def generate_thresholds(max_rank):
    """generate_thresholds (src/collections/ranked.rs:359-375), from the library."""
    L = _lib.lib()
    n = L.dto_generate_thresholds(int(max_rank), None, 0)
    out = np.zeros(max(n, 1), dtype=np.uint32)
    L.dto_generate_thresholds(int(max_rank), _lib.u32p(out), n)
    return out[:n]


def philox_permutation(seed, job, perm, list_id, n):
    """The Philox permutation the device draws for (seed, job, perm, list)."""
    out = np.zeros(max(n, 1), dtype=np.uint32)
    _lib.check(_lib.lib().dto_philox_permutation(int(seed) & (2**64 - 1), int(job), int(perm), int(list_id), int(n),
                                                  _lib.u32p(out)))
    return out[:n]
