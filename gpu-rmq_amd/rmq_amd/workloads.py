"""Synthetic workloads of the batched and serving forms (no reference counterpart).

Arrays: float32 uniform [0, 1) from the library's counter-based generator (rmq_fill_uniform,
SplitMix64 of (seed, array, position) -> 24-bit mantissa), filled on the device.
Queries: numpy PCG64, seeded per array, inclusive (l, r) pairs whose lengths follow the
reference's distributions relative to the array's own size n_a (src/rand.cuh:86-170):
  large  L ~ U_int[1, n_a]
  medium L = trunc(lognormal(ln n_a^0.6, 0.3)), redrawn outside [1, n_a]
  small  L = trunc(lognormal(ln n_a^0.3, 0.3)), redrawn outside [1, n_a]
  mixed  one third each, drawn per query
then l ~ U_int[0, n_a - L], r = l + L - 1.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import List

import numpy as np

LENGTH_KINDS = ("large", "medium", "small", "mixed")


def query_lengths(rng: np.random.Generator, n: int, q: int, kind: str) -> np.ndarray:
    if kind == "mixed":
        cat = rng.integers(0, 3, size=q)
        out = np.empty(q, dtype=np.int64)
        for c, k in enumerate(("large", "medium", "small")):
            m = cat == c
            out[m] = query_lengths(rng, n, int(m.sum()), k)
        return out
    if kind == "large":
        return rng.integers(1, n, size=q, endpoint=True)
    expo = {"medium": 0.6, "small": 0.3}[kind]
    mu = np.log(max(1, int(n ** expo)))
    out = np.empty(q, dtype=np.int64)
    todo = np.arange(q)
    while todo.size:
        L = np.floor(rng.lognormal(mu, 0.3, size=todo.size)).astype(np.int64)
        ok = (L >= 1) & (L <= n)
        out[todo[ok]] = L[ok]
        todo = todo[~ok]
    return out


def make_queries(rng: np.random.Generator, n: int, q: int, kind: str) -> np.ndarray:
    L = query_lengths(rng, n, q, kind)
    l = (rng.random(q) * (n - L + 1)).astype(np.int64)
    l = np.minimum(l, n - L)
    out = np.empty((q, 2), dtype=np.uint32)
    out[:, 0] = l
    out[:, 1] = l + L - 1
    return out


@dataclass
class BatchWorkload:
    sizes: np.ndarray            # n_a per array
    offsets: np.ndarray          # element offset of array a in the concatenated buffer
    qcounts: np.ndarray          # q_a per array
    qoffsets: np.ndarray         # query offset of array a (len = arrays + 1)
    queries: np.ndarray          # uint32 [sum q_a, 2], local (l, r) per array
    kinds: List[str]             # length distribution per array
    seed: int

    @property
    def total_elements(self) -> int:
        return int(self.sizes.sum())

    @property
    def total_queries(self) -> int:
        return int(self.qcounts.sum())


def batch_workload(arrays: int = 4096, log_n=(10, 24), log_q=(6, 16), kind: str = "mixed",
                   seed: int = 27722, align: int = 64) -> BatchWorkload:
    """Sizes log-uniform in [2^log_n[0], 2^log_n[1]], query counts log-uniform in
    [2^log_q[0], 2^log_q[1]]; every array starts at a multiple of `align` elements."""
    rng = np.random.default_rng(seed)
    sizes = np.exp2(rng.uniform(log_n[0], log_n[1], size=arrays)).astype(np.int64)
    qcounts = np.exp2(rng.uniform(log_q[0], log_q[1], size=arrays)).astype(np.int64)
    padded = (sizes + align - 1) // align * align
    offsets = np.concatenate([[0], np.cumsum(padded)[:-1]]).astype(np.int64)
    qoffsets = np.concatenate([[0], np.cumsum(qcounts)]).astype(np.int64)
    queries = np.empty((int(qoffsets[-1]), 2), dtype=np.uint32)
    kinds = []
    for a in range(arrays):
        k = kind if kind != "per-array" else LENGTH_KINDS[a % 4]
        kinds.append(k)
        r = np.random.default_rng((seed, a))
        queries[qoffsets[a]:qoffsets[a + 1]] = make_queries(r, int(sizes[a]), int(qcounts[a]), k)
    return BatchWorkload(sizes, offsets, qcounts, qoffsets, queries, kinds, seed)


def splitmix_uniform(seed: int, start: int, count: int) -> np.ndarray:
    """Host restatement of rmq_fill_uniform for checking: SplitMix64 finaliser of
    seed + (start + i) * golden gamma, top 24 bits / 2^24."""
    with np.errstate(over="ignore"):
        i = np.arange(start, start + count, dtype=np.uint64)
        z = np.uint64(seed) + i * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(40)).astype(np.float32) * np.float32(1.0 / (1 << 24))
