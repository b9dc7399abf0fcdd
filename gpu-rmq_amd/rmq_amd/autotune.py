"""Autotuner for the multi-array (batched / serving) query kernel.

For each (array size bucket, range-length class) it times the execution variants (one lane per
query on an R = 4, 8 or 32 hierarchy, or 4 / 8 cooperating lanes or a whole wavefront per query on
an R = 32 hierarchy) and the scan
threshold t (hierarchy depth) on synthetic arrays, and persists the fastest choice as JSON
(tuning_gfx950.json next to this file). `choose` maps an array and its query lengths to the
persisted choice. Results never depend on the choice; only speed does.

    python -m rmq_amd.autotune --write          # on the GPU; rewrites the table
"""

from __future__ import annotations

import argparse
import json
import math
import os
import time
from typing import Dict, Tuple

import numpy as np

from . import workloads

TABLE_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning_gfx950.json")
VARIANTS = ("lane4", "lane8", "lane", "group4", "group8", "wave")
THRESHOLDS = (64, 256, 1024, 4096, 32768)
TOTAL_QUERIES = 1 << 23
SIZE_LOGS = tuple(range(10, 29, 2))
KINDS = ("small", "medium", "large", "mixed")

# Used when no table has been written yet: one lane per query on an R = 8 hierarchy.
DEFAULT = {"variant": "lane8", "t": 1024}

_table_cache = None


def classify(n: int, lengths: np.ndarray) -> str:
    """Range-length class of a query batch relative to its array size n."""
    if lengths.size == 0:
        return "mixed"
    logn = math.log2(max(n, 2))
    x = np.log2(np.maximum(lengths, 1)) / logn  # exponent of the length in units of log n
    small = float(np.mean(x < 0.45))
    large = float(np.mean(x > 0.8))
    medium = 1.0 - small - large
    if small >= 0.75:
        return "small"
    if large >= 0.75:
        return "large"
    if medium >= 0.75:
        return "medium"
    return "mixed"


def _bucket(n: int) -> int:
    lg = math.log2(max(n, 2))
    return min(SIZE_LOGS, key=lambda s: abs(s - lg))


def load_table(path: str = "") -> Dict:
    """The persisted table ($RMQ_TUNING overrides the packaged file)."""
    global _table_cache
    path = path or os.environ.get("RMQ_TUNING") or TABLE_PATH
    if _table_cache is None:
        if os.path.exists(path):
            with open(path) as f:
                _table_cache = json.load(f)
        else:
            _table_cache = {"entries": {}}
    return _table_cache


def choose(n: int, lengths: np.ndarray) -> Tuple[str, int]:
    """(variant, t) for an array of n elements answering queries of these lengths."""
    e = load_table()["entries"].get(f"{_bucket(n)}:{classify(n, lengths)}", DEFAULT)
    return e["variant"], int(e["t"])


def _setup(torch, rmq_amd, size_log: int, kind: str, copies: int, per_array_q: int, seed: int):
    n = 1 << size_log
    sizes = np.full(copies, n, dtype=np.int64)
    offsets = np.arange(copies, dtype=np.int64) * ((n + 63) // 64 * 64)
    rng = np.random.default_rng(seed)
    qs = [workloads.make_queries(rng, n, per_array_q, kind) for _ in range(copies)]
    qoff = np.concatenate([[0], np.cumsum([per_array_q] * copies)]).astype(np.uint64)
    dA = torch.empty(int(offsets[-1] + n + 64), dtype=torch.float32, device="cuda")
    rmq_amd.fill_uniform(dA, seed=seed)
    dQ = torch.from_numpy(np.concatenate(qs).view(np.int32)).cuda()
    return dA, dQ, offsets, sizes, qoff


def _time_batch(torch, rmq_amd, data, variant: str, t: int) -> float:
    dA, dQ, offsets, sizes, qoff = data
    copies = len(sizes)
    with rmq_amd.Batch(dA, offsets, sizes, [t] * copies, [variant] * copies) as b:
        plan = b.plan(qoff)
        out_i = torch.empty(int(qoff[-1]), dtype=torch.int32, device="cuda")
        plan.query(dQ, False, True, None, out_i)
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(5):
            plan.query(dQ, False, True, None, out_i)
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / 5


def tune(total_queries: int = TOTAL_QUERIES, verbose: bool = True) -> Dict:
    import torch

    import rmq_amd

    entries = {}
    t0 = time.time()
    for size_log in SIZE_LOGS:
        copies = max(1, min(256, (1 << 31) // (1 << size_log) // 4))
        per_array_q = max(64, total_queries // copies)
        for kind in KINDS:
            data = _setup(torch, rmq_amd, size_log, kind, copies, per_array_q,
                          seed=size_log * 101 + len(kind))
            best = None
            timings = {}
            for t in THRESHOLDS:
                if t > (1 << size_log) and t != THRESHOLDS[0]:
                    continue
                for variant in VARIANTS:
                    ms = _time_batch(torch, rmq_amd, data, variant, t)
                    timings[f"{variant}/{t}"] = round(ms, 4)
                    if best is None or ms < best[2]:
                        best = (variant, t, ms)
            del data
            q = copies * per_array_q
            entries[f"{size_log}:{kind}"] = {
                "variant": best[0], "t": best[1], "ms": round(best[2], 4),
                "gqs": round(q / best[2] / 1e6, 3), "queries": q, "arrays": copies,
                "timings": timings,
            }
            if verbose:
                print(f"n=2^{size_log} {kind:6s}: {best[0]:7s} t={best[1]:6d} "
                      f"{q / best[2] / 1e6:7.2f} Gq/s ({time.time() - t0:.0f}s)", flush=True)
    return {
        "device": torch.cuda.get_device_name(0),
        "created": time.strftime("%Y-%m-%dT%H:%M:%S"),
        "command": "python -m rmq_amd.autotune --write",
        "method": "batched kernel, index mode, 5 timed launches after 1 warm-up, HIP events",
        "entries": entries,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--write", action="store_true", help="persist the table")
    ap.add_argument("--out", default=TABLE_PATH, help="where --write puts it")
    ap.add_argument("--queries", type=int, default=TOTAL_QUERIES)
    a = ap.parse_args()
    table = tune(a.queries)
    if a.write:
        with open(a.out, "w") as f:
            json.dump(table, f, indent=1)
        print("wrote", a.out)


if __name__ == "__main__":
    main()
