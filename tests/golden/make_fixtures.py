"""Writes the committed golden fixtures under tests/golden/. Run: python tests/golden/make_fixtures.py

The reference ships no fixtures, golden vectors or tests (its only correctness path is the
runtime --check flag, src/main.cu:194-283), so every expected value here is produced
independently of the code under test:
  * generator_fixtures.json — seeded inputs from tests/golden/pygen.py (published algorithms,
    not libstdc++), as a prefix plus the SHA-256 of the full output;
  * tie_vectors.json — hand-built arrays with ties, +0/-0, +inf and denormals, with answers from
    numpy (min over the slice; argmin returns the first occurrence = the leftmost rule of
    src/cpu_methods.h:42-53);
  * metadata_kats.json — level sizes/offsets of the recurrence src/xxx.h:225-243 written out as
    literal numbers (derived by hand from ceil-division and 32-alignment).
"""

from __future__ import annotations

import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import pygen  # noqa: E402

GENERATOR_CASES = [
    # (q, lr, n, seed, trivial_check)
    (4096, -1, 1 << 20, 27722, False),
    (4096, -2, 1 << 20, 27722, False),
    (4096, -3, 1 << 20, 27722, False),
    (4096, -6, 1 << 20, 27722, False),
    (2048, -2, 1 << 26, 27722, False),
    (2048, -1, 1 << 26, 833241, False),
    (2048, -3, 1 << 30, 913, False),
    (2048, -6, 1 << 30, 913, False),
    (1024, -4, 1 << 22, 97528, False),
    (1024, -5, 1 << 24, 97528, False),
    (1024, 33, 5000, 1, False),
    (1024, -6, 997, 26, False),
    (3000, -1, 1 << 16, 4, True),
]
ARRAY_CASES = [(1 << 16, 27722), (1 << 16, 833241), (1000, 0), (4099, 26)]


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def generator_fixtures():
    qs = []
    for q, lr, n, seed, triv in GENERATOR_CASES:
        out = pygen.gen_queries(q, lr, n, seed, triv)
        qs.append(dict(q=q, lr=lr, n=n, seed=seed, trivial_check=triv,
                       head=out[:16].tolist(), sha256=sha(out)))
    arrs = []
    for n, seed in ARRAY_CASES:
        a = pygen.gen_array(n, seed)
        arrs.append(dict(n=n, seed=seed, head_bits=a[:16].view(np.uint32).tolist(),
                         sha256=sha(a)))
    samples = []
    for orig, k, seed in [(1 << 20, 512, 27722), (300, 512, 5)]:
        s = pygen.sample_indices(orig, k, seed)
        samples.append(dict(original_size=orig, sample_size=k, seed=seed, sha256=sha(s),
                            head=s[:16].tolist()))
    return dict(queries=qs, arrays=arrs, samples=samples)


def _answers(A: np.ndarray, lr: np.ndarray):
    idx = np.array([l + int(np.argmin(A[l:r + 1])) for l, r in lr], dtype=np.uint32)
    val_bits = A[idx].view(np.uint32)
    return idx.tolist(), val_bits.tolist()


def tie_vectors():
    rng = np.random.default_rng(20240613)
    cases = []
    f32 = np.float32
    # 1. all equal
    A = np.full(300, 0.5, dtype=f32)
    cases.append(("all_equal", A))
    # 2. repeated minima at block boundaries (R = 32 and 16)
    A = rng.random(1000, dtype=f32) + f32(1.0)
    for p in (31, 32, 63, 64, 95, 128, 511, 512, 513, 999):
        A[p] = f32(0.25)
    cases.append(("boundary_ties", A))
    # 3. +0 and -0 mixed (the value must be the leftmost zero's bit pattern)
    A = rng.random(777, dtype=f32) + f32(1.0)
    A[[5, 40, 41, 300, 301, 640]] = f32(0.0)
    A[[6, 39, 299, 302, 641, 700]] = f32(-0.0)
    cases.append(("signed_zeros", A))
    # 4. +inf everywhere with a few finite values
    A = np.full(513, np.inf, dtype=f32)
    A[[100, 400]] = f32(3.0)
    cases.append(("mostly_inf", A))
    # 5. denormals and negative values
    A = (rng.random(640, dtype=f32) - f32(0.5)) * f32(1e-38)
    A[[17, 18, 321]] = A.min()
    cases.append(("denormal_negative", A))
    # 6. few distinct values (many ties), n not a multiple of 4 or 32
    A = rng.integers(0, 3, size=1027).astype(f32)
    cases.append(("three_values", A))
    out = []
    for name, A in cases:
        n = A.size
        pairs = [(0, n - 1), (0, 0), (n - 1, n - 1)]
        for _ in range(120):
            l, r = sorted(int(x) for x in rng.integers(0, n, size=2))
            pairs.append((l, r))
        for L in (1, 2, 3, 4, 5, 31, 32, 33, 63, 64, 65, 127, 128, 129):
            if L <= n:
                l = int(rng.integers(0, n - L + 1))
                pairs.append((l, l + L - 1))
        lr = np.array(pairs, dtype=np.uint32)
        idx, vbits = _answers(A, lr)
        out.append(dict(name=name, array_bits=A.view(np.uint32).tolist(), lr=lr.tolist(),
                        index=idx, value_bits=vbits))
    return out


# Level sizes and offsets of src/xxx.h:225-243: size_{k+1} = ceil(size_k / R) until
# size <= t or 20 levels; offset_k = align32(offset_{k-1} + size_{k-1}), level 0 offset ~0u.
METADATA_KATS = [
    dict(n=1 << 20, R_log=5, t=1 << 15, sizes=[1 << 20, 32768], offsets=[0xFFFFFFFF, 0], floats=32768),
    dict(n=1 << 26, R_log=5, t=1 << 15, sizes=[1 << 26, 2097152, 65536, 2048],
         offsets=[0xFFFFFFFF, 0, 2097152, 2162688], floats=2164736),
    dict(n=1 << 26, R_log=5, t=1 << 10, sizes=[1 << 26, 2097152, 65536, 2048, 64],
         offsets=[0xFFFFFFFF, 0, 2097152, 2162688, 2164736], floats=2164800),
    dict(n=1 << 26, R_log=5, t=64, sizes=[1 << 26, 2097152, 65536, 2048, 64],
         offsets=[0xFFFFFFFF, 0, 2097152, 2162688, 2164736], floats=2164800),
    dict(n=1 << 26, R_log=5, t=1 << 16, sizes=[1 << 26, 2097152, 65536],
         offsets=[0xFFFFFFFF, 0, 2097152], floats=2162688),
    dict(n=1 << 30, R_log=5, t=64, sizes=[1 << 30, 1 << 25, 1 << 20, 1 << 15, 1024, 32],
         offsets=[0xFFFFFFFF, 0, 33554432, 34603008, 34635776, 34636800], floats=34636832),
    dict(n=1 << 30, R_log=5, t=1 << 15, sizes=[1 << 30, 1 << 25, 1 << 20, 1 << 15],
         offsets=[0xFFFFFFFF, 0, 33554432, 34603008], floats=34635776),
    dict(n=100000, R_log=5, t=64, sizes=[100000, 3125, 98, 4], offsets=[0xFFFFFFFF, 0, 3136, 3264],
         floats=3268),
    dict(n=1000, R_log=5, t=64, sizes=[1000, 32], offsets=[0xFFFFFFFF, 0], floats=32),
    dict(n=1 << 26, R_log=4, t=1 << 15, sizes=[1 << 26, 1 << 22, 1 << 18, 1 << 14],
         offsets=[0xFFFFFFFF, 0, 4194304, 4456448], floats=4472832),
    dict(n=1 << 26, R_log=3, t=32,
         sizes=[1 << 26, 1 << 23, 1 << 20, 1 << 17, 1 << 14, 2048, 256, 32],
         offsets=[0xFFFFFFFF, 0, 8388608, 9437184, 9568256, 9584640, 9586688, 9586944],
         floats=9586976),
    # n <= t: one level, no level buffer
    dict(n=5000, R_log=5, t=1 << 15, sizes=[5000], offsets=[0xFFFFFFFF], floats=0),
    # 20-level cap (R = 8 cannot shrink below t = 1 before 20 levels only if n is huge; use R=8,
    # n = 2^31-1, t = 1: sizes shrink by 8 until 1 after 11 levels -> cap not hit)
    dict(n=2147483647, R_log=3, t=1,
         sizes=[2147483647, 268435456, 33554432, 4194304, 524288, 65536, 8192, 1024, 128, 16, 2, 1],
         offsets=[0xFFFFFFFF, 0, 268435456, 301989888, 306184192, 306708480, 306774016,
                  306782208, 306783232, 306783360, 306783392, 306783424],
         floats=306783425),
]


def main():
    with open(os.path.join(HERE, "generator_fixtures.json"), "w") as f:
        json.dump(generator_fixtures(), f, indent=1)
    with open(os.path.join(HERE, "tie_vectors.json"), "w") as f:
        json.dump(tie_vectors(), f)
    with open(os.path.join(HERE, "metadata_kats.json"), "w") as f:
        json.dump(METADATA_KATS, f, indent=1)
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()
