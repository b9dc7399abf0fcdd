"""Writes tests/golden/nan_vectors.json: arrays holding NaN, with queries and their answers under
CPU BASE's rule (src/cpu_methods.h:7-16, 42-53: the scan starts from A[l] and a later element
replaces the minimum only if strictly smaller), stated directly rather than by scanning:

  if A[l] is NaN:  index l
  else:            the leftmost position of the smallest non-NaN element of [l, r]

(a NaN compares false with everything, so it never replaces the running minimum, and A[l] is
never replaced when it is a NaN). Every NaN carries its own payload, so the value bits name the
position. Run: python tests/golden/make_nan_vectors.py
"""

import json
import os

import numpy as np

N = 2053


def nan_with_payload(i):
    return np.uint32(0x7FC00000 | (i & 0x3FFFFF)).view(np.float32)


def answer(A, l, r):
    if np.isnan(A[l]):
        return l
    seg = A[l:r + 1]
    ok = ~np.isnan(seg)
    m = seg[ok].min()
    return l + int(np.nonzero(ok & (seg == m))[0][0])


def case(name, A, rng):
    n = A.size
    nan_pos = np.nonzero(np.isnan(A))[0]
    pairs = {(0, n - 1), (0, 0), (n - 1, n - 1)}
    for p in nan_pos[:: max(1, nan_pos.size // 40)].tolist():
        for L in (1, 2, 5, 33, 300, n):
            pairs.add((p, min(n - 1, p + L - 1)))  # starting at a NaN
            pairs.add((max(0, p - L + 1), p))      # ending at one
            if p > 0:
                pairs.add((p - 1, min(n - 1, p + L)))  # holding one after a real first element
    for l, r in np.sort(rng.integers(0, n, size=(150, 2)), axis=1).tolist():
        pairs.add((l, r))
    lr = sorted(pairs)
    idx = [answer(A, l, r) for l, r in lr]
    return {"name": name, "array_bits": A.view(np.uint32).tolist(), "lr": [list(p) for p in lr],
            "index": idx, "value_bits": [int(A.view(np.uint32)[i]) for i in idx]}


def main():
    rng = np.random.default_rng(2053)
    cases = []
    A = rng.random(N, dtype=np.float32)
    for p in rng.choice(N, 60, replace=False):
        A[p] = nan_with_payload(p)
    cases.append(case("scattered_nan", A, rng))
    A = rng.integers(0, 4, size=N).astype(np.float32)  # ties everywhere
    for start, length in ((0, 4), (8, 8), (64, 32), (256, 128), (1024, 1024)):
        for p in range(start, min(N, start + length)):
            A[p] = nan_with_payload(p)
    cases.append(case("aligned_nan_blocks", A, rng))
    A = np.full(N, np.inf, dtype=np.float32)
    for p in range(N):
        if (p // 7) % 3 != 1:
            A[p] = nan_with_payload(p)
    cases.append(case("nan_and_inf", A, rng))
    A = np.array([nan_with_payload(p) for p in range(N)], dtype=np.float32)
    cases.append(case("all_nan", A, rng))
    A = np.where(rng.random(N) < 0.5, np.float32(0.0), np.float32(-0.0)).astype(np.float32)
    for p in rng.choice(N, 200, replace=False):
        A[p] = nan_with_payload(p)
    cases.append(case("nan_and_signed_zeros", A, rng))
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "nan_vectors.json")
    with open(out, "w") as f:
        json.dump(cases, f)
    print("wrote", out, [len(c["lr"]) for c in cases])


if __name__ == "__main__":
    main()
