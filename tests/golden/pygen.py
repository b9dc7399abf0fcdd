"""Independent pure-Python restatement of the reference's seeded input generators.

Test infrastructure: it pins the product's C++ generator (gpu-rmq_amd/csrc/rmq_gen.cpp), which
drives libstdc++ <random>, by re-deriving every draw from published algorithms instead of
calling libstdc++:
  * MT19937 with init_genrand seeding -> numpy's MT19937 after ``_legacy_seeding`` (checked
    against the C++ standard's known answer: the 10000th output of a default-seeded
    std::mt19937 is 4123659995, [rand.predef]);
  * generate_canonical<float, 24> / <double, 53> over a 32-bit engine: one / two draws,
    sum * 2^-32 / 2^-64, clamped below 1;
  * uniform_int_distribution over a 32-bit engine: Lemire's nearly-divisionless method
    (ACM TOMACS 29(1), 2019), which libstdc++ 11 uses for this case;
  * normal_distribution: Marsaglia's polar method with one cached value; lognormal = exp(s*z+m).
The call order follows src/rand.cuh:49-240 (see the C++ file for line-by-line citations).
"""

from __future__ import annotations

import math

import numpy as np

TWO32 = 4294967296.0
TWO64 = 18446744073709551616.0
MASK32 = 0xFFFFFFFF


class MT:
    """std::mt19937(seed) as a Python callable returning 32-bit outputs."""

    def __init__(self, seed: int):
        self.bg = np.random.MT19937()
        self.bg._legacy_seeding(seed & MASK32)
        self.buf = []
        self.pos = 0

    def __call__(self) -> int:
        if self.pos >= len(self.buf):
            self.buf = self.bg.random_raw(4096).tolist()
            self.pos = 0
        v = self.buf[self.pos]
        self.pos += 1
        return v


def canonical_float(g: MT) -> np.float32:
    f = np.float32(g()) / np.float32(TWO32)
    if f >= np.float32(1.0):
        f = np.nextafter(np.float32(1.0), np.float32(0.0))
    return f


def canonical_double(g: MT) -> float:
    s = float(g())
    s = s + float(g()) * TWO32
    r = s / TWO64
    if r >= 1.0:
        r = math.nextafter(1.0, 0.0)
    return r


def uniform_int(g: MT, a: int, b: int) -> int:
    urange = b - a
    if urange == MASK32:
        return g() + a
    if urange > MASK32:
        # ranges wider than the generator's 32 bits (64-bit positions above 2^32): a draw of
        # the high part over [0, urange / 2^32], times 2^32, plus one raw output; redrawn while
        # the sum leaves the range or wraps 64 bits
        while True:
            tmp = ((1 << 32) * uniform_int(g, 0, urange >> 32)) & ((1 << 64) - 1)
            ret = (tmp + g()) & ((1 << 64) - 1)
            if not (ret > urange or ret < tmp):
                return ret + a
    erange = urange + 1
    product = g() * erange
    low = product & MASK32
    if low < erange:
        threshold = ((1 << 32) - erange) % erange
        while low < threshold:
            product = g() * erange
            low = product & MASK32
    return (product >> 32) + a


class Normal:
    """normal_distribution<double>(0, 1) with the cached second value of the polar method."""

    def __init__(self):
        self.saved = None

    def __call__(self, g: MT) -> float:
        if self.saved is not None:
            r, self.saved = self.saved, None
            return r
        while True:
            x = 2.0 * canonical_double(g) - 1.0
            y = 2.0 * canonical_double(g) - 1.0
            r2 = x * x + y * y
            if not (r2 > 1.0 or r2 == 0.0):
                break
        mult = math.sqrt(-2.0 * math.log(r2) / r2)
        self.saved = x * mult
        return y * mult


class LogNormal:
    def __init__(self, m: float, s: float):
        self.m, self.s, self.nd = m, s, Normal()

    def __call__(self, g: MT) -> float:
        return math.exp(self.s * self.nd(g) + self.m)


def gen_array(n: int, seed: int) -> np.ndarray:
    g = MT(seed)
    return np.array([canonical_float(g) for _ in range(n)], dtype=np.float32)


def _place(g: MT, n: int, length: int):
    l = uniform_int(g, 0, n - 1 - (length - 1))
    return l, l + length - 1


def _lognormal_len(g: MT, d: LogNormal, n: int) -> int:
    while True:
        v = int(d(g))  # truncation, as the (nType) cast
        if 0 < v <= n:
            return v


def gen_queries(q: int, lr: int, n: int, seed: int, trivial_check: bool = False,
                dtype=np.uint32) -> np.ndarray:
    """dtype uint64: the reference's IS_LONG build (64-bit positions; the same draws while
    n <= 2^32)."""
    qgen = 1024 if (trivial_check and q > 1024) else q
    out = np.empty((qgen, 2), dtype=dtype)
    if lr == -6:
        g = MT(seed)
        med = LogNormal(math.log(int(math.pow(float(n), 0.6))), 0.3)
        small = LogNormal(math.log(int(math.pow(float(n), 0.3))), 0.3)
        for i in range(qgen):
            c = uniform_int(g, 0, 2)
            if c == 0:
                length = uniform_int(g, 1, n)
            elif c == 1:
                length = _lognormal_len(g, med, n)
            else:
                length = _lognormal_len(g, small, n)
            out[i] = _place(g, n, length)
    else:
        g = MT(0)  # seed * tid with tid = 0
        if lr > 0:
            for i in range(qgen):
                out[i] = _place(g, n, lr)
        elif lr == -1:
            for i in range(qgen):
                length = uniform_int(g, 1, n)
                out[i] = _place(g, n, length)
        else:
            scale = {
                -2: int(math.pow(float(n), 0.6)),
                -3: int(math.pow(float(n), 0.3)),
                -4: max(1, n // (1 << 8)),
                -5: max(1, n // (1 << 15)),
            }[lr]
            d = LogNormal(math.log(scale), 0.3)
            for i in range(qgen):
                out[i] = _place(g, n, _lognormal_len(g, d, n))
    if qgen != q:
        out = out[np.arange(q) & 1023]
    return out


def sample_indices(original_size: int, sample_size: int, seed: int) -> np.ndarray:
    if original_size <= sample_size:
        return np.arange(original_size, dtype=np.uint32)
    g = MT(seed)
    return np.array([uniform_int(g, 0, original_size - 1) for _ in range(sample_size)],
                    dtype=np.uint32)
