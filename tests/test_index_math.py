"""The integer and coordinate shortcuts the kernels take, checked against what they replace (CPU).

* Div16 (nntc_common.h): q = (n * ceil(2^32 / d)) >> 32 for n, d < 2^16 is the integer quotient.
* Div31 (nntc_common.h): q = (n * m) >> s, s = 31 + ceil(log2 d), m = ceil(2^s / d), for n, d < 2^31.
* tap_coord (nntc_common.h): on power-of-two planes, ldexp(p + 0.5 + d, log2 n - log2 n0) - 0.5 equals the fp32
  expression ((p + 0.5 + d) * n) / n0 - 0.5 bit for bit (src/sample.cuh:43-72 is the expression).
"""

import random

import numpy as np


def div16(n, d):
    if d <= 1:
        return n
    m = ((1 << 32) + d - 1) // d
    return (n * m) >> 32


def div31(n, d):
    s = 31 + (d - 1).bit_length() if d > 1 else 31
    m = ((1 << s) + d - 1) // d
    return (n * m) >> s


def test_div16_matches_integer_division():
    rng = random.Random(16)
    for d in list(range(1, 1100)) + [rng.randrange(1100, 1 << 16) for _ in range(200)] + [(1 << 16) - 1]:
        ns = list(range(0, 1 << 16, 97)) + [(1 << 16) - 1, d - 1, d, d + 1]
        for n in ns:
            if n < (1 << 16):
                assert div16(n, d) == n // d, (n, d)


def test_div31_matches_integer_division():
    rng = random.Random(31)
    ds = [1, 2, 3, 5, 7, 1023, 1024, 1025, (1 << 30) - 1, 1 << 30, (1 << 31) - 1]
    ds += [rng.randrange(1, 1 << 31) for _ in range(300)]
    for d in ds:
        ns = [0, 1, d - 1, d, d + 1, 2 * d - 1, 2 * d, (1 << 31) - 1] + [rng.randrange(0, 1 << 31) for _ in range(300)]
        for n in ns:
            if 0 <= n < (1 << 31):
                assert div31(n, d) == n // d, (n, d)


def test_tap_coord_power_of_two_path_is_the_division():
    f32 = np.float32
    offsets = [0.0, -0.375, -0.125, 0.125, 0.375, 0.25, -0.25]
    for log_n0 in range(0, 13):
        n0 = 1 << log_n0
        for log_n in range(0, log_n0 + 1):
            n = 1 << log_n
            p = np.arange(-2, n0 + 2, dtype=np.int64)
            for d in offsets:
                a = (p.astype(f32) + f32(0.5)) + f32(d)
                slow = (a * f32(n)) / f32(n0) - f32(0.5)
                fast = np.ldexp(a, log_n - log_n0).astype(f32) - f32(0.5)
                assert slow.dtype == np.float32 and fast.dtype == np.float32
                np.testing.assert_array_equal(slow.view(np.uint32), fast.view(np.uint32),
                                              err_msg=f"n0 {n0} n {n} d {d}")
