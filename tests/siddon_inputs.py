"""The segment set of the reference's test/SiddonTest.cpp:8-34, restated.

That test draws 10000 segments from std::default_random_engine seeded with 0
and std::uniform_real_distribution<double>(-500, 500), six draws per segment
in the order p1.x, p1.y, p1.z, p2.x, p2.y, p2.z (SiddonTest.cpp:27-32).  With
libstdc++ the engine is minstd_rand0 (x <- 16807 x mod 2^31-1; seed 0 maps to
state 1) and the distribution is a + (b - a) * generate_canonical<double, 53>,
which combines k = 2 draws: s = (u1 - 1) + (u2 - 1) * R, R = 2^31 - 2, divided by
R * R and clamped below 1.  This module restates both in numpy FP64, so the
inputs exist on machines without a C++ compiler; tests/golden/siddon_test.json
holds the SHA-256 of the bytes that tests/golden/siddon_test_inputs.cpp (the
real libstdc++ calls) writes, which pins this restatement.
"""
import numpy as np

_M = 2147483647
_A = 16807
_R = float(_M - 1)  # max - min + 1 of minstd_rand0 (min 1, max 2^31 - 2)


def minstd_rand0(n, seed=0):
    """n successive outputs of std::minstd_rand0(seed)."""
    x = seed % _M
    if x == 0:
        x = 1
    out = np.empty(n, np.int64)
    for i in range(n):
        x = (x * _A) % _M
        out[i] = x
    return out


def uniform_doubles(n, lo=-500.0, hi=500.0, seed=0):
    u = minstd_rand0(2 * n, seed).astype(np.float64)
    s = np.float64(0.0) + (u[0::2] - 1.0) * np.float64(1.0)
    s = s + (u[1::2] - 1.0) * np.float64(_R)
    c = s / (np.float64(_R) * np.float64(_R))
    c = np.where(c >= 1.0, np.nextafter(1.0, 0.0), c)
    return c * (hi - lo) + lo


def siddon_test_segments(n=10000):
    """(p1, p2), each (n, 3) float64, exactly as SiddonTest.cpp draws them."""
    v = uniform_doubles(6 * n).reshape(n, 6)
    return v[:, :3].copy(), v[:, 3:].copy()
