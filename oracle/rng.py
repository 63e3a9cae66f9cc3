"""Counter-based random streams for the synthetic workloads (oracle side).

SplitMix64 (Steele, Lea, Flood, "Fast splittable pseudorandom number
generators", OOPSLA 2014): state_k = seed + (k+1) * 0x9E3779B97F4A7C15, then the
published finaliser.  Uniform doubles take the top 53 bits.  Normal deviates
use Box-Muller on two consecutive uniforms.  The product carries its own
implementation; tests check that both agree bit for bit.
"""

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed, counters):
    counters = np.asarray(counters, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + (counters + np.uint64(1)) * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed, counters):
    """U[0,1) doubles from the top 53 bits."""
    return (splitmix64(seed, counters) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def normal(seed, n):
    """n standard normal deviates, Box-Muller on uniforms 2i and 2i+1."""
    i = np.arange(n, dtype=np.uint64)
    u1 = uniform(seed, 2 * i)
    u2 = uniform(seed, 2 * i + 1)
    return np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * np.pi * u2)
