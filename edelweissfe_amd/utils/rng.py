"""Seeded counter-based random streams for synthetic workloads.

SplitMix64 (Steele, Lea & Flood, OOPSLA 2014): the k-th value of stream `seed`
is the SplitMix64 finaliser applied to seed + (k + 1) * 0x9E3779B97F4A7C15.
Uniform doubles take the top 53 bits; normal deviates use Box-Muller on the
uniforms 2i and 2i + 1.
"""

import numpy as np

_INC = np.uint64(0x9E3779B97F4A7C15)


def _mix(z):
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def bits(seed, counter):
    with np.errstate(over="ignore"):
        return _mix(np.uint64(seed) + (np.asarray(counter, dtype=np.uint64) + np.uint64(1)) * _INC)


def uniform(seed, counter):
    return (bits(seed, counter) >> np.uint64(11)).astype(np.float64) / 9007199254740992.0


def normal(seed, n):
    k = np.arange(n, dtype=np.uint64)
    u1 = uniform(seed, np.uint64(2) * k)
    u2 = uniform(seed, np.uint64(2) * k + np.uint64(1))
    return np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * np.pi * u2)


def lognormal(seed, n, median, sigma_log):
    return median * np.exp(sigma_log * normal(seed, n))
