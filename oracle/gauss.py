"""CPU oracle of RBMD's GAUSS velocity creation -- TEST INFRASTRUCTURE ONLY.

Only tests/ may import this module.  Pure-Python restatement of
LammpsFullReader::ApplyGaussianVelocitiesIfRequested and MixVelocitySeed
(src/reader/src/format/lammps/lammps_full_reader.cpp:134-143,710-866), which
draw from the C++ standard library; the library algorithms are restated from
their published definitions:

* std::mt19937_64: the 64-bit Mersenne Twister with the parameters of the C++
  standard ([rand.predef]; Matsumoto and Nishimura);
* std::normal_distribution<double> as libstdc++ implements it: the Marsaglia
  polar method on 2 u - 1 from std::generate_canonical (one 64-bit draw / 2^64
  per uniform), returning y * mult first and keeping x * mult for the next call.

Pins (tests/test_oracle_gauss.py): the standard's check value (the 10000th
output of a default-seeded std::mt19937_64 is 9981545732273789042); the
polar method's draws are standard normal (moments, and every pair lies on the
circle the method maps from); the created velocities have zero total momentum
and exactly the requested temperature with 3N - 3 degrees of freedom.
Pure Python: a few thousand atoms per second.
"""
import math

import numpy as np

M64 = (1 << 64) - 1


class MT19937_64:
    N, M = 312, 156
    A = 0xB5026F5AA96619E9
    UPPER, LOWER = M64 ^ ((1 << 31) - 1), (1 << 31) - 1

    def __init__(self, seed=5489):
        mt = [seed & M64]
        for i in range(1, self.N):
            prev = mt[-1]
            mt.append((6364136223846793005 * (prev ^ (prev >> 62)) + i) & M64)
        self.mt, self.i = mt, self.N

    def _twist(self):
        mt, N, M = self.mt, self.N, self.M
        for k in range(N):
            x = (mt[k] & self.UPPER) | (mt[(k + 1) % N] & self.LOWER)
            xa = x >> 1
            if x & 1:
                xa ^= self.A
            mt[k] = mt[(k + M) % N] ^ xa
        self.i = 0

    def __call__(self):
        if self.i >= self.N:
            self._twist()
        y = self.mt[self.i]
        self.i += 1
        y ^= (y >> 29) & 0x5555555555555555
        y ^= (y << 17) & 0x71D67FFFEDA60000
        y ^= (y << 37) & 0xFFF7EEE000000000
        y ^= y >> 43
        return y & M64


def canonical(gen):
    """std::generate_canonical<double, 53> over a 64-bit engine."""
    r = float(gen()) / 18446744073709551616.0
    return r if r < 1.0 else math.nextafter(1.0, 0.0)


class NormalPolar:
    """std::normal_distribution<double>(0, 1) as libstdc++ draws it."""

    def __init__(self):
        self.saved = None

    def __call__(self, gen):
        if self.saved is not None:
            r, self.saved = self.saved, None
            return r
        while True:
            x = 2.0 * canonical(gen) - 1.0
            y = 2.0 * canonical(gen) - 1.0
            r2 = x * x + y * y
            if not (r2 > 1.0 or r2 == 0.0):
                break
        mult = math.sqrt(-2.0 * math.log(r2) / r2)
        self.saved = x * mult
        return y * mult


def mix_velocity_seed(seed, atom_id):
    v = (seed + 0x9E3779B97F4A7C15 + ((atom_id << 6) & M64) + (atom_id >> 2)) & M64
    v ^= v >> 30
    v = (v * 0xBF58476D1CE4E5B9) & M64
    v ^= v >> 27
    v = (v * 0x94D049BB133111EB) & M64
    v ^= v >> 31
    return v


def gauss_velocities(ids, types, mass, temperature, seed, kb=1.0, mvv2e=1.0):
    """(N, 3) velocities for 0-based ids / types and per-type masses."""
    n = len(ids)
    v = np.zeros((n, 3))
    m = np.array([mass[t] for t in types], dtype=float)
    mt, p = 0.0, [0.0, 0.0, 0.0]
    for i in range(n):
        gen = MT19937_64(mix_velocity_seed(seed, int(ids[i]) + 1))
        normal = NormalPolar()
        f = 1.0 / math.sqrt(m[i])
        for d in range(3):
            v[i, d] = normal(gen) * f
        mt += m[i]
        for d in range(3):
            p[d] += m[i] * v[i, d]
    bias = [p[d] / mt for d in range(3)]
    ts = 0.0
    for i in range(n):
        for d in range(3):
            v[i, d] = v[i, d] - bias[d]
        ts += m[i] * (v[i, 0] * v[i, 0] + v[i, 1] * v[i, 1] + v[i, 2] * v[i, 2])
    ts *= mvv2e
    scale = math.sqrt(temperature * float(3 * n - 3) * kb / ts)
    return v * scale
