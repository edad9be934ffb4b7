"""The draws of std::mt19937 + std::uniform_real_distribution<double>(0, 1)
as libstdc++ makes them, in Python, so that tests can feed the library the
uniforms the dumpconfig program draws (DumpConfig/utilityb.cpp:13-31 uses
the same pair).

MT19937: Matsumoto & Nishimura (1998), 32-bit, seeded by init_genrand as
std::mersenne_twister_engine::seed(value) is specified.  The double: the
C++ standard's generate_canonical<double, 53> with two 32-bit draws,
(x1 + x2 2^32) / 2^64 evaluated in double, 1 replaced by the largest double
below 1 (the libstdc++ form).
"""


class MT19937:
    N, M = 624, 397

    def __init__(self, seed=5489):
        mt = [0] * self.N
        mt[0] = seed & 0xFFFFFFFF
        for i in range(1, self.N):
            mt[i] = (1812433253 * (mt[i - 1] ^ (mt[i - 1] >> 30)) + i) & 0xFFFFFFFF
        self.mt = mt
        self.i = self.N

    def _twist(self):
        mt = self.mt
        for k in range(self.N):
            y = (mt[k] & 0x80000000) | (mt[(k + 1) % self.N] & 0x7FFFFFFF)
            mt[k] = mt[(k + self.M) % self.N] ^ (y >> 1) ^ (0x9908B0DF if y & 1 else 0)
        self.i = 0

    def next_u32(self):
        if self.i >= self.N:
            self._twist()
        y = self.mt[self.i]
        self.i += 1
        y ^= y >> 11
        y ^= (y << 7) & 0x9D2C5680
        y ^= (y << 15) & 0xEFC60000
        y ^= y >> 18
        return y

    def uniform(self):
        s = float(self.next_u32())
        s += float(self.next_u32()) * 4294967296.0
        r = s / 18446744073709551616.0
        return r if r < 1.0 else 0.99999999999999989


def uniforms(seed, count):
    m = MT19937(seed)
    return [m.uniform() for _ in range(count)]
