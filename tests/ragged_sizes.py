"""Size list of the ragged-batch workload: target T = 16 * 4096^u with u
uniform from the seed-0 SplitMix64 stream; with probability 0.4 the smallest
prime >= T (<= 65521), else the smallest 5-smooth number (2^a 3^b 5^c) >= T
(<= 65536)."""
import numpy as np

import oracle_lib as ol


def _is_prime(n):
    if n < 2:
        return False
    if n % 2 == 0:
        return n == 2
    r = int(n ** 0.5)
    for d in range(3, r + 1, 2):
        if n % d == 0:
            return False
    return True


def _smooth_list(limit=65536):
    out = []
    a = 1
    while a <= limit:
        b = a
        while b <= limit:
            c = b
            while c <= limit:
                out.append(c)
                c *= 5
            b *= 3
        a *= 2
    return np.array(sorted(out))


_SMOOTH = _smooth_list()


def ragged_sizes(count, seed=0):
    u = (ol.uniform_np(seed, 2 * count) + 1.0) / 2.0  # [0, 1)
    sizes = []
    for i in range(count):
        t = 16.0 * 4096.0 ** u[2 * i]
        if u[2 * i + 1] < 0.4:
            n = int(np.ceil(t))
            while not _is_prime(n):
                n += 1
            n = min(n, 65521)
        else:
            k = np.searchsorted(_SMOOTH, t)
            n = int(_SMOOTH[min(k, len(_SMOOTH) - 1)])
        sizes.append(n)
    return sizes
