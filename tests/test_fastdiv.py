"""The division by invariant integers used to split block and thread indices
of the tiled sweeps (FastDiv in csrc/lmx_batch.h): q = umulhi(n, mul) >> shr
with l = ceil(log2 d), mul = ceil(2^(31+l) / d), shr = l - 1, for n < 2^31.
Checked here against integer division for every divisor up to 2^13 and
divisors around powers of two up to 2^30, at the quotient boundaries and the
largest dividend; the GPU parity tests run the kernels that use it on many
lattice shapes."""
import numpy as np


def make_fastdiv(d):
    if d == 1:
        return 0, 0
    l = (d - 1).bit_length()          # ceil(log2 d)
    mul = ((1 << (31 + l)) + d - 1) // d
    assert mul < 1 << 32
    return mul, l - 1


def fast_div(n, d, mul, shr):
    return n if d == 1 else ((n * mul) >> 32) >> shr


def check(d, ns):
    mul, shr = make_fastdiv(d)
    for n in ns:
        if 0 <= n < 1 << 31:
            assert fast_div(n, d, mul, shr) == n // d, (n, d)


def boundaries(d):
    top = (1 << 31) - 1
    ks = list(range(0, 64)) + [top // d - 1, top // d, top // d + 1]
    ns = {top, 0, 1}
    for k in ks:
        for e in (-1, 0, 1):
            ns.add(k * d + e)
    return ns


def test_small_divisors_exhaustive_quotient_boundaries():
    for d in range(1, 1 << 13):
        check(d, boundaries(d))


def test_divisors_near_powers_of_two():
    for l in range(1, 31):
        for d in ((1 << l) - 1, 1 << l, (1 << l) + 1):
            check(d, boundaries(d))


def test_random_dividends():
    rng = np.random.default_rng(3)
    for d in rng.integers(1, 1 << 31, 200):
        d = int(d)
        check(d, [int(n) for n in rng.integers(0, 1 << 31, 200)] + list(boundaries(d)))
