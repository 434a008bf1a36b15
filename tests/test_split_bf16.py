"""The split-bf16 arithmetic of PREC_BF16X3, restated on the CPU.

A value x is held as hi = bf16(x), lo = bf16(x - hi) (round to nearest even),
and a product x.m as hi.m_hi + hi.m_lo + lo.m_hi with FP32 accumulation.  These
tests pin the two bounds the kernels rely on: the pair carries x to 2^-17
relative, and a dot product of the three partial products matches the FP64
dot product to about 2^-16 of the sum of |x_k m_k| (plus FP32 accumulation).
"""
import numpy as np


def bf16(x):
    """Round FP32 to bf16 (nearest even), returned as FP32."""
    b = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    b = (b + 0x7FFF + ((b >> 16) & 1)) & 0xFFFF0000
    return b.astype(np.uint32).view(np.float32)


def split(x):
    x = np.asarray(x, dtype=np.float32)
    hi = bf16(x)
    lo = bf16((x - hi).astype(np.float32))
    return hi, lo


def test_bf16_rounding_is_nearest_even():
    # 1 + 2^-8 is halfway between 1 and 1 + 2^-7: ties go to the even mantissa.
    assert bf16(np.float32(1 + 2 ** -8)) == np.float32(1.0)
    assert bf16(np.float32(1 + 3 * 2 ** -8)) == np.float32(1 + 2 ** -6)
    assert bf16(np.float32(1 + 2 ** -8 + 2 ** -20)) == np.float32(1 + 2 ** -7)
    # Values with 8 significant bits are kept exactly.
    v = np.float32(1.5 + 2 ** -7)
    assert bf16(v) == v


def test_pair_carries_seventeen_bits():
    rng = np.random.default_rng(0)
    x = (rng.lognormal(0, 3, 200000) * rng.choice([-1, 1], 200000)).astype(np.float32)
    hi, lo = split(x)
    rel = np.abs((hi.astype(np.float64) + lo) - x) / np.abs(x)
    assert rel.max() <= 2.0 ** -17
    # hi alone is the bf16 value (2^-9), so the lo half matters.
    assert (np.abs(hi.astype(np.float64) - x) / np.abs(x)).max() > 2.0 ** -10


def test_three_products_against_fp64():
    rng = np.random.default_rng(1)
    for P in (32, 256, 1024):
        n = rng.uniform(0, 5000, (64, P)).astype(np.float32)          # populations
        m = rng.dirichlet(np.ones(P), P).astype(np.float32)            # rows sum to 1
        nh, nl = split(n)
        mh, ml = split(m)
        f = lambda a, b: a.astype(np.float32) @ b.astype(np.float32)
        got = f(nh, mh) + f(nh, ml) + f(nl, mh)
        ref = n.astype(np.float64) @ m.astype(np.float64)
        scale = np.abs(n.astype(np.float64)) @ np.abs(m.astype(np.float64))
        err = np.abs(got - ref) / scale
        # 2^-16 from the split, plus FP32 accumulation over P terms.
        assert err.max() <= 2.0 ** -16 + P * 2.0 ** -24
        # Dropping hi.m_lo would cost about 2^-9: the three products are needed.
        two = f(nh, mh) + f(nl, mh)
        assert (np.abs(two - ref) / scale).max() > 2.0 ** -12
