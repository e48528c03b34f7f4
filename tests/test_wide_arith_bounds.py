"""Bounds of the 64-bit modular products of csrc/modarith.hpp, replayed on the
host with exact integers (and IEEE doubles where the device uses them), on
random and adversarial operands.

- The approximate Shoup quotient (three 32-bit partial products, the low x low
  product and the carries out of the low word dropped) is at most 3 short, so
  the remainder x w + qh (2^64 - Q) mod 2^64 lies in [0, 5Q) and is congruent
  to x w.
- reduce_small's quotient from a double multiply by 1/Q is the true quotient
  or one off for every x < 2^62, so x - qh Q lies in [-Q, 2Q) and the
  corrections give x mod Q.
"""
import random

import pytest

M64 = (1 << 64) - 1
PRIMES = [0x3FFFFFFFFED001,      # the 54-bit sets (2^54 - 77823)
          0x3FFFFFFFFEB001,      # the next 54-bit prime = 1 mod 4096
          0x1FFFFFFFF9001,       # 49-bit, = 1 mod 4096
          0x1FFFFC0001]          # 37-bit (the reference's 192-bit app set width)


def shoup64(w, Q):
    return (w << 64) // Q


def mul_shoup_lazy(x, w, wp, Q):
    xl, xh = x & 0xFFFFFFFF, x >> 32
    pl, ph = wp & 0xFFFFFFFF, wp >> 32
    mid = ((xl * ph) >> 32) + ((xh * pl) >> 32)
    qh = (xh * ph + mid) & M64
    return (x * w + qh * ((1 << 64) - Q)) & M64


def reduce_small(x, Q):
    qh = int(float(x) * (1.0 / float(Q)))  # trunc(fl(fl(x) * fl(1/Q)))
    r = x - qh * Q
    assert -Q <= r < 2 * Q, (x, Q, r)
    if r < 0:
        r += Q
    if r >= Q:
        r -= Q
    if r < 0:
        r += Q
    return r


def operands(Q, rng, n):
    edge = [0, 1, Q - 1, Q, Q + 1, M64, M64 - 1, (1 << 63), (1 << 60) - 1, 56 * Q - 1]
    for x in edge:
        yield x & M64
    for _ in range(n):
        yield rng.getrandbits(64)
        yield rng.randrange(0, 56 * Q)  # the forward transform's lazy bound


@pytest.mark.parametrize("Q", PRIMES)
def test_lazy_shoup_product_below_5Q(Q):
    rng = random.Random(Q)
    ws = [0, 1, Q - 1, Q // 2, 1 << 53 if (1 << 53) < Q else Q - 2] + [rng.randrange(Q) for _ in range(40)]
    for w in ws:
        wp = shoup64(w, Q)
        for x in operands(Q, rng, 60):
            r = mul_shoup_lazy(x, w, wp, Q)
            assert r < 5 * Q, (x, w, Q)
            assert r % Q == (x * w) % Q, (x, w, Q)


@pytest.mark.parametrize("Q", PRIMES)
def test_reduce_small_quotient_at_most_one_off(Q):
    rng = random.Random(Q + 1)
    top = 1 << 62
    xs = [0, 1, Q - 1, Q, Q + 1, top - 1, (top // Q) * Q, (top // Q) * Q - 1]
    # multiples of Q and their neighbours across the whole range, where the
    # double estimate is most likely to land on the wrong side
    for _ in range(3000):
        k = rng.randrange(0, top // Q)
        for d in (-2, -1, 0, 1, 2):
            x = k * Q + d
            if 0 <= x < top:
                xs.append(x)
    xs += [rng.randrange(top) for _ in range(3000)]
    for x in xs:
        assert reduce_small(x, Q) == x % Q
