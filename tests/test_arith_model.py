"""CPU model of the short-latency square root and reciprocal of the reflector and rotation chains
(gevp-pdhgeqz_amd/csrc/gqz_dev.h: d_sqrtf, d_rcpn; k_ht_blk.hip: ht_lartg).

The model runs the same instruction sequence with exact rounding.  Every product and every fused
multiply-add is rounded once from the exact rational value.  The hardware seed (v_rsq_f64 / v_rcp_f64)
is the true value perturbed by a relative error of up to 2^-23, the precision those instructions are
specified to.

Over the operand range the chains use (1e-280 .. 1e280), both results must be correctly rounded on
random operands, so the rounding of tau and of the rotations stays unbiased.  Operands crafted next
to a rounding midpoint (1 - 2^-53 for the reciprocal) may miss by one ulp, never more.  Dropping the
square root's residual correction, or one Newton step of the reciprocal, must break this.

This pins the numerical method, not the device: the GPU tests compare the device results with the
oracle and LAPACK."""
import math
import random
from fractions import Fraction as F

SEED_ERR = 2.0 ** -23


def _fma(a, b, c):
    return float(F(a) * F(b) + F(c))


def _mul(a, b):
    return float(F(a) * F(b))


def sqrt_model(x, e, correct=True):
    y = (1.0 / math.sqrt(x)) * (1.0 + e)          # v_rsq_f64 seed
    g, h = _mul(x, y), _mul(0.5, y)
    r = _fma(-g, h, 0.5)
    g, h = _fma(g, r, g), _fma(h, r, h)
    return _fma(_fma(-g, g, x), h, g) if correct else g


def rcp_model(x, e, steps=2):
    r = (1.0 / x) * (1.0 + e)                     # v_rcp_f64 seed
    for _ in range(steps):
        r = _fma(r, _fma(-x, r, 1.0), r)
    return r


SPECIAL = [1.0000000001e-280, 9.999999e279, 1.0, 2.0, 0.5, 3.0, 1.0 - 2.0 ** -53, 1.0 + 2.0 ** -52,
           2.0 - 2.0 ** -52, (1.0 + 2.0 ** -52) ** 2]


def _samples(n, seed):
    rng = random.Random(seed)
    xs = [10.0 ** rng.uniform(-280, 280) * rng.uniform(1.0, 10.0) for _ in range(n)]
    es = [rng.uniform(-SEED_ERR, SEED_ERR) for _ in xs]
    return xs, es


def _ulps(a, b):
    return abs(a - b) / math.ulp(b)


def test_square_root_is_correctly_rounded():
    xs, es = _samples(3000, 1)
    for x, e in zip(xs, es):
        for ee in (e, SEED_ERR, -SEED_ERR):
            assert sqrt_model(x, ee) == math.sqrt(x), (x, ee)
    for x in SPECIAL:
        for ee in (0.0, SEED_ERR, -SEED_ERR):
            assert _ulps(sqrt_model(x, ee), math.sqrt(x)) <= 1.0, (x, ee)


def test_reciprocal_is_correctly_rounded():
    xs, es = _samples(3000, 2)
    for x, e in zip(xs, es):
        for ee in (e, SEED_ERR, -SEED_ERR):
            assert rcp_model(x, ee) == 1.0 / x, (x, ee)
    for x in SPECIAL:
        for ee in (0.0, SEED_ERR, -SEED_ERR):
            assert _ulps(rcp_model(x, ee), 1.0 / x) <= 1.0, (x, ee)


def test_planted_shortcuts_fail():
    xs, es = _samples(400, 3)
    no_corr = sum(sqrt_model(x, e, correct=False) != math.sqrt(x) for x, e in zip(xs, es))
    one_step = sum(rcp_model(x, e, steps=1) != 1.0 / x for x, e in zip(xs, es))
    assert no_corr > len(xs) // 10
    assert one_step > len(xs) // 2
