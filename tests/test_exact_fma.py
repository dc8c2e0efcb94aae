"""The fused multiply-add of an exact product (polymage_amd/cprint.py:
exact_product, Printer(exact_fma=True)): l + a*k with a a float (24
significand bits) or an integer of at most 16 bits and k a double literal
of at most 53 - bits(a) significand bits is printed as fma(a, k, l).  The
product is then exact in double, so the separate multiply-then-add and the
fused form round the same exact sum once.  These tests check that claim in
exact rational arithmetic on random operands, that it fails one bit past
the bound (so the bound is the one that matters), and that the printer
applies the rewrite only inside the bound.  Pipeline parity with the
rewrite in the kernels is covered by the emulator and GPU suites against
the oracle."""
import random
import struct
from fractions import Fraction

from polymage_amd.cprint import Printer, exact_product, _const_bits
from polymage_amd.constructs import Variable, Image, Function, Interval, Cast, Parameter
from polymage_amd.expr_types import Int, UShort, Float, Double


def _f32(v):
    return struct.unpack('f', struct.pack('f', v))[0]


def _fused(a, k, l):
    # RN(a*k + l) with one rounding: Fraction -> float rounds to nearest even
    return float(Fraction(a) * Fraction(k) + Fraction(l))


def _separate(a, k, l):
    return (a * k) + l


def _rand_f32(rng):
    e = rng.randint(-30, 30)
    return _f32(rng.uniform(-1, 1) * 2.0 ** e)


def test_const_bits():
    assert _const_bits(4.0) == 1
    assert _const_bits(6.0) == 2
    assert _const_bits(0.0625) == 1
    assert _const_bits(0.375) == 2
    assert _const_bits(0.1) == 52   # 0x3FB999999999999A: one trailing zero
    assert _const_bits(-3.0) == 2
    assert _const_bits(0.0) is None
    assert _const_bits(float('inf')) is None
    assert _const_bits(2.0 ** 900) is None


def test_fused_equals_separate_inside_the_bound():
    rng = random.Random(7)
    ks = [4.0, 6.0, 0.0625, 0.375, -3.0, 12345.0, float((1 << 29) - 1)]
    for k in ks:
        assert 24 + _const_bits(k) <= 53
        for _ in range(3000):
            a, l = _rand_f32(rng), rng.uniform(-1, 1) * 2.0 ** rng.randint(-40, 40)
            assert _fused(a, k, l) == _separate(a, k, l), (a, k, l)
        for _ in range(500):
            a, l = float(rng.randint(0, 65535)), rng.uniform(-1e6, 1e6)
            assert _fused(a, k, l) == _separate(a, k, l), (a, k, l)


def test_fused_differs_past_the_bound():
    # k of 30 significand bits with a float of 24: the product needs up to
    # 54 bits and is rounded by the separate multiply, so some sums differ
    rng = random.Random(11)
    k = float((1 << 30) - 1)
    assert _const_bits(k) == 30
    diff = 0
    for _ in range(2000):
        a = _f32(rng.uniform(1, 2))
        l = rng.uniform(-1, 1)
        diff += _fused(a, k, l) != _separate(a, k, l)
    assert diff > 0


def test_printer_rewrites_only_exact_products():
    x = Variable(Int, 'x')
    img = Image(Float, 'img', [16])
    u16 = Image(UShort, 'u16', [16])
    i32 = Image(Int, 'i32', [16])
    dbl = Image(Double, 'dbl', [16])
    w = Parameter(Float, 'w')
    f = Function(([x], [Interval(Int, 0, 15)]), Float, 'f')
    cb = lambda r, p, pro: '%s[x]' % r.objectRef.name
    pr = Printer({x: 'x'}, cb, 'double', exact_fma=True)
    assert pr.expr(img(x) + img(x + 1) * 4.0) == \
        'pmx_fmad((double)(img[x]), 4.0, (double)(img[x]))'
    assert pr.expr(img(x) * 6.0 + f(x)).startswith('pmx_fmad((double)(img[x]), 6.0, ')
    assert pr.expr(f(x) - img(x) * 4.0).startswith('pmx_fmad((double)(img[x]), -4.0, ')
    assert pr.expr(img(x) * 4.0 - f(x)) == 'pmx_fmad((double)(img[x]), 4.0, -(double)(f[x]))'
    assert pr.expr(u16(x) * 3.0 + 1.5).startswith('pmx_fmad(')
    assert pr.expr(Cast(Float, i32(x)) * 0.5 + img(x)).startswith('pmx_fmad(')
    assert pr.expr(w * 2.0 + img(x)).startswith('pmx_fmad(')
    # inexact products, wide operands, integer literals, multiply alone
    assert 'pmx_fmad' not in pr.expr(img(x) * 0.1 + img(x))
    assert 'pmx_fmad' not in pr.expr(i32(x) * 4.0 + img(x))
    assert 'pmx_fmad' not in pr.expr(dbl(x) * 4.0 + img(x))
    assert 'pmx_fmad' not in pr.expr(img(x) * 4 + img(x))
    assert 'pmx_fmad' not in pr.expr(img(x) * 4.0)
    assert exact_product(img(x) * float((1 << 29) - 1)) is not None
    assert exact_product(img(x) * float((1 << 30) - 1)) is None
    # off by default, and never in float-literal mode
    assert 'pmx_fmad' not in Printer({x: 'x'}, cb, 'double').expr(img(x) + img(x) * 4.0)
    assert 'pmx_fmad' not in Printer({x: 'x'}, cb, 'float12', exact_fma=True).expr(
        img(x) + img(x) * 4.0)
