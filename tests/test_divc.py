"""The exact rewrite of a narrow integer divided by a double literal
(polymage_amd/cprint.py: divc_exact, pmx_divc): the check is exhaustive
over the integer type and rejects what is not exact, and the printer
uses the rewrite only for such quotients.  Device and host parity of the
pipelines that use it (the local Laplacian's colour / 65535.0) is covered
by the oracle comparisons of the GPU, emulator and CPU-backend suites."""
from fractions import Fraction

from polymage_amd.cprint import Printer, divc_exact, narrow_int_range
from polymage_amd.constructs import Variable, Image, Function, Interval, Cast
from polymage_amd.expr_types import Int, UShort, Float


def _two_fma(n, d):
    r = 1.0 / d
    q = n * r
    e = float(n - Fraction(q) * Fraction(d))
    return float(Fraction(e) * Fraction(r) + Fraction(q))


def test_check_accepts_exact_and_the_naive_product_is_not():
    assert divc_exact(65535.0, 0, 65535)
    assert divc_exact(255.0, -32768, 32767)
    # the rewrite without its correction step differs on 88 of the 65536
    # u16 values, so a check that accepted it would be wrong
    r = 1.0 / 65535.0
    bad = [n for n in range(65536) if n * r != n / 65535.0]
    assert len(bad) == 88
    assert all(_two_fma(n, 65535.0) == n / 65535.0 for n in bad)


def test_check_rejects_zero_and_non_finite_divisors():
    assert not divc_exact(0.0, 0, 255)
    assert not divc_exact(float('inf'), 0, 255)
    assert not divc_exact(float('nan'), 0, 255)


def test_printer_rewrites_only_narrow_integer_over_double_literal():
    x = Variable(Int, 'x')
    img = Image(UShort, 'img', [16])
    f = Function(([x], [Interval(Int, 0, 15)]), Float, 'f')
    pr = Printer({x: 'x'}, lambda r, p, pro: 'img[x]', 'double')
    assert narrow_int_range(img(x)) == (0, 65535)
    assert pr.expr(img(x) / 65535.0).startswith('pmx_divc(')
    assert pr.expr(Cast(UShort, f(x)) / 3.0).startswith('pmx_divc(')
    # a float numerator, an int numerator, an integer divisor, float literals
    assert 'pmx_divc' not in pr.expr(f(x) / 65535.0)
    assert 'pmx_divc' not in pr.expr(Cast(Int, f(x)) / 65535.0)
    assert 'pmx_divc' not in pr.expr(img(x) / 255)
    pf = Printer({x: 'x'}, lambda r, p, pro: 'img[x]', 'float12')
    assert 'pmx_divc' not in pf.expr(img(x) / 65535.0)
