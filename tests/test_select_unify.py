"""Merging the branches of a select (polymage_amd/cprint.py:
SelectUnifier, Printer(unify=...)): two branches built by the same
operators over the same C types share their arithmetic and select only
where they differ, when what is left to select is index arithmetic.  These
tests check the printed forms: the upsampling pattern of the pyramid
pipelines merges into one read at a selected index, branches whose C types
differ (a double literal against an int) or whose difference holds reads
stay two branches, and per-reference decisions follow the merged
reference.  The values are checked bit for bit against the oracle by the
emulator and GPU suites (local Laplacian, pyramid blending,
interpolation)."""
from polymage_amd.cprint import Printer, SelectUnifier, CSelect, same_expr
from polymage_amd.constructs import (Variable, Image, Function, Interval, Condition, Select,
                                     Parameter)
from polymage_amd.expr_types import Int, Float


def _setup():
    x = Variable(Int, 'x')
    y = Variable(Int, 'y')
    f = Image(Float, 'f', [64, 64])
    g = Image(Float, 'g', [64, 64])
    return x, y, f, g


def _cb(r, p, pro):
    return '%s[%s]' % (r.objectRef.name, ', '.join(p.expr(a, pro) for a in r.arguments))


def _printer(x, y, unify=lambda r: False):
    return Printer({x: 'q0', y: 'q1'}, _cb, 'double', lazy=True, unify=unify)


def test_upsampling_merges_into_one_selected_read():
    x, y, f, _ = _setup()
    even = Condition(y % 2, '==', 0)
    e = Select(even, 0.25 * f(x, y / 2 - 1) + 0.75 * f(x, y / 2),
               0.25 * f(x, y / 2 + 1) + 0.75 * f(x, y / 2))
    out = _printer(x, y).expr(e, [])
    # one read of each neighbour: the shared centre and the selected one
    assert out.count('f[') == 2, out
    assert '(((q1 / 2) - 1)) : (((q1 / 2) + 1))' in out, out
    plain = Printer({x: 'q0', y: 'q1'}, _cb, 'double', lazy=True).expr(e, [])
    assert plain.count('f[') == 4


def test_branches_of_different_c_type_stay_apart():
    x, y, f, _ = _setup()
    c = Condition(x, '<', 3)
    # 2.0 prints as a double literal, 2 as an int: the products differ in C type
    e = Select(c, f(x, y) * 2.0, f(x, y) * 2)
    u = SelectUnifier('double', {x: 'q0', y: 'q1'}).unify(c, e.true_expression,
                                                          e.false_expression)
    assert isinstance(u, CSelect) or u is None
    out = _printer(x, y).expr(e, [])
    assert out.count('f[') == 2


def test_differences_holding_reads_stay_branches():
    x, y, f, g = _setup()
    c = Condition(y % 2, '==', 0)
    e = Select(c, f(x, y / 2) + g(x, y), f(x, y / 2) + 4.0 * g(x, y))
    out = _printer(x, y).expr(e, [])
    # the difference (g vs 4*g) holds a read: printed as the two branches
    assert out.count('f[') == 2 and out.count('g[') == 2


def test_cheap_references_are_not_merged():
    x, y, f, _ = _setup()
    c = Condition(y % 2, '==', 0)
    e = Select(c, 0.5 * f(x, y - 1), 0.5 * f(x, y + 1))
    cheap = _printer(x, y, unify=lambda r: True).expr(e, [])
    assert cheap.count('f[') == 2
    merged = _printer(x, y).expr(e, [])
    assert merged.count('f[') == 1


def test_merged_reference_remembers_its_origins():
    x, y, f, _ = _setup()
    c = Condition(y % 2, '==', 0)
    a, b = f(x, y / 2 - 1), f(x, y / 2 + 1)
    u = SelectUnifier('double', {x: 'q0', y: 'q1'}).unify(c, a, b)
    assert u.objectRef is f and u.unified_from == (a, b)
    assert same_expr(u.arguments[0], x)
    assert isinstance(u.arguments[1], CSelect)


def test_same_expr():
    x, y, f, _ = _setup()
    w = Parameter(Float, 'w')
    assert same_expr(f(x, y + 1) * w, f(x, y + 1) * w)
    assert not same_expr(f(x, y + 1), f(x, y - 1))
    assert not same_expr(f(x, y) * 2, f(x, y) * 2.0)
    assert not same_expr(Variable(Int, 'x'), x)
