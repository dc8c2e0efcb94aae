"""DSL front-end known answers.

Same facts the reference's own unit tests pin (sandbox/tests/
test_constructs.py:12-172, test_expression.py:10-47, test_stencil.py):
string forms, conjunct splitting, affine coefficients, overload shortcuts
and Stencil macro expansion."""
from fractions import Fraction

from polymage_amd.constructs import *  # noqa: F401,F403
from polymage_amd.expression import (isAffine, get_affine_var_and_param_coeff,
                                     get_constant_from_expr, getType)


def s(e):
    return str(e).replace(' ', '')


def test_param_variable_types():
    N = Parameter(Float, "N")
    assert N.has(Variable) is False and N.has(Parameter) is True
    assert N.name == "N" and N.typ == Float
    x = Variable(UInt, "x")
    assert x in x.collect(Variable) and x not in x.collect(Parameter)
    for t in (Float, Double):
        assert Value(3.0, t) == 3.0
    for t in (Int, UInt, Char, UChar, Short, UShort, Long, ULong):
        assert Value(3, t) == 3


def test_interval_image():
    I = Interval(UInt, Value(3, UInt), Value(5, UInt))
    assert I.typ == UInt and I.lowerBound == 3 and I.upperBound == 5
    N = Parameter(UInt, "N")
    img = Image(UChar, "Input", [N, N, 3])
    assert img.typ == UChar and img.name == "Input"
    assert img.dimensions[0] is N and img.dimensions[2] == 3


def test_expr_strings():
    N = Parameter(UInt, "N")
    x, y, z = Variable(UInt, "x"), Variable(UInt, "y"), Variable(UInt, "z")
    w = x * (y * z) - N
    assert N in w.collect(Parameter)
    assert s(w) == "((x*(y*z))-N)"
    w = (+(((-(((x - (y + z)) / z) % 2) ^ 3) >> z) << 5) | 1) & y
    assert s(w) == "((+((((-((((x-(y+z))/z)%2))^3)>>z)<<5))|1)&y)"
    assert s(x * 3) == "(x*3)"


def test_conditions_and_conjuncts():
    N = Parameter(UInt, "N")
    x, y = Variable(UInt, "x"), Variable(UInt, "y")
    c1 = Condition(x + N, '>', y)
    assert s(c1) == "((x+N)>y)"
    assert s(c1 & Condition(x, '<', 5)) == "(((x+N)>y)&&(x<5))"
    cond = (Condition(x, '<', y) & Condition(x, '>', N)) | \
        (Condition(x, '<', 3) | Condition(x + y, '<=', N)) & Condition(x, '!=', 2 * N)
    got = [''.join(str(c) for c in conj).replace(' ', '') for conj in cond.split_to_conjuncts()]
    assert got == ["(x<y)(x>N)", "(x<3)(x<(2*N))", "(x<3)(x>(2*N))",
                   "((x+y)<=N)(x<(2*N))", "((x+y)<=N)(x>(2*N))"]


def test_case_function():
    N = Parameter(UInt, "N")
    x, y = Variable(UInt, "x"), Variable(UInt, "y")
    c = Case(Condition(x + 1, '>', y), x - y)
    assert s(c.condition) == "((x+1)>y)" and s(c.expression) == "(x-y)"
    r = Interval(UInt, 0, N - 1)
    f = Function(([x, y], [r, r]), UInt, "max")
    f.defn = [Case(Condition(x, '>', y), x), Case(Condition(x, '<=', y), y)]
    assert s(f.defn[0]) == "Case((x>y)){x}" and s(f.defn[1]) == "Case((x<=y)){y}"
    g = Function(([x, y], [r, r]), UInt, "add")
    g.defn = [x + y]
    assert r in g.getObjects(Interval)


def test_affine():
    N = Parameter(UInt, "N")
    x, y = Variable(UInt, "x"), Variable(UInt, "y")
    assert isAffine(x + y) and isAffine(3) and not isAffine(x * y)
    assert isAffine(-x + N + 3 * y) and isAffine(2 * x + N / 2 + 3 * y)
    assert isAffine(Condition(x, '<', 2 * y)) and not isAffine(Condition(x, '>=', x * y))
    assert not isAffine(Condition(x * N, '!=', y))
    e = (2 * x + N // 2 + 3 * y) * 3 // 2
    co = get_affine_var_and_param_coeff(e)
    assert co[x] == Fraction(3) and co[N] == Fraction(3, 4) and co[y] == Fraction(9, 2)
    assert get_constant_from_expr((N - 1) // 2, affine=True) == Fraction(-1, 2)
    co = get_affine_var_and_param_coeff(2 * (x * 3 + y + N + x + y - 5) + 3 * (-x) + 4 * (-y) + N)
    assert co[x] == 5 and co[y] == 0 and co[N] == 3
    assert get_affine_var_and_param_coeff(N * x + y) == {}


def test_overload_shortcuts():
    N = Parameter(UInt, "N")
    assert s(1 // N) == "(1/N)"
    for e in (N // 1, 1 * N, N * 1, N + 0, N >> 0, N << 0):
        assert s(e) == "N"
    assert s(N * 0) == "0" and s(0 * N) == "0"
    x = Variable(Int, "x")
    assert s(0 - x) == "-(x)"
    assert s(x * 1.0) == "x"          # 1.0 == 1: no promotion (expr_ast.py:72-77)


def test_types_and_select():
    x = Variable(Int, "x")
    img = Image(Float, "img", [10])
    assert getType(img(x) * 2) is Float
    assert getType(img(x) * 0.5) is Float       # the C code computes this in double
    assert getType(Cast(Short, x) + Cast(UChar, x)) is Short
    m = Max(img(x), img(x + 1))
    assert s(m) == "((img(x)>img((x+1)))?img(x):img((x+1)))"
    assert getType(Select(Condition(x, '>', 0), Cast(UChar, x), Cast(UChar, 1))) is UChar


def test_stencil_expansion():
    # same kernels and expected strings as sandbox/tests/test_stencil.py:18-57
    R, C = Parameter(Int, "R"), Parameter(Int, "C")
    x, y = Variable(Int, "x"), Variable(Int, "y")
    xrow, xcol = Interval(Int, 1, R), Interval(Int, 1, C)
    bounds = Condition(x, '>=', 2) & Condition(x, '<=', R - 2) & \
        Condition(y, '>=', 2) & Condition(y, '<=', C - 2)
    img = Image(Float, "input", [R + 2, C + 2])
    k1 = [[1, 0, 0], [0, 2, 0], [0, 0, 1]]
    st1 = Stencil(img, [x, y], k1)
    assert st1.sizes == [3, 3] and st1.origin == [1, 1] and st1.kernel == k1
    assert str(Stencil.macro_expand(st1)) == \
        '((input((x + -1), (y + -1)) + (input(x, y) * 2)) + input((x + 1), (y + 1)))'
    f1 = Function(([x, y], [xrow, xcol]), Float, "f1")
    f1.defn = [Case(bounds, st1)]
    k2 = [[-1, 2, -3.8], [0, -1, 0]]
    st2 = Stencil(f1, [x, y], k2)
    assert st2.sizes == [2, 3] and st2.origin == [0, 1]
    assert str(Stencil.macro_expand(st2)) == \
        '((((f1(x, (y + -1)) * -1) + (f1(x, y) * 2)) + (f1(x, (y + 1)) * -3.8)) + (f1((x + 1), y) * -1))'
    # a Stencil inside a Case is expanded when the Case is built
    assert str(f1.defn[0].expression).startswith('((input((x + -1)')


def test_reduction_rejects_unbound_variables():
    """A Reduce may only name its reduction variables: the accumulation
    loops over the reduction domain, so any other variable has no value.
    The reference's max_filter writes such a reduction
    (polymage_maxfilter.py:60-66: vert_log(x, y, c, t) over (x, rx, c, ry))."""
    import pytest
    R = Parameter(Int, "R")
    x, y, rx = Variable(Int, "x"), Variable(Int, "y"), Variable(Int, "rx")
    img = Image(Float, "img", [R, R])
    rows = Interval(Int, 0, R - 1)
    ok = Reduction(([x], [rows]), ([x, rx], [rows, rows]), Float, "rowmax")
    ok.defn = [Reduce(ok(x), img(x, rx), Op.Max)]
    bad = Reduction(([x, y], [rows, rows]), ([x, rx], [rows, rows]), Float, "bad")
    with pytest.raises(TypeError, match="y not among its reduction variables"):
        bad.defn = [Reduce(bad(x, y), img(x, rx), Op.Max)]
    bad2 = Reduction(([x], [rows]), ([x, rx], [rows, rows]), Float, "bad2")
    with pytest.raises(TypeError, match="y not among"):
        bad2.defn = [Case(Condition(x, '>', 0), Reduce(bad2(x), img(y, rx), Op.Sum))]
    # conditions on the function's own variables select target points
    sel = Reduction(([y, x], [Interval(Int, 0, 1), rows]), ([x, rx], [rows, rows]), Float, "sel")
    sel.defn = [Case(Condition(y, '==', 0), Reduce(sel(0, x), img(x, rx), Op.Sum)),
                Case(Condition(y, '==', 1), Reduce(sel(1, x), 1, Op.Sum))]
