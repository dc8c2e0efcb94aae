"""Expression IR of the PolyMage DSL.

API-compatible restatement of sandbox/expr_ast.py:31-374 and the expression
classes of sandbox/constructs.py (Variable/Parameter 322-345, Reference
380-424, Cast 213-253, Select 256-310, Max/Min 312-319, the math builtins
77-211, Condition 722-803).  Expressions are trees built by Python operator
overloading on DSL objects; the same algebraic shortcuts as the reference
are applied while building (``e + 0 -> e``, ``e * 1 -> e``, ``e * 0 -> 0``,
``e / 1 -> e``, ``0 - e -> -e``; expr_ast.py:47-143), because they change the
emitted C and therefore the arithmetic types.

Printing (``str``) follows the reference's forms, which its own tests pin
(tests/test_constructs.py:53-172).
"""

from fractions import Fraction

from .expr_types import *  # noqa: F401,F403
from .expr_types import Int, Float, Rational, UInt


def _unique(objs):
    """Order-preserving de-duplication by identity."""
    seen = set()
    out = []
    for o in objs:
        if id(o) not in seen:
            seen.add(id(o))
            out.append(o)
    return out


def _as_expr(v):
    return Value.numericToValue(v)


class AbstractExpression(object):
    """Base of every expression node; overloads the arithmetic operators."""

    def _binop(self, other, cls, swap=False):
        other = _as_expr(other)
        if not isinstance(other, AbstractExpression):
            return NotImplemented
        return cls(other, self) if swap else cls(self, other)

    def __add__(self, other):
        other = _as_expr(other)
        if isinstance(other, Value) and other.value == 0:
            return self
        return self._binop(other, Add)

    def __radd__(self, other):
        other = _as_expr(other)
        if isinstance(other, Value) and other.value == 0:
            return self
        return self._binop(other, Add, swap=True)

    def __sub__(self, other):
        other = _as_expr(other)
        if isinstance(other, Value) and other.value == 0:
            return self
        return self._binop(other, Sub)

    def __rsub__(self, other):
        other = _as_expr(other)
        if isinstance(other, Value) and other.value == 0:
            return UnaryMinus(self)
        return self._binop(other, Sub, swap=True)

    def __mul__(self, other):
        other = _as_expr(other)
        if isinstance(other, Value):
            if other.value == 1:
                return self
            if other.value == 0:
                return Value(0, Int)
        return self._binop(other, Mul)

    def __rmul__(self, other):
        other = _as_expr(other)
        if isinstance(other, Value):
            if other.value == 1:
                return self
            if other.value == 0:
                return Value(0, Int)
        return self._binop(other, Mul, swap=True)

    def __truediv__(self, other):
        other = _as_expr(other)
        if isinstance(other, Value) and other.value == 1:
            return self
        return self._binop(other, Div)

    def __rtruediv__(self, other):
        return self._binop(other, Div, swap=True)

    # the DSL has a single (C) division; // and / build the same node
    __floordiv__ = __truediv__
    __rfloordiv__ = __rtruediv__

    def __lshift__(self, other):
        other = _as_expr(other)
        if isinstance(other, Value) and other.value == 0:
            return self
        return self._binop(other, LShift)

    def __rlshift__(self, other):
        return self._binop(other, LShift, swap=True)

    def __rshift__(self, other):
        other = _as_expr(other)
        if isinstance(other, Value) and other.value == 0:
            return self
        return self._binop(other, RShift)

    def __rrshift__(self, other):
        return self._binop(other, RShift, swap=True)

    def __mod__(self, other):
        return self._binop(other, Mod)

    def __rmod__(self, other):
        return self._binop(other, Mod, swap=True)

    def __and__(self, other):
        return self._binop(other, LAnd)

    def __rand__(self, other):
        return self._binop(other, LAnd, swap=True)

    def __xor__(self, other):
        return self._binop(other, Xor)

    def __rxor__(self, other):
        return self._binop(other, Xor, swap=True)

    def __or__(self, other):
        return self._binop(other, LOr)

    def __ror__(self, other):
        return self._binop(other, LOr, swap=True)

    def __neg__(self):
        return UnaryMinus(self)

    def __pos__(self):
        return UnaryPlus(self)

    __hash__ = object.__hash__

    # -- traversal --------------------------------------------------------
    def children(self):
        return []

    def collect(self, objType):
        objs = []
        for c in self.children():
            objs += c.collect(objType)
        if type(self) is objType:
            objs.append(self)
        return _unique(objs)

    def has(self, objType):
        return len(self.collect(objType)) > 0

    def macro_expand(self):
        return self

    def clone(self):
        return self


class Value(AbstractExpression):
    """A literal.  Python ints become Int, floats become Float (printed
    later as C double literals, expr_ast.py:193-203, targetc.py:80-86)."""

    @classmethod
    def numericToValue(cls, v):
        if isinstance(v, bool):
            return Value(int(v), Int)
        if type(v) is int:
            return Value(v, Int)
        if type(v) is float:
            return Value(v, Float)
        if type(v) is Fraction:
            return Value(v, Rational)
        return v

    def __init__(self, _value, _typ):
        self._value = _value
        self._typ = _typ

    @property
    def value(self):
        return self._value

    @property
    def typ(self):
        return self._typ

    def clone(self):
        return Value(self._value, self._typ)

    def __eq__(self, other):
        if isinstance(other, Value):
            return self._value == other._value
        return self._value == other

    def __ne__(self, other):
        return not self.__eq__(other)

    __hash__ = object.__hash__

    def __str__(self):
        return str(self._value)

    def __repr__(self):
        return 'Value(%r)' % (self._value,)


class AbstractBinaryOpNode(AbstractExpression):
    def __init__(self, _left, _right, _op=None):
        self._left = _left
        self._right = _right
        self._op = _op

    @property
    def left(self):
        return self._left

    @property
    def right(self):
        return self._right

    @property
    def op(self):
        return self._op

    def children(self):
        return [self._left, self._right]

    def clone(self):
        return type(self)(self._left.clone(), self._right.clone()) \
            if type(self) is not AbstractBinaryOpNode else \
            AbstractBinaryOpNode(self._left.clone(), self._right.clone(), self._op)

    def macro_expand(self):
        self._left = self._left.macro_expand()
        self._right = self._right.macro_expand()
        return self

    def __str__(self):
        return "(" + str(self._left) + " " + self._op + " " + str(self._right) + ")"


def _binop_class(name, sym):
    def __init__(self, _left, _right):
        AbstractBinaryOpNode.__init__(self, _left, _right, sym)
    return type(name, (AbstractBinaryOpNode,), {'__init__': __init__})


Add = _binop_class('Add', '+')
Sub = _binop_class('Sub', '-')
Mul = _binop_class('Mul', '*')
Div = _binop_class('Div', '/')
Mod = _binop_class('Mod', '%')
LShift = _binop_class('LShift', '<<')
RShift = _binop_class('RShift', '>>')
LAnd = _binop_class('LAnd', '&')
LOr = _binop_class('LOr', '|')
Xor = _binop_class('Xor', '^')


class AbstractUnaryOpNode(AbstractExpression):
    def __init__(self, _child, _op=None):
        self._child = _child
        self._op = _op

    @property
    def child(self):
        return self._child

    @property
    def op(self):
        return self._op

    def children(self):
        return [self._child]

    def clone(self):
        if type(self) is AbstractUnaryOpNode:
            return AbstractUnaryOpNode(self._child.clone(), self._op)
        return type(self)(self._child.clone())

    def macro_expand(self):
        self._child = self._child.macro_expand()
        return self

    def __str__(self):
        return self._op + "(" + str(self._child) + ")"


class UnaryPlus(AbstractUnaryOpNode):
    def __init__(self, _child):
        AbstractUnaryOpNode.__init__(self, _child, '+')


class UnaryMinus(AbstractUnaryOpNode):
    def __init__(self, _child):
        AbstractUnaryOpNode.__init__(self, _child, '-')


# ---------------------------------------------------------------------------
# Variables, parameters, references
# ---------------------------------------------------------------------------

class Variable(AbstractExpression):
    def __init__(self, _typ, _name):
        self._typ = _typ
        self._name = _name

    @property
    def name(self):
        return self._name

    @property
    def typ(self):
        return self._typ

    def clone(self):
        return self

    def __str__(self):
        return str(self._name)

    def __repr__(self):
        return 'Variable(%s)' % self._name


class Parameter(Variable):
    def __repr__(self):
        return 'Parameter(%s)' % self._name


class Reference(AbstractExpression):
    """``f(args)``: a read of a Function/Image/Reduction at a point."""

    def __init__(self, _obj, _args):
        _args = [_as_expr(a) for a in _args]
        for a in _args:
            assert isinstance(a, AbstractExpression), a
        self._obj = _obj
        self._args = list(_args)

    @property
    def objectRef(self):
        return self._obj

    @property
    def arguments(self):
        return self._args

    def _replace_ref_object(self, obj):
        self._obj = obj

    def children(self):
        return list(self._args)

    def clone(self):
        return Reference(self._obj, [a.clone() for a in self._args])

    def macro_expand(self):
        self._args = [a.macro_expand() for a in self._args]
        return self

    def __str__(self):
        return self._obj.name + "(" + ", ".join(str(a) for a in self._args) + ")"


# ---------------------------------------------------------------------------
# Conditions, select, cast, builtins
# ---------------------------------------------------------------------------

class Condition(object):
    """Comparison or conjunction/disjunction (constructs.py:722-803)."""

    _CMP = ['<', '<=', '>', '>=', '==', '!=']

    def __init__(self, _left, _cond, _right):
        _left = _as_expr(_left)
        _right = _as_expr(_right)
        assert _cond in self._CMP + ['&&', '||'], _cond
        if _cond in self._CMP:
            assert isinstance(_left, AbstractExpression)
            assert isinstance(_right, AbstractExpression)
        else:
            assert isinstance(_left, Condition)
            assert isinstance(_right, Condition)
        self._left = _left
        self._right = _right
        self._cond = _cond

    @property
    def lhs(self):
        return self._left

    @property
    def rhs(self):
        return self._right

    @property
    def conditional(self):
        return self._cond

    def clone(self):
        return Condition(self._left.clone(), self._cond, self._right.clone())

    def collect(self, objType):
        if type(self) is objType:
            return [self]
        return _unique(self._left.collect(objType) + self._right.collect(objType))

    def split_to_conjuncts(self):
        """Disjunctive normal form as a list of lists of comparisons;
        ``a != b`` splits into ``a < b`` or ``a > b`` (constructs.py:764-786)."""
        c = self._cond
        if c in ['<', '<=', '>', '>=', '==']:
            return [[self]]
        if c == '!=':
            return [[Condition(self._left, '<', self._right)],
                    [Condition(self._left, '>', self._right)]]
        if c == '||':
            return self._left.split_to_conjuncts() + self._right.split_to_conjuncts()
        if c == '&&':
            out = []
            for l in self._left.split_to_conjuncts():
                for r in self._right.split_to_conjuncts():
                    out.append(l + r)
            return out
        raise AssertionError(c)

    def macro_expand(self):
        if self._cond in self._CMP:
            self._left = self._left.macro_expand()
            self._right = self._right.macro_expand()
        else:
            self._left.macro_expand()
            self._right.macro_expand()
        return self

    def __and__(self, other):
        assert isinstance(other, Condition)
        return Condition(self, '&&', other)

    def __or__(self, other):
        assert isinstance(other, Condition)
        return Condition(self, '||', other)

    def __str__(self):
        return "(" + str(self._left) + " " + self._cond + " " + str(self._right) + ")"


class InbuiltFunction(AbstractExpression):
    _cname = None

    def __init__(self, *_args):
        _args = [_as_expr(a) for a in _args]
        for a in _args:
            assert isinstance(a, AbstractExpression)
        self._args = list(_args)

    @property
    def arguments(self):
        return self._args

    def children(self):
        return list(self._args)

    def clone(self):
        return type(self)(*[a.clone() for a in self._args])

    def macro_expand(self):
        self._args = [a.macro_expand() for a in self._args]
        return self

    def getType(self):
        raise TypeError(self)

    def __str__(self):
        return self._cname + "(" + ", ".join(str(a) for a in self._args) + ")"


def _builtin(name, cname, rtype, nargs):
    """Math builtin whose result type is fixed (constructs.py:77-197)."""
    def __init__(self, *args):
        assert len(args) == nargs, (name, args)
        InbuiltFunction.__init__(self, *args)

    def getType(self):
        return rtype
    return type(name, (InbuiltFunction,),
                {'__init__': __init__, 'getType': getType, '_cname': cname})


from .expr_types import Double  # noqa: E402

Pow = _builtin('Pow', 'pow', Double, 2)
Powf = _builtin('Powf', 'powf', Float, 2)
Log = _builtin('Log', 'log', Double, 1)
Exp = _builtin('Exp', 'std::exp', Double, 1)
Sin = _builtin('Sin', 'std::sin', Double, 1)
Cos = _builtin('Cos', 'std::cos', Double, 1)
Sqrt = _builtin('Sqrt', 'std::sqrt', Double, 1)
Sqrtf = _builtin('Sqrtf', 'std::sqrtf', Float, 1)


class RandomFloat(InbuiltFunction):
    """Pseudo-random float in [0,1] (constructs.py:121-133)."""
    _cname = 'rand'

    def __init__(self):
        InbuiltFunction.__init__(self)

    def getType(self):
        return Float

    def __str__(self):
        return "(static_cast<float> (rand()) / static_cast<float> (RAND_MAX))"


class Abs(InbuiltFunction):
    _cname = 'std::abs'

    def __init__(self, _expr):
        InbuiltFunction.__init__(self, _expr)

    def getType(self):
        from .expression import getType
        return getType(self._args[0])


class Cast(AbstractExpression):
    def __init__(self, _typ, _expr):
        _expr = _as_expr(_expr)
        from .expr_types import ALL_TYPES
        assert _typ in ALL_TYPES, _typ
        assert isinstance(_expr, AbstractExpression)
        self._typ = _typ
        self._expr = _expr

    @property
    def typ(self):
        return self._typ

    @property
    def expression(self):
        return self._expr

    def children(self):
        return [self._expr]

    def clone(self):
        return Cast(self._typ, self._expr.clone())

    def macro_expand(self):
        self._expr = self._expr.macro_expand()
        return self

    def __str__(self):
        return "(" + self._typ.c_name() + ") (" + str(self._expr) + ")"


class Select(AbstractExpression):
    """``cond ? t : f``.  Both branches are evaluated (codegen.py:618-642)."""

    def __init__(self, _cond, _true_expr, _false_expr, typeCheck=True):
        assert isinstance(_cond, Condition)
        _true_expr = _as_expr(_true_expr)
        _false_expr = _as_expr(_false_expr)
        assert isinstance(_true_expr, AbstractExpression)
        assert isinstance(_false_expr, AbstractExpression)
        if typeCheck:
            from .expression import getType
            tt, ft = getType(_true_expr), getType(_false_expr)
            assert tt == ft, str(_true_expr) + " == " + str(_false_expr)
        self._cond = _cond
        self._true_expr = _true_expr
        self._false_expr = _false_expr

    @property
    def condition(self):
        return self._cond

    @property
    def true_expression(self):
        return self._true_expr

    @property
    def false_expression(self):
        return self._false_expr

    def children(self):
        return [self._true_expr, self._false_expr]

    def collect(self, objType):
        objs = self._cond.collect(objType)
        objs += self._true_expr.collect(objType)
        objs += self._false_expr.collect(objType)
        if type(self) is objType:
            objs.append(self)
        return _unique(objs)

    def clone(self):
        return Select(self._cond.clone(), self._true_expr.clone(),
                      self._false_expr.clone(), False)

    def macro_expand(self):
        self._cond.macro_expand()
        self._true_expr = self._true_expr.macro_expand()
        self._false_expr = self._false_expr.macro_expand()
        return self

    def __str__(self):
        return "(" + str(self._cond) + "? " + str(self._true_expr) + ": " + \
            str(self._false_expr) + ")"


class Max(Select):
    def __init__(self, _leftExpr, _rightExpr, typeCheck=True):
        Select.__init__(self, Condition(_leftExpr, '>', _rightExpr),
                        _leftExpr, _rightExpr, typeCheck)

    def clone(self):
        return Max(self._true_expr.clone(), self._false_expr.clone(), False)


class Min(Select):
    def __init__(self, _leftExpr, _rightExpr, typeCheck=True):
        Select.__init__(self, Condition(_leftExpr, '<', _rightExpr),
                        _leftExpr, _rightExpr, typeCheck)

    def clone(self):
        return Min(self._true_expr.clone(), self._false_expr.clone(), False)
