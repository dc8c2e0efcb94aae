"""Expression analysis: types, affinity, affine coefficients, substitution.

Restates sandbox/expression.py (getType 355-379, isAffine 283-353,
get_affine_var_and_param_coeff 29-79, get_constant_from_expr 114-135,
is_constant_expr 137-156, simplify_expr 158-199, substitute_refs 201-238,
substitute_vars 240-281).  Constant folding uses Python semantics
(``Fraction`` for affine division) exactly where the reference does:
only on index and bound constants, never on emitted arithmetic.
"""

from fractions import Fraction

from .expr_ast import (AbstractExpression, AbstractBinaryOpNode,
                       AbstractUnaryOpNode, Value, Variable, Parameter,
                       Reference, Condition, Select, Cast, InbuiltFunction,
                       Add, Sub, Mul, Div, Mod, UnaryMinus)
from .expr_types import Int, Rational, result_type


def getType(expr):
    expr = Value.numericToValue(expr)
    assert isinstance(expr, AbstractExpression), expr
    if isinstance(expr, Value):
        return expr.typ
    if isinstance(expr, Variable):
        return expr.typ
    if isinstance(expr, Reference):
        return expr.objectRef.typ
    if isinstance(expr, AbstractBinaryOpNode):
        return result_type(getType(expr.left), getType(expr.right))
    if isinstance(expr, AbstractUnaryOpNode):
        return getType(expr.child)
    if isinstance(expr, Cast):
        return expr.typ
    if isinstance(expr, Select):
        t, f = getType(expr.true_expression), getType(expr.false_expression)
        assert t == f
        return t
    if isinstance(expr, InbuiltFunction):
        return expr.getType()
    raise TypeError(type(expr))


def _has_var_or_param(e):
    return e.has(Variable) or e.has(Parameter)


def isAffine(expr, include_div=True, include_modulo=False):
    expr = Value.numericToValue(expr)
    if isinstance(expr, Value):
        return expr.typ is Int or (include_div and expr.typ is Rational)
    if isinstance(expr, Variable):
        return True
    if isinstance(expr, Reference):
        return False
    if isinstance(expr, AbstractBinaryOpNode):
        if not (isAffine(expr.left, include_div, include_modulo) and
                isAffine(expr.right, include_div, include_modulo)):
            return False
        if expr.op in ['+', '-']:
            return True
        if expr.op == '*':
            return (not _has_var_or_param(expr.left)) or \
                   (not _has_var_or_param(expr.right))
        if include_div and expr.op == '/':
            return not _has_var_or_param(expr.right)
        if include_modulo and expr.op == '%':
            if not _has_var_or_param(expr.right):
                return isAffine(expr.left, include_div, False)
            return False
        return False
    if isinstance(expr, AbstractUnaryOpNode):
        return isAffine(expr.child, include_div, include_modulo)
    if isinstance(expr, Condition):
        return isAffine(expr.lhs, include_div, include_modulo) and \
               isAffine(expr.rhs, include_div, include_modulo)
    if isinstance(expr, (Select, Cast, InbuiltFunction)):
        return False
    raise TypeError(type(expr))


def _eval_binop(l, r, op, affine):
    if op == '/':
        return Fraction(l, r) if affine else l / r
    if op == '+':
        return l + r
    if op == '-':
        return l - r
    if op == '*':
        return l * r
    if op == '%':
        return l % r
    if op == '>>':
        return l >> r
    if op == '<<':
        return l << r
    if op == '|':
        return l | r
    if op == '^':
        return l ^ r
    if op == '&':
        return l & r
    raise TypeError(op)


def is_constant_expr(expr, affine=False):
    expr = Value.numericToValue(expr)
    if isinstance(expr, Value):
        return (not affine) or expr.typ in (Int, Rational)
    if isinstance(expr, AbstractBinaryOpNode):
        return is_constant_expr(expr.left, affine) and \
               is_constant_expr(expr.right, affine)
    if isinstance(expr, AbstractUnaryOpNode):
        return is_constant_expr(expr.child, affine)
    return False


def get_constant_from_expr(expr, affine=False):
    """Constant part of an expression: variables, references, selects,
    casts and builtins count as 0 (expression.py:114-135)."""
    expr = Value.numericToValue(expr)
    if isinstance(expr, Value):
        if affine:
            assert expr.typ in (Int, Rational)
        return expr.value
    if isinstance(expr, (Variable, Reference, Select, Cast, InbuiltFunction)):
        return 0
    if isinstance(expr, AbstractBinaryOpNode):
        return _eval_binop(get_constant_from_expr(expr.left, affine),
                           get_constant_from_expr(expr.right, affine),
                           expr.op, affine)
    if isinstance(expr, AbstractUnaryOpNode):
        v = get_constant_from_expr(expr.child, affine)
        return -v if expr.op == '-' else v
    raise TypeError(type(expr))


def get_affine_var_and_param_coeff(expr):
    """Map {Variable/Parameter: Fraction coefficient} of an affine expr."""
    expr = Value.numericToValue(expr)
    if not isAffine(expr) or isinstance(expr, Value) or is_constant_expr(expr):
        return {}
    if isinstance(expr, Variable):
        return {expr: 1}
    if isinstance(expr, AbstractBinaryOpNode):
        lc = get_affine_var_and_param_coeff(expr.left)
        rc = get_affine_var_and_param_coeff(expr.right)
        keys = list(lc) + [k for k in rc if k not in lc]
        if expr.op == '+':
            return dict((k, lc.get(k, 0) + rc.get(k, 0)) for k in keys)
        if expr.op == '-':
            return dict((k, lc.get(k, 0) - rc.get(k, 0)) for k in keys)
        if expr.op == '*':
            lconst = is_constant_expr(expr.left, affine=True)
            rconst = is_constant_expr(expr.right, affine=True)
            assert not (lconst and rconst)
            if lconst:
                c = get_constant_from_expr(expr.left, affine=True)
                return dict((k, c * v) for k, v in rc.items())
            if rconst:
                c = get_constant_from_expr(expr.right, affine=True)
                return dict((k, c * v) for k, v in lc.items())
            return {}
        if expr.op == '/':
            assert is_constant_expr(expr.right, affine=True)
            d = get_constant_from_expr(expr.right, affine=True)
            return dict((k, Fraction(1, 1) / d * v) for k, v in lc.items())
        return {}
    if isinstance(expr, AbstractUnaryOpNode):
        cc = get_affine_var_and_param_coeff(expr.child)
        if expr.op == '-':
            return dict((k, -v) for k, v in cc.items())
        return cc
    raise TypeError(type(expr))


def simplify_expr(expr):
    """Fold constant sub-trees of an (index) expression."""
    expr = Value.numericToValue(expr)
    if isinstance(expr, Value):
        return expr
    if is_constant_expr(expr, affine=True):
        v = get_constant_from_expr(expr, affine=True)
        if isinstance(v, Fraction) and v.denominator == 1:
            v = int(v)
        return Value(v, Int if isinstance(v, int) else Rational)
    if isinstance(expr, AbstractBinaryOpNode):
        return type(expr)(simplify_expr(expr.left), simplify_expr(expr.right)) \
            if type(expr) is not AbstractBinaryOpNode else expr
    if isinstance(expr, AbstractUnaryOpNode):
        return type(expr)(simplify_expr(expr.child))
    return expr


def substitute_vars(expr, var_to_expr_map):
    """Return a copy of ``expr`` with variables replaced."""
    expr = Value.numericToValue(expr)
    if isinstance(expr, Variable):
        return var_to_expr_map.get(expr, expr)
    if isinstance(expr, Value):
        return expr
    if isinstance(expr, AbstractBinaryOpNode):
        return _rebuild_bin(expr, substitute_vars(expr.left, var_to_expr_map),
                            substitute_vars(expr.right, var_to_expr_map))
    if isinstance(expr, AbstractUnaryOpNode):
        return type(expr)(substitute_vars(expr.child, var_to_expr_map))
    if isinstance(expr, Reference):
        return Reference(expr.objectRef,
                         [substitute_vars(a, var_to_expr_map) for a in expr.arguments])
    if isinstance(expr, Cast):
        return Cast(expr.typ, substitute_vars(expr.expression, var_to_expr_map))
    if isinstance(expr, Select):
        return Select(_subst_cond(expr.condition, lambda e: substitute_vars(e, var_to_expr_map)),
                      substitute_vars(expr.true_expression, var_to_expr_map),
                      substitute_vars(expr.false_expression, var_to_expr_map), False)
    if isinstance(expr, InbuiltFunction):
        return type(expr)(*[substitute_vars(a, var_to_expr_map) for a in expr.arguments])
    raise TypeError(type(expr))


def substitute_refs(expr, ref_to_expr_map):
    """Return a copy of ``expr`` with Reference objects (by identity)
    replaced by expressions."""
    expr = Value.numericToValue(expr)
    if isinstance(expr, Reference):
        if expr in ref_to_expr_map:
            return ref_to_expr_map[expr]
        return Reference(expr.objectRef,
                         [substitute_refs(a, ref_to_expr_map) for a in expr.arguments])
    if isinstance(expr, (Value, Variable)):
        return expr
    if isinstance(expr, AbstractBinaryOpNode):
        return _rebuild_bin(expr, substitute_refs(expr.left, ref_to_expr_map),
                            substitute_refs(expr.right, ref_to_expr_map))
    if isinstance(expr, AbstractUnaryOpNode):
        return type(expr)(substitute_refs(expr.child, ref_to_expr_map))
    if isinstance(expr, Cast):
        return Cast(expr.typ, substitute_refs(expr.expression, ref_to_expr_map))
    if isinstance(expr, Select):
        return Select(_subst_cond(expr.condition, lambda e: substitute_refs(e, ref_to_expr_map)),
                      substitute_refs(expr.true_expression, ref_to_expr_map),
                      substitute_refs(expr.false_expression, ref_to_expr_map), False)
    if isinstance(expr, InbuiltFunction):
        return type(expr)(*[substitute_refs(a, ref_to_expr_map) for a in expr.arguments])
    raise TypeError(type(expr))


def _rebuild_bin(node, l, r):
    if type(node) is AbstractBinaryOpNode:
        return AbstractBinaryOpNode(l, r, node.op)
    return type(node)(l, r)


def _subst_cond(cond, f):
    if cond.conditional in ('&&', '||'):
        return Condition(_subst_cond(cond.lhs, f), cond.conditional,
                         _subst_cond(cond.rhs, f))
    return Condition(f(cond.lhs), cond.conditional, f(cond.rhs))
