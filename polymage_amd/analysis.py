"""Static analysis used by the planner and the emitters.

The reference does this with isl (sandbox/poly.py): every Case becomes a
"poly part" whose domain is the function domain intersected with the affine
conjuncts of its condition, non-affine conditions become predicates
(poly.py:657-742), and each reference gets a dependence vector
(poly.py:291-429) that is '*' when the access is not a uniform affine map.

Our domains are boxes with bounds affine in the parameters and our accesses
are monotone single-variable maps (``x-1``, ``2*x+1``, ``x/2``), so box and
interval arithmetic with C integer semantics replace the polyhedral library:

* ``ceval``            evaluates an integer expression with C semantics
                       (truncating ``/`` and ``%``, expr_ast / targetc A2);
* ``classify_access``  turns a reference argument into a monotone map of one
                       consumer variable, a constant, or '*' (data dependent);
* ``case_parts``       splits a definition into parts (box, predicate, expr);
* ``interval_image``   image of an interval under a monotone access.
"""

from fractions import Fraction
import math

from .expr_ast import (AbstractExpression, AbstractBinaryOpNode,
                       AbstractUnaryOpNode, Value, Variable, Parameter,
                       Reference, Condition, Select, Cast, InbuiltFunction)
from .expression import isAffine, get_affine_var_and_param_coeff
from .expr_types import Int, Rational


class PlanError(Exception):
    pass


# ---------------------------------------------------------------------------
# C-semantics integer evaluation
# ---------------------------------------------------------------------------

def c_div(a, b):
    """C99 integer division (truncation toward zero)."""
    if b == 0:
        raise ZeroDivisionError
    q = abs(a) // abs(b)
    return q if (a >= 0) == (b >= 0) else -q


def c_mod(a, b):
    return a - b * c_div(a, b)


def ceval(expr, env):
    """Evaluate an integer expression; ``env`` maps Variable/Parameter
    objects to ints.  Fractions (affine-analysis results) are floored."""
    expr = Value.numericToValue(expr)
    if isinstance(expr, Value):
        v = expr.value
        if isinstance(v, Fraction):
            return v
        return v
    if isinstance(expr, Variable):
        if expr not in env:
            raise PlanError("no value for %s" % expr.name)
        return env[expr]
    if isinstance(expr, AbstractBinaryOpNode):
        l = ceval(expr.left, env)
        r = ceval(expr.right, env)
        op = expr.op
        if isinstance(l, Fraction) or isinstance(r, Fraction) or \
                isinstance(l, float) or isinstance(r, float):
            if op == '+':
                return l + r
            if op == '-':
                return l - r
            if op == '*':
                return l * r
            if op == '/':
                return Fraction(l) / Fraction(r) if not isinstance(l, float) else l / r
            raise PlanError("non-integer operand to %s" % op)
        if op == '+':
            return l + r
        if op == '-':
            return l - r
        if op == '*':
            return l * r
        if op == '/':
            return c_div(l, r)
        if op == '%':
            return c_mod(l, r)
        if op == '<<':
            return l << r
        if op == '>>':
            return l >> r
        if op == '&':
            return l & r
        if op == '|':
            return l | r
        if op == '^':
            return l ^ r
        raise PlanError(op)
    if isinstance(expr, AbstractUnaryOpNode):
        v = ceval(expr.child, env)
        return -v if expr.op == '-' else v
    raise PlanError("cannot evaluate %s statically" % (expr,))


def ceval_np(expr, env):
    """``ceval`` over numpy int64 arrays (C truncating ``/`` and ``%``)."""
    import numpy as np
    expr = Value.numericToValue(expr)
    if isinstance(expr, Value):
        return expr.value
    if isinstance(expr, Variable):
        if expr not in env:
            raise PlanError("no value for %s" % expr.name)
        return env[expr]
    if isinstance(expr, AbstractBinaryOpNode):
        l = ceval_np(expr.left, env)
        r = ceval_np(expr.right, env)
        op = expr.op
        if op == '+':
            return l + r
        if op == '-':
            return l - r
        if op == '*':
            return l * r
        if op in ('/', '%'):
            l = np.asarray(l, dtype=np.int64)
            r = np.asarray(r, dtype=np.int64)
            q = np.abs(l) // np.abs(r)
            q = np.where((l >= 0) == (r >= 0), q, -q)
            return q if op == '/' else l - r * q
        raise PlanError(op)
    if isinstance(expr, AbstractUnaryOpNode):
        v = ceval_np(expr.child, env)
        return -v if expr.op == '-' else v
    raise PlanError("cannot evaluate %s statically" % (expr,))


def ceval_int(expr, env):
    v = ceval(expr, env)
    if isinstance(v, Fraction):
        if v.denominator != 1:
            raise PlanError("non-integral bound %s" % (v,))
        return int(v)
    return int(v)


# ---------------------------------------------------------------------------
# Access classification
# ---------------------------------------------------------------------------

class Access(object):
    """How one argument of a reference depends on the consumer's point.

    kind == 'const': ``expr`` depends on parameters only.
    kind == 'var':   ``expr`` is a monotone function of consumer variable
                     ``var`` (index ``dim``); ``slope`` is its rational
                     derivative (e.g. 2 for ``2*x+1``, 1/2 for ``x/2``) and
                     ``identity`` is True for exactly ``var``.
    kind == '*':     anything else (data dependent or non-monotone).
    """

    def __init__(self, kind, expr, var=None, dim=None, slope=None,
                 identity=False, increasing=True):
        self.kind = kind
        self.expr = expr
        self.var = var
        self.dim = dim
        self.slope = slope
        self.identity = identity
        self.increasing = increasing

    def __repr__(self):
        if self.kind == 'var':
            return 'Access(%s, slope=%s%s)' % (self.expr, self.slope,
                                                ', id' if self.identity else '')
        return 'Access(%s, %s)' % (self.kind, self.expr)


def _monotone(expr, var):
    """Return (ok, slope) if ``expr`` is built from ``var`` (appearing once),
    parameters and integer constants with + - * (by constant) and / by a
    positive constant; slope is the rational coefficient of ``var``."""
    expr = Value.numericToValue(expr)
    if isinstance(expr, Value):
        return (expr.typ in (Int, Rational)), Fraction(0)
    if isinstance(expr, Parameter):
        return True, Fraction(0)
    if isinstance(expr, Variable):
        return True, Fraction(1 if expr is var else 0)
    if isinstance(expr, AbstractUnaryOpNode):
        ok, s = _monotone(expr.child, var)
        return ok, (-s if expr.op == '-' else s)
    if isinstance(expr, AbstractBinaryOpNode):
        lok, ls = _monotone(expr.left, var)
        rok, rs = _monotone(expr.right, var)
        if not (lok and rok):
            return False, None
        if expr.op in ('+', '-'):
            if ls != 0 and rs != 0:
                return False, None
            return True, (ls + rs if expr.op == '+' else ls - rs)
        if expr.op == '*':
            if ls != 0 and rs != 0:
                return False, None
            if ls != 0:
                if expr.right.has(Variable) or expr.right.has(Parameter):
                    return False, None
                c = ceval(expr.right, {})
                return True, ls * c
            if rs != 0:
                if expr.left.has(Variable) or expr.left.has(Parameter):
                    return False, None
                c = ceval(expr.left, {})
                return True, rs * c
            return True, Fraction(0)
        if expr.op == '/':
            if rs != 0:
                return False, None
            if expr.right.has(Variable) or expr.right.has(Parameter):
                return False, None
            d = ceval(expr.right, {})
            if not isinstance(d, int) or d <= 0:
                return False, None
            return True, ls / d
        return False, None
    return False, None


def classify_access(arg, consumer_vars):
    """Classify one reference argument against the consumer variables."""
    arg = Value.numericToValue(arg)
    used = [v for v in consumer_vars if arg.has(Variable) and v in arg.collect(Variable)]
    if isinstance(arg, Variable) and arg in consumer_vars:
        return Access('var', arg, var=arg, dim=consumer_vars.index(arg),
                      slope=Fraction(1), identity=True)
    if arg.has(Reference) or arg.has(Select) or arg.has(Cast) or \
            arg.has(InbuiltFunction):
        return Access('*', arg)
    if len(used) == 0:
        if all(isinstance(v, Parameter) for v in arg.collect(Variable)) and \
                not any(type(v) is Variable for v in arg.collect(Variable)):
            ok, _ = _monotone(arg, None)
            if ok:
                return Access('const', arg)
        return Access('*', arg)
    if len(used) > 1:
        return Access('*', arg)
    v = used[0]
    ok, slope = _monotone(arg, v)
    if not ok or slope == 0:
        return Access('*', arg)
    return Access('var', arg, var=v, dim=consumer_vars.index(v), slope=slope,
                  identity=False, increasing=(slope > 0))


_NON_SHIFT_OPS = ('/', '%', '<<', '>>', '&', '|', '^')


def shift_access(arg, consumer_vars):
    """``(k, off)`` when ``arg`` is exactly ``consumer_vars[k] + off`` with
    ``off`` built from parameters and integer constants by + - * only (no
    truncating operator, so the identity holds in C integer arithmetic);
    ``('const', arg)`` for a parameter/constant argument; else None.
    ``off`` is an int when it is a plain constant."""
    from .expression import substitute_vars, is_constant_expr, get_constant_from_expr
    arg = Value.numericToValue(arg)
    if arg.has(Reference) or arg.has(Select) or arg.has(Cast) or arg.has(InbuiltFunction):
        return None
    stack = [arg]
    while stack:
        e = stack.pop()
        if isinstance(e, AbstractBinaryOpNode) and e.op in _NON_SHIFT_OPS:
            return None
        if hasattr(e, 'children'):
            stack.extend(e.children())
    allv = arg.collect(Variable) if arg.has(Variable) else []
    used = [v for v in consumer_vars if v in allv]
    others = [v for v in allv if v not in consumer_vars and not isinstance(v, Parameter)]
    if others or len(used) > 1:
        return None
    if not used:
        return ('const', arg)
    v = used[0]
    if not isAffine(arg, include_div=False):
        return None
    coeff = get_affine_var_and_param_coeff(arg)
    if coeff.get(v) != 1:
        return None
    off = substitute_vars(arg, {v: Value(0, Int)})
    off = Value.numericToValue(off)
    if is_constant_expr(off):
        c = get_constant_from_expr(off)
        if c == int(c):
            off = int(c)
    return (consumer_vars.index(v), off)


def scaled_access(arg, consumer_vars):
    """``(k, a, off)`` when ``arg`` is exactly ``a * consumer_vars[k] + off``
    with integers a >= 1 and off (no truncating operator); ``('const', 0,
    arg)`` for a parameter/constant argument; else None.  The shifts of
    shift_access are the case a = 1 (with an integer offset); a = 2 is a
    downsampling read (``f(2 * x + 1)``)."""
    from .expression import substitute_vars, is_constant_expr, get_constant_from_expr
    arg = Value.numericToValue(arg)
    if arg.has(Reference) or arg.has(Select) or arg.has(Cast) or arg.has(InbuiltFunction):
        return None
    stack = [arg]
    while stack:
        e = stack.pop()
        if isinstance(e, AbstractBinaryOpNode) and e.op in _NON_SHIFT_OPS:
            return None
        if hasattr(e, 'children'):
            stack.extend(e.children())
    allv = arg.collect(Variable) if arg.has(Variable) else []
    used = [v for v in consumer_vars if v in allv]
    others = [v for v in allv if v not in consumer_vars and not isinstance(v, Parameter)]
    if others or len(used) > 1:
        return None
    if not used:
        return ('const', 0, arg)
    v = used[0]
    if not isAffine(arg, include_div=False):
        return None
    coeff = get_affine_var_and_param_coeff(arg)
    a = coeff.get(v)
    if a is None or a != int(a) or int(a) < 1 or any(
            k is not v and k != 1 and c for k, c in coeff.items()):
        return None
    off = Value.numericToValue(substitute_vars(arg, {v: Value(0, Int)}))
    if not is_constant_expr(off):
        return None
    c = get_constant_from_expr(off)
    if c != int(c):
        return None
    return (consumer_vars.index(v), int(a), int(c))


def interval_image(acc, lo, hi, env):
    """Exact [min, max] of a monotone access over consumer values [lo, hi]
    (other variables bound in env)."""
    e1 = dict(env)
    e2 = dict(env)
    e1[acc.var] = lo
    e2[acc.var] = hi
    a = ceval_int(acc.expr, e1)
    b = ceval_int(acc.expr, e2)
    return (a, b) if a <= b else (b, a)


# ---------------------------------------------------------------------------
# Case parts
# ---------------------------------------------------------------------------

class Part(object):
    """One piece of a definition: ``expr`` applies on the points of the
    function domain inside ``box`` (per-dimension (lo, hi) bound expressions
    affine in parameters, None = unbounded) where ``pred`` (a list of
    Conditions that could not be made into box bounds) holds."""

    def __init__(self, box, pred, expr, case_index):
        self.box = box
        self.pred = pred
        self.expr = expr
        self.case_index = case_index

    def __repr__(self):
        return 'Part(box=%s, pred=%s, expr=%s)' % (
            [(str(l), str(h)) for (l, h) in self.box],
            [str(p) for p in self.pred], self.expr)


def _linear_in(expr, var):
    """Return (True, a) when expr == a*var + g(parameters) syntactically,
    with ``var`` not under a division, modulo or non-affine operator."""
    expr = Value.numericToValue(expr)
    if isinstance(expr, Value):
        return expr.typ in (Int, Rational), 0
    if isinstance(expr, Variable):
        if expr is var:
            return True, 1
        return isinstance(expr, Parameter), 0
    if isinstance(expr, AbstractUnaryOpNode):
        ok, a = _linear_in(expr.child, var)
        return ok, (-a if expr.op == '-' else a)
    if isinstance(expr, AbstractBinaryOpNode):
        lok, la = _linear_in(expr.left, var)
        rok, ra = _linear_in(expr.right, var)
        if not (lok and rok):
            return False, 0
        if expr.op == '+':
            return True, la + ra
        if expr.op == '-':
            return True, la - ra
        if expr.op == '*':
            if la != 0 and ra != 0:
                return False, 0
            if la != 0:
                if expr.right.has(Variable):
                    return False, 0
                return True, la * ceval_int(expr.right, {})
            if ra != 0:
                if expr.left.has(Variable):
                    return False, 0
                return True, ra * ceval_int(expr.left, {})
            return True, 0
        if la == 0 and ra == 0 and expr.op in ('/', '%'):
            return True, 0
        return False, 0
    return False, 0


def _bound_from_cmp(cond, var):
    """If ``cond`` is ``k*var + g OP h`` with k = +-1 and g, h parameter
    expressions, return ('lo'|'hi'|'eq', bound expression) else None."""
    from .expression import substitute_vars
    lhs, rhs, op = cond.lhs, cond.rhs, cond.conditional
    lok, la = _linear_in(lhs, var)
    rok, ra = _linear_in(rhs, var)
    if not (lok and rok):
        return None
    k = la - ra
    if k not in (1, -1):
        return None
    others = [v for v in (lhs.collect(Variable) + rhs.collect(Variable))
              if type(v) is Variable and v is not var]
    if others:
        return None
    L0 = substitute_vars(lhs, {var: Value(0, Int)})
    R0 = substitute_vars(rhs, {var: Value(0, Int)})
    # k*var OP B  with B = R0 - L0
    B = Sub_(R0, L0)
    if k == -1:
        # -var OP B  <=>  var OP' -B
        B = Neg_(B)
        op = {'<': '>', '<=': '>=', '>': '<', '>=': '<=', '==': '=='}[op]
    if op == '==':
        return ('eq', B)
    if op == '<=':
        return ('hi', B)
    if op == '<':
        return ('hi', Sub_(B, Value(1, Int)))
    if op == '>=':
        return ('lo', B)
    if op == '>':
        return ('lo', Add_(B, Value(1, Int)))
    return None


def Sub_(a, b):
    from .expr_ast import Sub
    if isinstance(b, Value) and b.value == 0:
        return a
    return Sub(a, b)


def Add_(a, b):
    from .expr_ast import Add
    if isinstance(b, Value) and b.value == 0:
        return a
    return Add(a, b)


def Neg_(a):
    from .expr_ast import UnaryMinus
    return UnaryMinus(a)


def _is_param_only(e):
    e = Value.numericToValue(e)
    return not any(type(v) is Variable for v in e.collect(Variable))


def case_parts(func, variables=None):
    """Split the definition of a (non-reduction) Function into Parts.

    Conjunct comparisons of the form ``a*x + params OP params`` become box
    bounds on x; everything else stays a runtime predicate.  Reference:
    poly.py:657-742 (affine conjuncts into the domain, non-affine
    conditions as predicates, ``!=`` and ``||`` split into several parts).
    """
    from .constructs import Case
    if variables is None:
        variables = func.variables
    nd = len(variables)
    parts = []
    for ci, case in enumerate(func.defn):
        if isinstance(case, Case):
            for conj in case.condition.split_to_conjuncts():
                box = [[None, None] for _ in range(nd)]
                pred = []
                for cond in conj:
                    handled = False
                    for d, v in enumerate(variables):
                        if cond.lhs.has(Variable) or cond.rhs.has(Variable):
                            if v in cond.lhs.collect(Variable) or v in cond.rhs.collect(Variable):
                                r = _bound_from_cmp(cond, v)
                                if r is not None:
                                    kind, b = r
                                    if kind in ('lo', 'eq'):
                                        box[d][0] = b if box[d][0] is None else _Max2(box[d][0], b)
                                    if kind in ('hi', 'eq'):
                                        box[d][1] = b if box[d][1] is None else _Min2(box[d][1], b)
                                    handled = True
                                break
                    if not handled:
                        pred.append(cond)
                parts.append(Part([tuple(b) for b in box], pred, case.expression, ci))
        else:
            parts.append(Part([(None, None)] * nd, [], case, ci))
    return parts


class _Max2(AbstractExpression):
    def __init__(self, a, b):
        self.a, self.b = Value.numericToValue(a), Value.numericToValue(b)

    def children(self):
        return [self.a, self.b]

    def __str__(self):
        return 'pmx_max(%s, %s)' % (self.a, self.b)


class _Min2(AbstractExpression):
    def __init__(self, a, b):
        self.a, self.b = Value.numericToValue(a), Value.numericToValue(b)

    def children(self):
        return [self.a, self.b]

    def __str__(self):
        return 'pmx_min(%s, %s)' % (self.a, self.b)


Max2 = _Max2
Min2 = _Min2


def eval_bexpr(expr, env):
    """Evaluate a bound expression built from parameter expressions, Max2
    and Min2."""
    if isinstance(expr, _Max2):
        return max(eval_bexpr(expr.a, env), eval_bexpr(expr.b, env))
    if isinstance(expr, _Min2):
        return min(eval_bexpr(expr.a, env), eval_bexpr(expr.b, env))
    return ceval_int(expr, env)


def refs_in(expr):
    """All Reference nodes of an expression (including inside Select
    conditions), in a deterministic order."""
    from .constructs import Reduce
    if isinstance(expr, Reduce):
        return expr.accumulate_ref.collect(Reference)[:-1] + \
            [r for r in expr.expression.collect(Reference)]
    return expr.collect(Reference)


def pred_refs(pred):
    out = []
    for c in pred:
        out += c.collect(Reference)
    return out


def unconditional_refs(expr):
    """References evaluated whatever the data: those not inside a Select
    branch (Select conditions and everything outside Selects count)."""
    from .expr_ast import (AbstractBinaryOpNode, AbstractUnaryOpNode, Cast,
                           InbuiltFunction)
    out = []

    def walk(e):
        e = Value.numericToValue(e)
        if isinstance(e, Reference):
            out.append(e)
            for a in e.arguments:
                walk(a)
        elif isinstance(e, Select):
            walk_cond(e.condition)
        elif isinstance(e, (AbstractBinaryOpNode, AbstractUnaryOpNode, Cast, InbuiltFunction)):
            for ch in e.children():
                walk(ch)

    def walk_cond(c):
        if c.conditional in ('&&', '||'):
            walk_cond(c.lhs)
            walk_cond(c.rhs)
        else:
            walk(c.lhs)
            walk(c.rhs)
    walk(expr)
    return out
