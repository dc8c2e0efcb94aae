"""Front-end constructs of the PolyMage DSL.

API-compatible restatement of sandbox/constructs.py: Interval (347-378),
Stencil (544-687), Case (806-840), Op/Reduce (842-904), Function
(907-1064), Image (1066-1090), Reduction (1092-1214).  Application files
written for the reference (``from constructs import *``) run unchanged
against this module through the compat package.
"""

import math

from .expr_types import *  # noqa: F401,F403
from .expr_ast import *  # noqa: F401,F403
from .expr_ast import (AbstractExpression, Value, Variable, Parameter,
                       Reference, Condition, _unique)
from .expression import *  # noqa: F401,F403
from .expression import isAffine, substitute_refs


class Interval(object):
    def __init__(self, _typ, _lb, _ub):
        _lb = Value.numericToValue(_lb)
        _ub = Value.numericToValue(_ub)
        assert isinstance(_lb, AbstractExpression)
        assert isinstance(_ub, AbstractExpression)
        self._typ = _typ
        self._lb = _lb
        self._ub = _ub

    @property
    def lowerBound(self):
        return self._lb

    @property
    def upperBound(self):
        return self._ub

    @property
    def typ(self):
        return self._typ

    def collect(self, objType):
        if type(self) is objType:
            return [self]
        return _unique(self._lb.collect(objType) + self._ub.collect(objType))

    def clone(self):
        return Interval(self._typ, self._lb.clone(), self._ub.clone())

    def __str__(self):
        return '(' + str(self._lb) + ', ' + str(self._ub) + ')'


class Case(object):
    """A guarded definition piece; points that no Case covers stay
    unwritten (poly.py:657-742)."""

    def __init__(self, _cond, _expr):
        _expr = Value.numericToValue(_expr)
        assert isinstance(_cond, Condition)
        assert isinstance(_expr, (AbstractExpression, Reduce))
        self._cond = _cond
        if isinstance(_expr, AbstractExpression):
            self._expr = _expr.macro_expand()
        else:
            self._expr = _expr

    @property
    def condition(self):
        return self._cond

    @property
    def expression(self):
        return self._expr

    def collect(self, objType):
        if type(self) is objType:
            return [self]
        return _unique(self._cond.collect(objType) + self._expr.collect(objType))

    def clone(self):
        return Case(self._cond.clone(), self._expr.clone())

    def __str__(self):
        return 'Case(' + str(self._cond) + ')' + '{ ' + str(self._expr) + ' }'


class Op(object):
    Sum = 0
    Mul = 1
    Min = 2
    Max = 3


class Reduce(object):
    """``red_ref op= expr`` inside a Reduction definition."""

    def __init__(self, _red_ref, _expr, _op_typ):
        assert isinstance(_red_ref, Reference)
        assert isinstance(_red_ref.objectRef, Reduction)
        _expr = Value.numericToValue(_expr)
        assert isinstance(_expr, AbstractExpression)
        assert _op_typ in [Op.Sum, Op.Mul, Op.Min, Op.Max]
        self._red_ref = _red_ref
        self._expr = _expr
        self._op_typ = _op_typ

    @property
    def accumulate_ref(self):
        return self._red_ref

    @property
    def expression(self):
        return self._expr

    @property
    def op_type(self):
        # The reference returns the expression here (constructs.py:866-868),
        # and its code generator then treats every reduction as a sum
        # (codegen.py:282).  We return the operator and honour it.
        return self._op_typ

    def collect(self, objType):
        if type(self) is objType:
            return [self]
        return _unique(self._red_ref.collect(objType) + self._expr.collect(objType))

    def clone(self):
        return Reduce(self._red_ref.clone(), self._expr.clone(), self._op_typ)

    def macro_expand(self):
        self._expr = self._expr.macro_expand()
        return self

    def __str__(self):
        op_str, sep = {Op.Sum: ('', ' + '), Op.Mul: ('', ' * '),
                       Op.Min: ('Min', ', '), Op.Max: ('Max', ', ')}[self._op_typ]
        r = str(self._red_ref)
        return 'Reduce [ ' + r + ' = ' + op_str + '(' + r + sep + str(self._expr) + ') ]'


class Function(object):
    def __init__(self, _varDom, _typ, _name, _const=""):
        self._name = _name
        self._typ = _typ
        self._const = (_const == "const")
        assert len(_varDom[0]) == len(_varDom[1])
        for v, d in zip(_varDom[0], _varDom[1]):
            assert isinstance(v, Variable)
            assert isinstance(d, Interval)
            assert v.typ == d.typ
        self._variables = list(_varDom[0])
        self._varDomain = list(_varDom[1])
        self._ndims = len(self._variables)
        self._body = []

    @property
    def name(self):
        return self._name

    @property
    def typ(self):
        return self._typ

    @property
    def is_const_func(self):
        return self._const

    @property
    def variableDomain(self):
        return (self._variables, self._varDomain)

    @property
    def domain(self):
        return self._varDomain

    @property
    def variables(self):
        return self._variables

    @property
    def ndims(self):
        return self._ndims

    @property
    def defn(self):
        return self._body

    @defn.setter
    def defn(self, _def):
        assert self._body == []
        assert len(_def) > 0, str(_def) + " " + str(self._name)
        n_case = n_plain = 0
        for case in _def:
            case = Value.numericToValue(case)
            assert isinstance(case, (Case, AbstractExpression)), str(case)
            if isinstance(case, Case):
                n_case += 1
            else:
                n_plain += 1
                case = case.macro_expand()
            assert n_plain <= 1
            assert n_case * n_plain == 0
            self._body.append(case)

    def __call__(self, *args):
        assert len(args) == len(self._variables), (self._name, args)
        return Reference(self, args)

    def replace_refs(self, ref_to_expr_map):
        for i, case in enumerate(self._body):
            if isinstance(case, Case):
                self._body[i] = Case(case.condition,
                                     substitute_refs(case.expression, ref_to_expr_map))
            else:
                self._body[i] = substitute_refs(case, ref_to_expr_map)

    def getObjects(self, objType):
        objs = []
        for case in self._body:
            objs += case.collect(objType)
        for interval in self._varDomain:
            objs += interval.collect(objType)
        return _unique(objs)

    def hasBoundedIntegerDomain(self):
        for d in self._varDomain:
            if not (isAffine(d.lowerBound) and isAffine(d.upperBound)):
                return False
        return True

    def clone(self):
        f = Function(([v.clone() for v in self._variables],
                      [d.clone() for d in self._varDomain]),
                     self._typ, self._name, "const" if self._const else "")
        f.defn = [c.clone() for c in self._body]
        return f

    def __str__(self):
        if not self._body:
            return self._name
        var_str = ", ".join(str(v) for v in self._variables)
        dom_str = ', '.join(str(v) + str(d) for v, d in
                            zip(self._variables, self._varDomain))
        case_str = "{ " + "\n ".join(str(c) for c in self._body) + " }"
        return "Domain: " + dom_str + '\n' + self._name + "(" + var_str + \
            ") = " + case_str + '\n'


class Image(Function):
    """A pipeline input of the given extents (constructs.py:1066-1090)."""

    def __init__(self, _typ, _name, _dims):
        _dims = [Value.numericToValue(d) for d in _dims]
        for d in _dims:
            assert isinstance(d, AbstractExpression)
        self._dims = _dims
        intervals = [Interval(UInt, 0, d - 1) for d in _dims]
        variables = [Variable(UInt, "_" + _name + str(i)) for i in range(len(_dims))]
        Function.__init__(self, (variables, intervals), _typ, _name)

    @property
    def dimensions(self):
        return tuple(self._dims)

    def clone(self):
        return self

    def __str__(self):
        return str(self._name) + "(" + ", ".join(str(d) for d in self._dims) + ")"


class Reduction(Function):
    def __init__(self, _varDom, _redDom, _typ, _name):
        Function.__init__(self, _varDom, _typ, _name)
        assert len(_redDom[0]) == len(_redDom[1])
        for v, d in zip(_redDom[0], _redDom[1]):
            assert isinstance(v, Variable)
            assert isinstance(d, Interval)
            assert v.typ == d.typ
        self._redVariables = list(_redDom[0])
        self._redDomain = list(_redDom[1])
        self._default = Value(0, _typ)

    @property
    def default(self):
        return self._default

    @default.setter
    def default(self, _expr):
        _expr = Value.numericToValue(_expr)
        assert isinstance(_expr, AbstractExpression)
        self._default = _expr

    @property
    def reductionDomain(self):
        return self._redDomain

    @property
    def reductionVariables(self):
        return self._redVariables

    @property
    def defn(self):
        return self._body

    @defn.setter
    def defn(self, _def):
        assert self._body == []
        bound = set(self._redVariables)
        for case in _def:
            case = Value.numericToValue(case)
            assert isinstance(case, (Case, Reduce)), str(case)
            # the accumulation runs over the reduction domain only, so in
            # the Reduce (target and value) a variable that is not a
            # reduction variable has no value (max_filter's vert_log,
            # polymage_maxfilter.py:60-66, reduces into vert_log(x, y, c, t)
            # over (x, rx, c, ry)); a Case condition may name the function's
            # own variables, which select target points (the bilateral grid's
            # c == 0 / c == 1)
            red = case.expression if isinstance(case, Case) else case
            free = [v for v in red.collect(Variable) if v not in bound]
            if free:
                raise TypeError("Reduction %s: %s not among its reduction variables %s"
                                % (self.name, ', '.join(sorted(v.name for v in free)),
                                   [v.name for v in self._redVariables]))
            self._body.append(case)

    def hasBoundedIntegerDomain(self):
        for d in list(self._varDomain) + list(self._redDomain):
            if not (isAffine(d.lowerBound) and isAffine(d.upperBound)):
                return False
        return True

    def getObjects(self, objType):
        objs = Function.getObjects(self, objType)
        for interval in self._redDomain:
            objs += interval.collect(objType)
        objs += self._default.collect(objType)
        return _unique(objs)

    def replace_refs(self, ref_to_expr_map):
        raise NotImplementedError("inlining into a Reduction is not supported")

    def clone(self):
        r = Reduction(([v.clone() for v in self._variables],
                       [d.clone() for d in self._varDomain]),
                      ([v.clone() for v in self._redVariables],
                       [d.clone() for d in self._redDomain]),
                      self._typ, self._name)
        r.defn = [c.clone() for c in self._body]
        r.default = self._default.clone()
        return r

    def __str__(self):
        if not self._body:
            return self._name
        var_str = ", ".join(str(v) for v in self._variables)
        dom_str = ', '.join(str(v) + str(d) for v, d in
                            zip(self._variables, self._varDomain))
        red_str = ', '.join(str(v) + str(d) for v, d in
                            zip(self._redVariables, self._redDomain))
        case_str = "{ " + "\n ".join(str(c) for c in self._body) + " }"
        return "Domain: " + dom_str + '\n' + "Reduction Domain: " + red_str + \
            '\n' + self._name + "(" + var_str + ") = " + case_str + '\n' + \
            "Default: " + str(self._default)


# ---------------------------------------------------------------------------
# Stencil sugar (constructs.py:423-687)
# ---------------------------------------------------------------------------

def get_inner_dimensions(lst):
    if isinstance(lst, list):
        return [len(x) if isinstance(x, list) else 0 for x in lst]
    return 0


def list_elements_equal(lst):
    return not lst or lst.count(lst[0]) == len(lst)


def _kernel_depth_ok(kernel, ndims):
    if isinstance(kernel, list):
        assert list_elements_equal(get_inner_dimensions(kernel)), \
            "kernel does not have equal dimensions: %s" % (kernel,)
        return all(_kernel_depth_ok(k, ndims - 1) for k in kernel)
    assert ndims == 0, ("kernel has %s dimensions than expected" %
                        ("more" if ndims < 0 else "fewer"))
    return True


def is_valid_kernel(kernel, num_dimensions):
    return _kernel_depth_ok(kernel, num_dimensions)


def get_valid_kernel_sizes(kernel):
    if isinstance(kernel, list):
        if len(kernel) == 0:
            return [0]
        return [len(kernel)] + get_valid_kernel_sizes(kernel[0])
    return []


def check_type(given, expected):
    if not isinstance(given, expected):
        raise TypeError("Expected %s to be of type %s, got %s" %
                        (given, expected, type(given)))


class Stencil(AbstractExpression):
    """Weighted neighbourhood sum over a Function, expanded into
    ``sum(ref * weight)`` when used in a definition."""

    def __init__(self, _input_fn, _iteration_vars, _kernel, _origin=None):
        check_type(_input_fn, Function)
        for v in _iteration_vars:
            check_type(v, Variable)
        assert is_valid_kernel(_kernel, len(_iteration_vars))
        self._input_fn = _input_fn
        self._iteration_vars = list(_iteration_vars)
        self._kernel = _kernel
        self._sizes = get_valid_kernel_sizes(_kernel)
        self._origin = _origin
        if self._origin is None:
            self._origin = [(s - 1) // 2 for s in self._sizes]

    @property
    def input_func(self):
        return self._input_fn

    @property
    def iter_vars(self):
        return self._iteration_vars

    @property
    def kernel(self):
        return self._kernel

    @property
    def sizes(self):
        return self._sizes

    @property
    def origin(self):
        return self._origin

    @staticmethod
    def _build_indexed_kernel(origin, iter_vars, kernel):
        """[(index expressions, weight)] in row-major kernel order."""
        assert is_valid_kernel(kernel, len(iter_vars))
        out = []

        def walk(sub, depth, idx):
            if depth == len(iter_vars):
                out.append((idx, sub))
                return
            for i in range(len(sub)):
                walk(sub[i], depth + 1,
                     idx + [iter_vars[depth] + (i - origin[depth])])
        walk(kernel, 0, [])
        return out

    def macro_expand(self):
        expr = 0
        for idx, w in self._build_indexed_kernel(self._origin,
                                                 self._iteration_vars,
                                                 self._kernel):
            expr += Reference(self._input_fn, idx) * w
        return 1 * expr

    def __str__(self):
        return ("Stencil object\n%s\n\tinput: %s\n\titeration vars: %s"
                "\n\tdimension sizes: %s\n\torigin: %s\n\tkernel: %s" %
                (str(Stencil.macro_expand(self)), self._input_fn,
                 [str(v) for v in self._iteration_vars], self._sizes,
                 self._origin, self._kernel))


class TStencil(object):
    """Time-iterated stencil declaration.  The reference leaves it
    unfinished (tests/test_t_stencil.py:11 is disabled); kept for API
    compatibility only."""

    def __init__(self, _var_domain, _kernel, _name, _origin=None, _timesteps=1):
        self.name = _name
        self.var_domain = _var_domain
        self.timesteps = int(_timesteps)
        assert is_valid_kernel(_kernel, len(_var_domain))
        self.size = get_valid_kernel_sizes(_kernel)
        self.kernel = _kernel
        self.origin = _origin if _origin is not None else \
            [int(math.floor(s / 2)) for s in self.size]
