"""DSL expression -> C/HIP expression text with the reference's semantics.

The reference emits every DSL operator verbatim as a C operator
(codegen.py:574-581), so C's own rules (integer promotion, truncating
``/`` and ``%``, float*double -> double) decide the arithmetic.  We keep
that, and keep the three places where the reference adds its own rules:

  * literals: Python ``str`` with no suffix, i.e. C doubles
    (targetc.py:80-86); 'float12' mode prints ``%.12g`` + ``f`` like the
    reference's checked-in generated files;
  * ``Select`` (and Max/Min, which are Selects) computes both branches into
    temporaries whose C type is the branch's DSL type before selecting
    (codegen.py:618-642); inside ``Powf`` arguments no temporaries are made
    (codegen.py:654-657);
  * ``Cast(T, e)`` -> ``(T) (e)``.

Math builtins are printed through small device helpers that pin the C++
overload the reference's g++ build resolves to (``log``/``pow`` always in
double, ``std::exp``/``std::sqrt``/``std::abs`` by argument type).
"""

from .expr_ast import (AbstractBinaryOpNode, AbstractUnaryOpNode, Value,
                       Variable, Parameter, Reference, Condition, Select, Cast,
                       InbuiltFunction, Pow, Powf, Log, Exp, Sin, Cos, Sqrt,
                       Sqrtf, Abs, RandomFloat)
from .expr_types import Float
from .expression import getType


def c_type(t):
    return t.c_name()


# value ranges of the integer types narrow enough for an exhaustive check
_NARROW = {'unsigned char': (0, 255), 'char': (-128, 127), 'signed char': (-128, 127),
           'unsigned short': (0, 65535), 'short': (-32768, 32767)}
_DIVC = {}


def narrow_int_range(e):
    """(lo, hi) when expression e is a read of a function or image, or a
    cast, of an integer type of at most 16 bits (its C value lies in the
    type's range even after promotion), else None."""
    if isinstance(e, (Reference, Cast)):
        t = e.objectRef.typ if isinstance(e, Reference) else e.typ
        return _NARROW.get(t.c_name())
    return None


def divc_exact(d, lo, hi):
    """True when, for every integer n in [lo, hi], the double quotient
    n / d (IEEE division, round to nearest) equals q + (n - q*d) * r with
    r = RN(1/d), q = RN(n*r) and both corrections fused multiply-adds
    (one Newton step).  Checked exhaustively in exact rational arithmetic,
    so the rewrite is used only where it gives the division's bits."""
    from fractions import Fraction
    key = (d, lo, hi)
    if key not in _DIVC:
        ok = d != 0 and d == d and abs(d) != float('inf')
        if ok:
            r = 1.0 / d
            fd, fr = Fraction(d), Fraction(r)
            for n in range(lo, hi + 1):
                q = n * r
                e = float(n - Fraction(q) * fd)
                q1 = float(Fraction(e) * fr + Fraction(q))
                if q1 != n / d or (q1 == 0 and str(q1) != str(n / d)):
                    ok = False
                    break
        _DIVC[key] = ok
    return _DIVC[key]


# significand bits of every value of a type that converts to double exactly
# (float's 24 with its exponent range; integers of at most 16 bits)
_EXACT_BITS = dict([('float', 24)] + [(t, 16) for t in _NARROW])


def _exact_bits(e):
    """Significand bits of the C value of e when e is a read, cast or
    parameter of a type in _EXACT_BITS (the printed text then has that C
    type: function reads print as typed loads or typed variables)."""
    if isinstance(e, Reference):
        t = e.objectRef.typ
    elif isinstance(e, (Cast, Parameter)):
        t = e.typ
    else:
        return None
    return _EXACT_BITS.get(t.c_name())


def _const_bits(v):
    """Significand bits of a finite nonzero double constant whose binary
    exponent lies within +-800 (so its product with a float or a narrow
    integer neither overflows nor leaves double's normal range), else
    None."""
    import math
    if not isinstance(v, float) or v == 0 or v != v or abs(v) == float('inf'):
        return None
    if abs(math.frexp(v)[1]) > 800:
        return None
    num, _ = abs(v).as_integer_ratio()
    while num % 2 == 0:
        num //= 2
    return num.bit_length()


def exact_product(e):
    """(a, k) when e is a * k (either order) with k a double literal and
    a an expression whose value has few enough significand bits that a * k
    is exact in double (bits(a) + bits(k) <= 53), else None.  Then
    l + a * k rounds once, exactly as fma(a, k, l) does."""
    if not (isinstance(e, AbstractBinaryOpNode) and e.op == '*'):
        return None
    for a, c in ((e.left, e.right), (e.right, e.left)):
        cv = Value.numericToValue(c)
        if isinstance(cv, Value) and isinstance(cv.value, float):
            ab, cb = _exact_bits(a), _const_bits(cv.value)
            if ab is not None and cb is not None and ab + cb <= 53:
                return a, cv.value
    return None


def same_expr(a, b):
    """Structural equality of two expressions (or conditions): the same
    operators over the same variables, objects and typed constants."""
    a, b = Value.numericToValue(a), Value.numericToValue(b)
    if isinstance(a, Condition) or isinstance(b, Condition):
        return isinstance(a, Condition) and isinstance(b, Condition) and \
            a.conditional == b.conditional and same_expr(a.lhs, b.lhs) and same_expr(a.rhs, b.rhs)
    if type(a) is not type(b):
        return False
    if isinstance(a, Value):
        return a.typ == b.typ and repr(a.value) == repr(b.value)
    if isinstance(a, Variable):
        return a is b
    if isinstance(a, AbstractBinaryOpNode):
        return a.op == b.op and same_expr(a.left, b.left) and same_expr(a.right, b.right)
    if isinstance(a, AbstractUnaryOpNode):
        return a.op == b.op and same_expr(a.child, b.child)
    if isinstance(a, Reference):
        return a.objectRef is b.objectRef and len(a.arguments) == len(b.arguments) and \
            all(same_expr(x, y) for x, y in zip(a.arguments, b.arguments))
    if isinstance(a, Cast):
        return a.typ == b.typ and same_expr(a.expression, b.expression)
    if isinstance(a, Select):
        return same_expr(a.condition, b.condition) and \
            same_expr(a.true_expression, b.true_expression) and \
            same_expr(a.false_expression, b.false_expression)
    if isinstance(a, InbuiltFunction):
        return len(a.arguments) == len(b.arguments) and \
            all(same_expr(x, y) for x, y in zip(a.arguments, b.arguments))
    return False


_RANK = {'int': 1, 'unsigned int': 2, 'long long': 3, 'unsigned long long': 4,
         'float': 5, 'double': 6}


def _promote(t):
    return 'int' if t in _NARROW else t


class CSelect(Select):
    """``cond ? a : b`` of two expressions whose printed C text has the same
    type ``ctype``: printed as a plain conditional (only the taken branch
    is evaluated), so the value keeps that C type."""

    def __init__(self, cond, a, b, ctype):
        Select.__init__(self, cond, a, b, False)
        self.ctype = ctype


class SelectUnifier(object):
    """An expression equal, point by point, to ``cond ? a : b`` in which the
    selection is pushed down to where the two branches differ: a and b
    built by the same operators over the same C types share every node
    above their differences, so, e.g., the 0.25/0.75 upsampling of two
    neighbours computes its arithmetic once and reads one of the two
    differing neighbours instead of both.  The C type of each printed node
    is tracked (double literals make it differ from the DSL type), and a
    node is shared only where both branches' nodes have the same C type,
    so a point performs exactly the arithmetic of its taken branch.
    ``cheap_ref(r)``: a reference whose load is cheap (a constant offset
    from a per-item base) is kept whole, two loads selected, rather than
    selecting its index."""

    def __init__(self, mode, var_names, cheap_ref=None):
        self.mode = mode
        self.var_names = var_names
        self.cheap_ref = cheap_ref

    def ctype(self, e):
        """C type of e's printed text, or None when not tracked."""
        e = Value.numericToValue(e)
        if isinstance(e, CSelect):
            return e.ctype
        if isinstance(e, Value):
            if isinstance(e.value, float):
                return 'float' if (self.mode == 'float12' and e.typ is Float) else 'double'
            return 'int' if -2 ** 31 <= e.value < 2 ** 31 else None
        if isinstance(e, Parameter):
            return _promote(e.typ.c_name())
        if isinstance(e, Variable):
            return 'int' if e in self.var_names else None
        if isinstance(e, Reference):
            return _promote(e.objectRef.typ.c_name())
        if isinstance(e, Cast):
            return _promote(e.typ.c_name())
        if isinstance(e, Select):
            # printed with a cast to the select's type
            return _promote(c_type(getType(e)))
        if isinstance(e, AbstractUnaryOpNode):
            return self.ctype(e.child) if e.op in ('-', '+') else None
        if isinstance(e, AbstractBinaryOpNode):
            if e.op == '/' and self.mode == 'double' and isinstance(
                    Value.numericToValue(e.right), Value) and narrow_int_range(e.left):
                return None     # may print as pmx_divc
            l, r = self.ctype(e.left), self.ctype(e.right)
            if l is None or r is None or l not in _RANK or r not in _RANK:
                return None
            if e.op in ('<<', '>>'):
                return l
            return l if _RANK[l] >= _RANK[r] else r
        return None

    def index_only(self, u):
        """True when every selection left in u chooses between two pure
        arithmetic expressions (indices, constants: no reads, calls or
        selects), which compile to a conditional move; merging is then
        worth it.  Otherwise the two branches are kept."""
        stack = [u]
        while stack:
            e = Value.numericToValue(stack.pop())
            if isinstance(e, CSelect):
                for x in (e.true_expression, e.false_expression):
                    x = Value.numericToValue(x)
                    if x.has(Reference) or x.has(Select) or x.has(InbuiltFunction):
                        return False
                continue
            if isinstance(e, Select):
                stack += [e.true_expression, e.false_expression]
            elif hasattr(e, 'children'):
                stack += list(e.children())
        return True

    def unify(self, cond, a, b):
        """The unified expression, or None when a and b have different
        (or untracked) C types."""
        a, b = Value.numericToValue(a), Value.numericToValue(b)
        if same_expr(a, b):
            return a
        ta, tb = self.ctype(a), self.ctype(b)
        if ta is None or ta != tb:
            return None
        leaf = CSelect(cond, a, b, ta)
        if type(a) is not type(b):
            return leaf
        if isinstance(a, AbstractBinaryOpNode) and a.op == b.op:
            l = self.unify(cond, a.left, b.left)
            r = self.unify(cond, a.right, b.right)
            if l is None or r is None:
                return leaf
            if type(a) is AbstractBinaryOpNode:
                return AbstractBinaryOpNode(l, r, a.op)
            return type(a)(l, r)
        if isinstance(a, Cast) and a.typ == b.typ:
            x = self.unify(cond, a.expression, b.expression)
            return leaf if x is None else Cast(a.typ, x)
        if isinstance(a, Reference) and a.objectRef is b.objectRef and \
                len(a.arguments) == len(b.arguments):
            if self.cheap_ref is not None and self.cheap_ref(a) and self.cheap_ref(b):
                return leaf
            args = [self.unify(cond, x, y) for x, y in zip(a.arguments, b.arguments)]
            if any(x is None for x in args):
                return leaf
            r = Reference(a.objectRef, args)
            # the references it stands for (their per-reference decisions,
            # such as clamping, apply to it)
            r.unified_from = getattr(a, 'unified_from', (a,)) + getattr(b, 'unified_from', (b,))
            return r
        if type(a) is Select and same_expr(a.condition, b.condition):
            t = self.unify(cond, a.true_expression, b.true_expression)
            f = self.unify(cond, a.false_expression, b.false_expression)
            if t is None or f is None:
                return leaf
            # printed with a cast to the select's (shared) type
            return Select(a.condition, t, f, False)
        return leaf


def literal(v, mode):
    if isinstance(v.value, float):
        if mode == 'float12' and v.typ is Float:
            s = '%.12g' % v.value
            if 'e' not in s and '.' not in s:
                s += '.0'
            return s + 'f'
        s = repr(v.value)
        if s in ('inf', '-inf', 'nan'):
            raise ValueError("non-finite literal")
        return s
    return str(v.value)


class Printer(object):
    """Prints expressions; ``ref_cb(ref, printer, prologue)`` returns the C
    text of a Reference, ``var_names`` maps Variables to C names."""

    def __init__(self, var_names, ref_cb, mode='double', param_prefix='', lazy=False,
                 exact_fma=False, unify=None):
        self.var_names = var_names
        self.ref_cb = ref_cb
        self.mode = mode
        self.param_prefix = param_prefix
        self.lazy = lazy
        # l +- a * k as one fused multiply-add where the product is exact
        # (exact_product): the same bits as the separate multiply and add
        self.exact_fma = exact_fma and mode == 'double'
        # lazy Selects whose branches share structure print through
        # SelectUnifier when only index arithmetic is left to select;
        # ``unify`` is its cheap_ref test (None: off)
        self.unify = unify
        self._tmp = [0]

    def tmp(self):
        self._tmp[0] += 1
        return '_t%d' % self._tmp[0]

    def expr(self, e, pro=None):
        e = Value.numericToValue(e)
        if isinstance(e, Value):
            return literal(e, self.mode)
        if isinstance(e, Parameter):
            return self.param_prefix + e.name
        if isinstance(e, Variable):
            return self.var_names[e]
        if isinstance(e, AbstractBinaryOpNode):
            if e.op == '/' and self.mode == 'double':
                # a narrow integer over a double literal (u16 / 65535.0):
                # a multiply and two FMAs give the division's exact bits
                # (checked for every value of the integer type)
                d = Value.numericToValue(e.right)
                rng = narrow_int_range(e.left)
                if rng is not None and isinstance(d, Value) and isinstance(d.value, float) \
                        and divc_exact(d.value, rng[0], rng[1]):
                    return 'pmx_divc((double)(%s), %r, %r)' % (
                        self.expr(e.left, pro), 1.0 / d.value, d.value)
            if self.exact_fma and e.op in ('+', '-'):
                f = self._fma(e, pro)
                if f is not None:
                    return f
            return '(' + self.expr(e.left, pro) + ' ' + e.op + ' ' + \
                self.expr(e.right, pro) + ')'
        if isinstance(e, AbstractUnaryOpNode):
            return e.op + '(' + self.expr(e.child, pro) + ')'
        if isinstance(e, Reference):
            return self.ref_cb(e, self, pro)
        if isinstance(e, Select):
            if isinstance(e, CSelect):
                return '((%s) ? (%s) : (%s))' % (self.cond(e.condition, pro),
                                                 self.expr(e.true_expression, pro),
                                                 self.expr(e.false_expression, pro))
            if self.lazy and pro is not None and self.unify is not None and type(e) is Select:
                uf = SelectUnifier(self.mode, self.var_names, self.unify)
                u = uf.unify(e.condition, e.true_expression, e.false_expression)
                if u is not None and not isinstance(u, CSelect) and uf.index_only(u):
                    return '((%s)(%s))' % (c_type(getType(e)), self.expr(u, pro))
            if self.lazy and pro is not None:
                # Same value as the reference's "both branches into typed
                # temporaries, then select", but only the taken branch runs,
                # so an untaken branch never reads outside its producer.
                c = self.cond(e.condition, pro)
                t = self.expr(e.true_expression, pro)
                f = self.expr(e.false_expression, pro)
                return '((%s)((%s) ? (%s)(%s) : (%s)(%s)))' % (
                    c_type(getType(e)), c, c_type(getType(e.true_expression)), t,
                    c_type(getType(e.false_expression)), f)
            c = self.cond(e.condition, pro)
            t = self.expr(e.true_expression, pro)
            f = self.expr(e.false_expression, pro)
            if pro is None:
                return '(' + c + '? ' + t + ': ' + f + ')'
            a, b, s = self.tmp(), self.tmp(), self.tmp()
            pro.append('%s %s = %s;' % (c_type(getType(e.true_expression)), a, t))
            pro.append('%s %s = %s;' % (c_type(getType(e.false_expression)), b, f))
            pro.append('%s %s = (%s? %s: %s);' % (c_type(getType(e)), s, c, a, b))
            return s
        if isinstance(e, Cast):
            return '(' + c_type(e.typ) + ') (' + self.expr(e.expression, pro) + ')'
        if isinstance(e, Powf):
            return 'powf(' + self.expr(e.arguments[0], None) + ', ' + \
                self.expr(e.arguments[1], None) + ')'
        if isinstance(e, Pow):
            return 'pow((double)(' + self.expr(e.arguments[0], pro) + '), (double)(' + \
                self.expr(e.arguments[1], pro) + '))'
        if isinstance(e, Log):
            return 'log((double)(' + self.expr(e.arguments[0], pro) + '))'
        if isinstance(e, Exp):
            return 'pmx_exp(' + self.expr(e.arguments[0], pro) + ')'
        if isinstance(e, Sin):
            return 'pmx_sin(' + self.expr(e.arguments[0], pro) + ')'
        if isinstance(e, Cos):
            return 'pmx_cos(' + self.expr(e.arguments[0], pro) + ')'
        if isinstance(e, Sqrt):
            return 'pmx_sqrt(' + self.expr(e.arguments[0], pro) + ')'
        if isinstance(e, Sqrtf):
            return 'sqrtf(' + self.expr(e.arguments[0], pro) + ')'
        if isinstance(e, Abs):
            return 'pmx_abs(' + self.expr(e.arguments[0], pro) + ')'
        if isinstance(e, RandomFloat):
            raise NotImplementedError("RandomFloat has no GPU lowering")
        raise TypeError(type(e))

    def _fma(self, e, pro):
        """l + a*k, a*k + l, l - a*k or a*k - l with an exact product a*k,
        as fma(a, +-k, +-l): RN(l + p) is what both forms compute."""
        r = exact_product(e.right)
        if r is not None:
            a, k = r
            return 'pmx_fmad((double)(%s), %r, (double)(%s))' % (
                self.expr(a, pro), k if e.op == '+' else -k, self.expr(e.left, pro))
        r = exact_product(e.left)
        if r is not None:
            a, k = r
            return 'pmx_fmad((double)(%s), %r, %s(double)(%s))' % (
                self.expr(a, pro), k, '' if e.op == '+' else '-', self.expr(e.right, pro))
        return None

    def cond(self, c, pro=None):
        if c.conditional in ('&&', '||'):
            return '(' + self.cond(c.lhs, pro) + ' ' + c.conditional + ' ' + \
                self.cond(c.rhs, pro) + ')'
        return '(' + self.expr(c.lhs, pro) + ' ' + c.conditional + ' ' + \
            self.expr(c.rhs, pro) + ')'


# Device helpers that pin the overloads libstdc++ picks on the host.
DEVICE_PRELUDE = r'''
#define PMX_BIG 536870912
#ifndef PMX_GLOBAL
#define PMX_GLOBAL __attribute__((address_space(1)))
#define PMX_CONST __attribute__((address_space(4)))
#define PMX_OPAQUE(x) __asm__ volatile("" : "+v"(x))
#endif
__device__ __forceinline__ int pmx_abs(int v) { return v < 0 ? -v : v; }
__device__ __forceinline__ long long pmx_abs(long long v) { return v < 0 ? -v : v; }
__device__ __forceinline__ float pmx_abs(float v) { return __builtin_fabsf(v); }
__device__ __forceinline__ double pmx_abs(double v) { return __builtin_fabs(v); }
__device__ __forceinline__ float pmx_exp(float v) { return expf(v); }
__device__ __forceinline__ double pmx_exp(double v) { return exp(v); }
__device__ __forceinline__ double pmx_exp(int v) { return exp((double)v); }
__device__ __forceinline__ float pmx_sqrt(float v) { return sqrtf(v); }
__device__ __forceinline__ double pmx_sqrt(double v) { return sqrt(v); }
__device__ __forceinline__ double pmx_sqrt(int v) { return sqrt((double)v); }
__device__ __forceinline__ float pmx_sin(float v) { return sinf(v); }
__device__ __forceinline__ double pmx_sin(double v) { return sin(v); }
__device__ __forceinline__ double pmx_sin(int v) { return sin((double)v); }
__device__ __forceinline__ float pmx_cos(float v) { return cosf(v); }
__device__ __forceinline__ double pmx_cos(double v) { return cos(v); }
__device__ __forceinline__ double pmx_cos(int v) { return cos((double)v); }
__device__ __forceinline__ int pmx_imin(int a, int b) { return a < b ? a : b; }
__device__ __forceinline__ int pmx_imax(int a, int b) { return a > b ? a : b; }
__device__ __forceinline__ int pmx_clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
// value of v in lane src of the wavefront (lane fusion): 32- and 64-bit
// types move as they are, narrower integers through int
template <typename T> __device__ __forceinline__ T pmx_shfl(T v, int src) { return __shfl(v, src, 64); }
__device__ __forceinline__ unsigned char pmx_shfl(unsigned char v, int src) { return (unsigned char)__shfl((int)v, src, 64); }
__device__ __forceinline__ signed char pmx_shfl(signed char v, int src) { return (signed char)__shfl((int)v, src, 64); }
__device__ __forceinline__ char pmx_shfl(char v, int src) { return (char)__shfl((int)v, src, 64); }
__device__ __forceinline__ unsigned short pmx_shfl(unsigned short v, int src) { return (unsigned short)__shfl((int)v, src, 64); }
__device__ __forceinline__ short pmx_shfl(short v, int src) { return (short)__shfl((int)v, src, 64); }
// n / d for n an integer of at most 16 bits, as q + (n - q*d) / d with
// q = n * (1/d): callers pass r = RN(1/d) only where this is checked to
// equal the IEEE quotient for every such n
__device__ __forceinline__ double pmx_fmad(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ double pmx_divc(double n, double r, double d) {
  const double q = n * r;
  return __builtin_fma(__builtin_fma(-q, d, n), r, q);
}
#ifndef PMX_LD4
#define PMX_LD4 1
// four consecutive 4-byte elements at a 16-byte aligned address, one
// 128-bit load (an LDS row run of a 4-wide block)
template <typename V, typename T> __device__ __forceinline__ V pmx_ld4(const T* p) { return *(const V*)p; }
#endif
'''
