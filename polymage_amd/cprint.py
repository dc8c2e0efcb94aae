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
                 exact_fma=False):
        self.var_names = var_names
        self.ref_cb = ref_cb
        self.mode = mode
        self.param_prefix = param_prefix
        self.lazy = lazy
        # l +- a * k as one fused multiply-add where the product is exact
        # (exact_product): the same bits as the separate multiply and add
        self.exact_fma = exact_fma and mode == 'double'
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
