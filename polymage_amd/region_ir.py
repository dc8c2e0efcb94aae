"""A tiny SSA integer program for per-tile regions.

The reference derives tile shapes with isl: integer tile slopes per group
and a trapezoid whose extent at level t is ``ts + (left-right)(h-t)``
(poly_schedule.py:128-219, 429-536).  We instead propagate the exact box
each stage needs, from the tile of the group's anchor stage backwards
through every reference (monotone maps such as ``x-2``, ``2*x+1``,
``x/2``).  The same program is (a) evaluated in Python over all tiles to
size LDS buffers, and (b) printed as scalar C that every workgroup runs in
its prologue, so host-side sizing and device-side indexing cannot drift.
"""

from .analysis import ceval_int, eval_bexpr, Max2, Min2
from .expr_ast import Value, Variable, Parameter, AbstractBinaryOpNode, AbstractUnaryOpNode

BIG = 1 << 29


class RegionIR(object):
    def __init__(self):
        self.ops = []
        self._memo = {}

    def _emit(self, key, op):
        if key in self._memo:
            return self._memo[key]
        self.ops.append(op)
        i = len(self.ops) - 1
        self._memo[key] = i
        return i

    # constructors --------------------------------------------------------
    def const(self, n):
        return self._emit(('c', int(n)), ('c', int(n)))

    def expr(self, e):
        """Parameter-only DSL bound expression (may hold Max2/Min2)."""
        e = Value.numericToValue(e)
        if isinstance(e, Value) and isinstance(e.value, int):
            return self.const(e.value)
        return self._emit(('e', id(e)), ('e', e))

    def access(self, e, var, a):
        """DSL index expression ``e`` with consumer variable ``var`` bound to
        SSA value ``a`` (other variables must not occur)."""
        if isinstance(e, Variable) and e is var:
            return a
        return self._emit(('acc', id(e), id(var), a), ('acc', e, var, a))

    def tile(self, j):
        return self._emit(('tile', j), ('tile', j))

    def add(self, a, b):
        return self._emit(('add', a, b), ('add', a, b))

    def sub(self, a, b):
        return self._emit(('sub', a, b), ('sub', a, b))

    def mul(self, a, k):
        if k == 1:
            return a
        return self._emit(('mul', a, k), ('mul', a, k))

    def fdiv(self, a, k):
        if k == 1:
            return a
        return self._emit(('fdiv', a, k), ('fdiv', a, k))

    def min(self, a, b):
        if a == b:
            return a
        return self._emit(('min', a, b), ('min', a, b))

    def max(self, a, b):
        if a == b:
            return a
        return self._emit(('max', a, b), ('max', a, b))

    def sel_valid(self, conds, a, b):
        """all(p <= q for (p, q) in conds) ? a : b"""
        if not conds:
            return a
        conds = tuple(conds)
        return self._emit(('selv', conds, a, b), ('selv', conds, a, b))

    def sel_eq(self, p, q, a, b):
        """(p == q) ? a : b"""
        return self._emit(('seleq', p, q, a, b), ('seleq', p, q, a, b))

    # evaluation ------------------------------------------------------------
    def evaluate(self, env, tiles):
        vals = []
        for op in self.ops:
            k = op[0]
            if k == 'c':
                v = op[1]
            elif k == 'e':
                v = eval_bexpr(op[1], env)
            elif k == 'acc':
                e2 = dict(env)
                e2[op[2]] = vals[op[3]]
                v = ceval_int(op[1], e2)
            elif k == 'tile':
                v = tiles[op[1]]
            elif k == 'add':
                v = vals[op[1]] + vals[op[2]]
            elif k == 'sub':
                v = vals[op[1]] - vals[op[2]]
            elif k == 'mul':
                v = vals[op[1]] * op[2]
            elif k == 'fdiv':
                v = vals[op[1]] // op[2]
            elif k == 'min':
                v = min(vals[op[1]], vals[op[2]])
            elif k == 'max':
                v = max(vals[op[1]], vals[op[2]])
            elif k == 'selv':
                ok = all(vals[p] <= vals[q] for p, q in op[1])
                v = vals[op[2]] if ok else vals[op[3]]
            elif k == 'seleq':
                v = vals[op[3]] if vals[op[1]] == vals[op[2]] else vals[op[4]]
            else:
                raise AssertionError(k)
            vals.append(v)
        return vals

    def evaluate_many(self, env, tiles):
        """Vectorised ``evaluate``: ``tiles`` maps anchor dims to int64
        arrays of equal length; returns one array (or scalar) per op."""
        import numpy as np
        from .analysis import ceval_np
        vals = []
        for op in self.ops:
            k = op[0]
            if k == 'c':
                v = op[1]
            elif k == 'e':
                v = eval_bexpr(op[1], env)
            elif k == 'acc':
                e2 = dict(env)
                e2[op[2]] = vals[op[3]]
                v = ceval_np(op[1], e2)
            elif k == 'tile':
                v = np.asarray(tiles[op[1]], dtype=np.int64)
            elif k == 'add':
                v = vals[op[1]] + vals[op[2]]
            elif k == 'sub':
                v = vals[op[1]] - vals[op[2]]
            elif k == 'mul':
                v = vals[op[1]] * op[2]
            elif k == 'fdiv':
                v = vals[op[1]] // op[2]
            elif k == 'min':
                v = np.minimum(vals[op[1]], vals[op[2]])
            elif k == 'max':
                v = np.maximum(vals[op[1]], vals[op[2]])
            elif k == 'selv':
                ok = True
                for p, q in op[1]:
                    ok = ok & (np.asarray(vals[p]) <= np.asarray(vals[q]))
                v = np.where(ok, vals[op[2]], vals[op[3]])
            elif k == 'seleq':
                v = np.where(np.asarray(vals[op[1]]) == np.asarray(vals[op[2]]),
                             vals[op[3]], vals[op[4]])
            else:
                raise AssertionError(k)
            vals.append(v)
        return vals

    # printing ----------------------------------------------------------------
    def to_c(self, printer_factory, live=None, prefix='r'):
        """C statements defining ``<prefix><i>`` for every op (or only those
        in ``live`` and their operands)."""
        need = set(range(len(self.ops))) if live is None else self._closure(live)
        lines = []
        for i, op in enumerate(self.ops):
            if i not in need:
                continue
            k = op[0]
            n = lambda j: '%s%d' % (prefix, j)
            if k == 'c':
                rhs = str(op[1])
            elif k == 'e':
                rhs = bound_c(op[1], printer_factory({}))
            elif k == 'acc':
                rhs = printer_factory({op[2]: n(op[3])}).expr(op[1])
            elif k == 'tile':
                rhs = 't%d' % op[1]
            elif k == 'add':
                rhs = '%s + %s' % (n(op[1]), n(op[2]))
            elif k == 'sub':
                rhs = '%s - %s' % (n(op[1]), n(op[2]))
            elif k == 'mul':
                rhs = '%s * %d' % (n(op[1]), op[2])
            elif k == 'fdiv':
                rhs = 'pmx_fdiv(%s, %d)' % (n(op[1]), op[2])
            elif k == 'min':
                rhs = 'pmx_imin(%s, %s)' % (n(op[1]), n(op[2]))
            elif k == 'max':
                rhs = 'pmx_imax(%s, %s)' % (n(op[1]), n(op[2]))
            elif k == 'selv':
                c = ' && '.join('%s <= %s' % (n(p), n(q)) for p, q in op[1])
                rhs = '(%s) ? %s : %s' % (c, n(op[2]), n(op[3]))
            elif k == 'seleq':
                rhs = '(%s == %s) ? %s : %s' % (n(op[1]), n(op[2]), n(op[3]), n(op[4]))
            lines.append('const int %s = %s;' % (n(i), rhs))
        return lines

    def _closure(self, live):
        need = set()
        stack = list(live)
        while stack:
            i = stack.pop()
            if i in need:
                continue
            need.add(i)
            op = self.ops[i]
            k = op[0]
            if k == 'acc':
                stack.append(op[3])
            elif k in ('add', 'sub', 'min', 'max'):
                stack += [op[1], op[2]]
            elif k in ('mul', 'fdiv'):
                stack.append(op[1])
            elif k == 'selv':
                for p, q in op[1]:
                    stack += [p, q]
                stack += [op[2], op[3]]
            elif k == 'seleq':
                stack += [op[1], op[2], op[3], op[4]]
        return need


def bound_c(e, printer):
    """Print a bound expression that may contain Max2/Min2."""
    if isinstance(e, Max2):
        return 'pmx_imax(%s, %s)' % (bound_c(e.a, printer), bound_c(e.b, printer))
    if isinstance(e, Min2):
        return 'pmx_imin(%s, %s)' % (bound_c(e.a, printer), bound_c(e.b, printer))
    return printer.expr(e)


REGION_PRELUDE = r'''
__device__ __forceinline__ int pmx_fdiv(int a, int k) { return a >= 0 ? a / k : -((-a + k - 1) / k); }
'''
