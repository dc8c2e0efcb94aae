"""Float64 numpy evaluator of DSL pipelines (ORACLE ONLY).

Evaluates every function over its whole domain with numpy, independent of
both C emitters: floating-point arithmetic in float64 throughout (no
float32 rounding of literals, intermediates or stored values), integer
arithmetic exact in int64 with C semantics (truncating ``/`` and ``%``)
and narrowing to the declared type on Cast and on store.  A point no Case
covers is NaN (floats) so accidental use of undefined data shows up.
Reductions accumulate with ``np.add.at`` (Sum), ``np.multiply.at``,
``np.minimum.at``, ``np.maximum.at``.

Used for the error statement of every float pipeline (float32 GPU output
against this float64 evaluation) and as an independent check of the
integer semantics.
"""

import numpy as np


def _C():
    import polymage_amd.constructs as C
    return C


_INFO = {'Char': (np.int8,), 'UChar': (np.uint8,), 'Short': (np.int16,),
         'UShort': (np.uint16,), 'Int': (np.int32,), 'UInt': (np.uint32,),
         'Long': (np.int64,), 'ULong': (np.uint64,)}


def _is_float(t):
    return t.__name__ in ('Float', 'Double')


def _narrow(v, t):
    """Convert a value to DSL type t: float types stay float64; integer
    types truncate toward zero and wrap like C conversions."""
    if _is_float(t):
        return np.asarray(v, dtype=np.float64)
    v = np.asarray(v)
    if v.dtype.kind == 'f':
        v = np.trunc(v)
        v = np.nan_to_num(v, nan=0.0, posinf=0.0, neginf=0.0)
    dt = _INFO[t.__name__][0]
    return np.asarray(v, dtype=np.int64).astype(dt).astype(np.int64)


def _cdiv(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    if a.dtype.kind == 'f' or b.dtype.kind == 'f':
        return a / b
    bb = np.where(b == 0, 1, b)
    q = np.abs(a) // np.abs(bb)
    return np.where((a >= 0) == (bb >= 0), q, -q)


def _cmod(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    if a.dtype.kind == 'f' or b.dtype.kind == 'f':
        return np.fmod(a, b)
    return a - b * _cdiv(a, b)


class Evaluator(object):
    def __init__(self, outputs, values):
        self.C = _C()
        self.values = values
        self.arrays = {}
        self.params = {}
        self.outputs = outputs
        for k, v in values.items():
            if not isinstance(v, np.ndarray):
                self.params[k] = v

    def ival(self, e, env):
        """Integer-valued bound/index expression."""
        return int(np.asarray(self.ev(e, env)))

    def domain(self, f):
        return [(self.ival(d.lowerBound, {}), self.ival(d.upperBound, {})) for d in f.domain]

    def array_of(self, f):
        C = self.C
        if f in self.arrays:
            return self.arrays[f]
        if isinstance(f, C.Image):
            a = np.asarray(self.values[f.name])
            a = a.astype(np.float64) if _is_float(f.typ) else a.astype(np.int64)
            self.arrays[f] = (a, [(0, s - 1) for s in a.shape])
            return self.arrays[f]
        self.compute(f)
        return self.arrays[f]

    def ev(self, e, env):
        C = self.C
        e = C.Value.numericToValue(e)
        if isinstance(e, C.Value):
            v = e.value
            return np.float64(v) if isinstance(v, float) else np.int64(int(v))
        if isinstance(e, C.Parameter):
            v = self.params[e.name]
            return np.float64(v) if _is_float(e.typ) else np.int64(v)
        if isinstance(e, C.Variable):
            return env[e]
        if isinstance(e, C.AbstractBinaryOpNode):
            l = self.ev(e.left, env)
            r = self.ev(e.right, env)
            op = e.op
            if op == '+':
                return l + r
            if op == '-':
                return l - r
            if op == '*':
                return l * r
            if op == '/':
                return _cdiv(l, r)
            if op == '%':
                return _cmod(l, r)
            if op == '<<':
                return np.left_shift(l, r)
            if op == '>>':
                return np.right_shift(l, r)
            if op == '&':
                return np.bitwise_and(l, r)
            if op == '|':
                return np.bitwise_or(l, r)
            if op == '^':
                return np.bitwise_xor(l, r)
            raise TypeError(op)
        if isinstance(e, C.AbstractUnaryOpNode):
            v = self.ev(e.child, env)
            return -v if e.op == '-' else v
        if isinstance(e, C.Reference):
            arr, dom = self.array_of(e.objectRef)
            idx = []
            valid = True
            for d, a in enumerate(e.arguments):
                q = np.asarray(self.ev(a, env))
                if q.dtype.kind == 'f':
                    q = np.trunc(q)
                q = q.astype(np.int64)
                lo, hi = dom[d]
                valid = valid & (q >= lo) & (q <= hi)
                idx.append(np.clip(q, lo, hi) - lo)
            v = arr[tuple(idx)]
            if v.dtype.kind == 'f':
                v = np.where(valid, v, np.nan)
            return v
        if isinstance(e, C.Select):
            c = self.cond(e.condition, env)
            t = self.ev(e.true_expression, env)
            f = self.ev(e.false_expression, env)
            return np.where(c, t, f)
        if isinstance(e, C.Cast):
            return _narrow(self.ev(e.expression, env), e.typ)
        name = type(e).__name__
        args = [np.asarray(self.ev(a, env), dtype=np.float64) if name != 'Abs' else self.ev(a, env)
                for a in e.arguments]
        with np.errstate(all='ignore'):
            if name in ('Pow', 'Powf'):
                return np.power(args[0], args[1])
            if name == 'Log':
                return np.log(args[0])
            if name == 'Exp':
                return np.exp(args[0])
            if name in ('Sqrt', 'Sqrtf'):
                return np.sqrt(args[0])
            if name == 'Sin':
                return np.sin(args[0])
            if name == 'Cos':
                return np.cos(args[0])
            if name == 'Abs':
                return np.abs(args[0])
        raise TypeError(name)

    def cond(self, c, env):
        if c.conditional == '&&':
            return self.cond(c.lhs, env) & self.cond(c.rhs, env)
        if c.conditional == '||':
            return self.cond(c.lhs, env) | self.cond(c.rhs, env)
        l = self.ev(c.lhs, env)
        r = self.ev(c.rhs, env)
        op = c.conditional
        return {'<': np.less, '<=': np.less_equal, '>': np.greater, '>=': np.greater_equal,
                '==': np.equal, '!=': np.not_equal}[op](l, r)

    def compute(self, f):
        C = self.C
        dom = self.domain(f)
        shape = tuple(max(0, hi - lo + 1) for lo, hi in dom)
        fl = _is_float(f.typ)
        if isinstance(f, C.Reduction):
            env0 = {}
            grids = np.meshgrid(*[np.arange(lo, hi + 1) for lo, hi in dom], indexing='ij')
            for v, g in zip(f.variables, grids):
                env0[v] = g
            arr = np.broadcast_to(_narrow(self.ev(f.default, env0), f.typ), shape).astype(
                np.float64 if fl else np.int64).copy()
            rdom = [(self.ival(d.lowerBound, {}), self.ival(d.upperBound, {})) for d in f.reductionDomain]
            rg = np.meshgrid(*[np.arange(lo, hi + 1) for lo, hi in rdom], indexing='ij')
            env = dict(zip(f.reductionVariables, [g.ravel() for g in rg]))
            for case in f.defn:
                if isinstance(case, C.Case):
                    red, cnd = case.expression, case.condition
                else:
                    red, cnd = case, None
                tgt = red.accumulate_ref.arguments
                idx = []
                for d, a in enumerate(tgt):
                    q = np.broadcast_to(np.asarray(self.ev(a, env)), env[f.reductionVariables[0]].shape)
                    idx.append(np.clip(q.astype(np.int64), dom[d][0], dom[d][1]) - dom[d][0])
                val = np.broadcast_to(np.asarray(self.ev(red.expression, env), dtype=np.float64 if fl else np.int64),
                                      idx[0].shape)
                m = np.ones(idx[0].shape, bool)
                if cnd is not None:
                    cenv = dict(env)
                    for v, a in zip(f.variables, tgt):
                        if v not in f.reductionVariables:
                            cenv[v] = np.broadcast_to(np.asarray(self.ev(a, env)), m.shape)
                    m = np.broadcast_to(self.cond(cnd, cenv), m.shape)
                ii = tuple(i[m] for i in idx)
                op = red.op_type
                ufunc = {C.Op.Sum: np.add, C.Op.Mul: np.multiply, C.Op.Min: np.minimum,
                         C.Op.Max: np.maximum}[op]
                ufunc.at(arr, ii, val[m])
            self.arrays[f] = (arr if fl else _narrow(arr, f.typ), dom)
            return
        arr = np.full(shape, np.nan if fl else 0, dtype=np.float64 if fl else np.int64)
        grids = np.meshgrid(*[np.arange(lo, hi + 1) for lo, hi in dom], indexing='ij')
        env = dict(zip(f.variables, grids))
        for case in f.defn:
            if isinstance(case, C.Case):
                m = np.broadcast_to(self.cond(case.condition, env), shape)
                with np.errstate(all='ignore'):
                    v = np.broadcast_to(_narrow(self.ev(case.expression, env), f.typ), shape)
                arr = np.where(m, v, arr)
            else:
                with np.errstate(all='ignore'):
                    arr = np.broadcast_to(_narrow(self.ev(case, env), f.typ), shape).copy()
        self.arrays[f] = (arr, dom)


def evaluate(outputs, values):
    """Returns {output name: float64/int64 array over its domain}."""
    ev = Evaluator(outputs, values)
    return dict((o.name, ev.array_of(o)[0]) for o in outputs)


def error_statement(got, ref64):
    """Errors of a float32 result against the float64 evaluation over the
    points where the evaluation is defined."""
    got = np.asarray(got, np.float64)
    m = np.isfinite(ref64)
    d = np.abs(got[m] - ref64[m])
    scale = float(np.max(np.abs(ref64[m]))) if m.any() else 1.0
    rel = d / np.maximum(np.abs(ref64[m]), 1e-12 * max(scale, 1e-30))
    return {'max_abs': float(d.max()) if d.size else 0.0,
            'max_rel_normwise': float(d.max() / scale) if d.size and scale > 0 else 0.0,
            'p999_rel': float(np.quantile(rel, 0.999)) if rel.size else 0.0,
            'points': int(m.sum())}
