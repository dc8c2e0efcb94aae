"""Kernels for the groups that are not fused stencils.

* ``ConstPlan``: ``Function(..., "const")`` tables (look-up tables, colour
  matrices).  The reference never fuses them (grouping.py:146-161).  A
  table that depends only on parameters is computed on the host inside the
  stub with the host C library (the same libm the reference's g++ build
  uses, so ``powf``/``exp`` tables are bit-identical) and uploaded once per
  parameter change; a table that reads images is computed by a small
  per-point kernel.
* ``ReductionPlan``: ``Reduction`` stages (also never fused).  When every
  target index along a tiled dimension is a monotone floor-division map of
  one reduction variable (``2 + x/8``) the reduction is computed by
  "owner cells": each thread owns one cell of those dimensions, walks its
  pre-image of the reduction domain in the reference's lexicographic order
  and accumulates every data-dependent dimension (``2 + (int)(img*10)``) in
  registers, so the float sums are the reference's serial sums, bit for
  bit, with no atomics.  Other reductions run the reference's serial loop
  in a single thread (exact, slow; histograms and other small cases).
"""

from fractions import Fraction

from .analysis import PlanError, classify_access, case_parts, ceval_int, eval_bexpr, interval_image
from .constructs import Reference, Image, Case, Reduce, Op, Reduction
from .cprint import Printer, c_type
from .expr_ast import Variable, Value
from .region_ir import bound_c


def _prod(xs):
    n = 1
    for x in xs:
        n *= x
    return n


def _extent_c(f, d, pr):
    return '(%s - %s + 1)' % (pr.expr(f.domain[d].upperBound), pr.expr(f.domain[d].lowerBound))


def _refs_of_parts(parts):
    refs = []
    for part in parts:
        refs += part.expr.collect(Reference)
        for c in part.pred:
            refs += c.collect(Reference)
    return refs


class ConstPlan(object):
    def __init__(self, pipe, group, gindex):
        self.pipe = pipe
        self.group = group
        self.gindex = gindex
        self.comp = group.comps[0]
        self.func = self.comp.func
        self.parts = self.comp.parts
        refs = _refs_of_parts(self.parts)
        self.on_host = len(refs) == 0
        self.params = pipe.sorted_parameters()
        self.upload_fn = 'pmx_upload_%s' % self.func.name

    def describe(self):
        return 'group %d: const table %s (%s)' % (self.gindex, self.func.name,
                                                  'host' if self.on_host else 'device')

    def host_code(self, pr):
        f = self.func
        T = c_type(f.typ)
        nd = len(f.variables)
        names = ['h%d' % i for i in range(nd)]
        vn = dict(zip(f.variables, names))
        hp = Printer(vn, None, self.pipe.literal_mode)
        L = []
        w = L.append
        w('static inline float pmx_exp(float v) { return std::exp(v); }')
        w('static inline double pmx_exp(double v) { return std::exp(v); }')
        w('static inline double pmx_exp(int v) { return std::exp((double)v); }')
        w('static inline float pmx_sqrt(float v) { return std::sqrt(v); }')
        w('static inline double pmx_sqrt(double v) { return std::sqrt(v); }')
        w('static inline float pmx_abs(float v) { return std::abs(v); }')
        w('static inline double pmx_abs(double v) { return std::abs(v); }')
        w('static inline int pmx_abs(int v) { return std::abs(v); }')
        w('static inline float pmx_sin(float v) { return std::sin(v); }')
        w('static inline double pmx_sin(double v) { return std::sin(v); }')
        w('static inline float pmx_cos(float v) { return std::cos(v); }')
        w('static inline double pmx_cos(double v) { return std::cos(v); }')
        sig = ', '.join('%s %s' % (c_type(p.typ), p.name) for p in self.params)
        w('static std::vector<%s> pmx_tab_%s;' % (T, f.name))
        w('static std::vector<unsigned char> pmx_tabkey_%s;' % f.name)
        w('static bool pmx_tab_fresh_%s[64];' % f.name)
        w('static int %s(%s%svoid* stream, void** out) {' % (self.upload_fn, sig, ', ' if sig else ''))
        # key = raw bytes of the parameters
        w('  std::vector<unsigned char> key;')
        for p in self.params:
            w('  { const unsigned char* q = (const unsigned char*)&%s; key.insert(key.end(), q, q + sizeof(%s)); }'
              % (p.name, p.name))
        size = ' * '.join('(long long)%s' % _extent_c(f, d, hp) for d in range(nd)) or '1'
        w('  long long n = %s;' % size)
        w('  if (key != pmx_tabkey_%s || (long long)pmx_tab_%s.size() != n) {' % (f.name, f.name))
        w('    pmx_tab_%s.assign((size_t)n, (%s)0);' % (f.name, T))
        for i in range(nd):
            w('    ' + '  ' * i + 'for (int %s = %s; %s <= %s; ++%s) {' % (
                names[i], hp.expr(f.domain[i].lowerBound), names[i],
                hp.expr(f.domain[i].upperBound), names[i]))
        ind = '    ' + '  ' * nd
        lin = None
        for i in range(nd):
            t = '(long long)(%s - (%s))' % (names[i], hp.expr(f.domain[i].lowerBound))
            lin = t if lin is None else '(%s * %s + %s)' % (lin, _extent_c(f, i, hp), t)
        for part in self.parts:
            conds = []
            for i, (blo, bhi) in enumerate(part.box):
                if blo is not None:
                    conds.append('%s >= %s' % (names[i], bound_c(blo, hp)))
                if bhi is not None:
                    conds.append('%s <= %s' % (names[i], bound_c(bhi, hp)))
            pro = []
            pc = [hp.cond(c, pro) for c in part.pred]
            val = hp.expr(part.expr, pro)
            w(ind + 'if (%s) {' % (' && '.join(conds) or 'true'))
            for l in pro:
                w(ind + '  ' + l)
            w(ind + '  if (%s) pmx_tab_%s[%s] = %s;' % (' && '.join(pc) or 'true', f.name, lin or '0', val))
            w(ind + '}')
        for i in reversed(range(nd)):
            w('    ' + '  ' * i + '}')
        w('    pmx_tabkey_%s = key;' % f.name)
        w('    for (int d = 0; d < 64; ++d) pmx_tab_fresh_%s[d] = false;' % f.name)
        w('  }')
        w('  int dev = 0; if (hipGetDevice(&dev) != hipSuccess) return -1;')
        w('  int rc = pmx_arena_get("%s/const/%s", (size_t)n * sizeof(%s), out); if (rc) return rc;'
          % (self.pipe.name, f.name, T))
        w('  if (!pmx_tab_fresh_%s[dev]) {' % f.name)
        w('    rc = pmx_memcpy_h2d(*out, pmx_tab_%s.data(), (size_t)n * sizeof(%s), stream); if (rc) return rc;'
          % (f.name, T))
        w('    rc = pmx_stream_sync(stream); if (rc) return rc;')
        w('    pmx_tab_fresh_%s[dev] = true;' % f.name)
        w('  }')
        w('  return 0;')
        w('}')
        return L


def emit_const_kernel(pipe, plan, kname, mode):
    """One thread per table point; the table reads images (e.g. a colour
    matrix interpolated from two input matrices)."""
    from .hip_emit import KernelInfo
    f = plan.func
    params = pipe.sorted_parameters()
    reads = []
    for r in _refs_of_parts(plan.parts):
        if r.objectRef not in reads:
            reads.append(r.objectRef)
    buffers = sorted(reads, key=lambda g: g.name) + [f]
    nd = len(f.variables)
    names = ['p%d' % i for i in range(nd)]
    base = Printer({}, None, mode)

    def cb(ref, pr, pro):
        g = ref.objectRef
        lin = None
        for d, arg in enumerate(ref.arguments):
            q = 'pmx_clampi(%s, %s, %s)' % (pr.expr(arg, pro), base.expr(g.domain[d].lowerBound),
                                             base.expr(g.domain[d].upperBound))
            t = '(%s - (%s))' % (q, base.expr(g.domain[d].lowerBound))
            lin = t if lin is None else '(%s * %s + %s)' % (lin, _extent_c(g, d, base), t)
        return 'G_%s[%s]' % (g.name, lin or '0')
    pr = Printer(dict(zip(f.variables, names)), cb, mode)
    args = ['const %s %s' % (c_type(p.typ), p.name) for p in params]
    for g in buffers:
        if g is f:
            args.append('%s* __restrict__ G_%s' % (c_type(g.typ), g.name))
        else:
            args.append('const %s* __restrict__ G_%s' % (c_type(g.typ), g.name))
    L = ['extern "C" __global__ __launch_bounds__(256) void %s(%s) {' % (kname, ', '.join(args))]
    w = lambda s: L.append('  ' + s)
    exts = [_extent_c(f, d, base) for d in range(nd)]
    w('long long pmx_n = %s;' % (' * '.join('(long long)%s' % e for e in exts) or '1'))
    w('long long idx = (long long)blockIdx.x * 256 + threadIdx.x;')
    w('if (idx >= pmx_n) return;')
    w('long long rem = idx;')
    for i in reversed(range(nd)):
        w('const int %s = (int)(rem %% %s) + (%s); rem /= %s;' % (
            names[i], exts[i], base.expr(f.domain[i].lowerBound), exts[i]))
    w('%s v = (%s)0; bool pmx_w = false;' % (c_type(f.typ), c_type(f.typ)))
    for part in plan.parts:
        conds = []
        for i, (blo, bhi) in enumerate(part.box):
            if blo is not None:
                conds.append('%s >= %s' % (names[i], bound_c(blo, base)))
            if bhi is not None:
                conds.append('%s <= %s' % (names[i], bound_c(bhi, base)))
        w('if (%s) {' % (' && '.join(conds) or 'true'))
        pro = []
        pc = [pr.cond(c, pro) for c in part.pred]
        val = pr.expr(part.expr, pro)
        for l in pro:
            w('  ' + l)
        w('  if (%s) { v = %s; pmx_w = true; }' % (' && '.join(pc) or 'true', val))
        w('}')
    w('if (pmx_w) G_%s[idx] = v;' % f.name)
    L.append('}')
    grid = '((%s + 255) / 256)' % (' * '.join('(long long)%s' % e for e in exts) or '1')
    return '\n'.join(L) + '\n', KernelInfo(kname, 'const', buffers, grid, 256, 0, plan,
                                           [g for g in buffers if g is not f], [f])


class ReductionPlan(object):
    """See module docstring."""

    def __init__(self, pipe, group, gindex):
        self.pipe = pipe
        self.group = group
        self.gindex = gindex
        self.comp = group.comps[0]
        self.func = self.comp.func
        raise PlanError("reductions are planned by ReductionPlan.build (not yet available)")


def emit_reduction_kernel(pipe, plan, kname, mode):
    raise PlanError("reduction emission unavailable")
