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
  bit, with no atomics.  Reduction variables that appear in no target
  index (a row reduction ``hist(c, x) op= img(x, y)`` over y) are looped
  inside each owner thread in the domain's order, so those sums keep the
  reference's order too.  Integer reductions whose target is data
  dependent (histograms: ``hist(bin(img(x, y))) += 1``) have a result that
  does not depend on the order of their updates (integer +, *, min and max
  are associative and commutative, modulo 2^32 for + and *), so they run
  one thread per reduction point with device atomics, after a kernel that
  writes the default value.  The remaining cases (a float reduction into
  data-dependent cells with no owned dimension) run the reference's serial
  loop in a single thread (exact, slow).
"""

from fractions import Fraction

from .analysis import PlanError, classify_access, case_parts, ceval_int, eval_bexpr, interval_image
from .constructs import Reference, Image, Case, Reduce, Op, Reduction, Parameter
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

    def host_code(self, pr, host_only=False):
        """C++ computing the table on the host (``<upload_fn>_host``) and,
        unless ``host_only``, uploading it to the device arena once per
        parameter key (``<upload_fn>``)."""
        f = self.func
        T = c_type(f.typ)
        nd = len(f.variables)
        names = ['h%d' % i for i in range(nd)]
        vn = dict(zip(f.variables, names))
        hp = Printer(vn, None, self.pipe.literal_mode)
        L = []
        w = L.append
        # host versions of the math helpers, once per translation unit
        w('#ifndef PMX_HOST_MATH')
        w('#define PMX_HOST_MATH')
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
        w('static inline double pmx_divc(double n, double r, double d) '
          '{ const double q = n * r; return std::fma(std::fma(-q, d, n), r, q); }')
        w('#endif')
        sig = ', '.join('%s %s' % (c_type(p.typ), p.name) for p in self.params)
        used = set(id(q) for q in f.getObjects(Parameter))
        keyp = [p for p in self.params if id(p) in used]
        w('static std::map<std::vector<unsigned char>, std::vector<%s> > pmx_tab_%s;' % (T, f.name))
        w('static std::set<std::pair<int, std::vector<unsigned char> > > pmx_tab_up_%s;' % f.name)
        # the host table (computed once per parameter key, kept for the
        # process) and its byte size
        w('static int %s_host(%s%sconst %s** data, size_t* bytes, std::vector<unsigned char>* keyp) {'
          % (self.upload_fn, sig, ', ' if sig else '', T))
        for p in self.params:
            w('  (void)%s;' % p.name)
        w('  std::vector<unsigned char> key;')
        for p in keyp:
            w('  { const unsigned char* q = (const unsigned char*)&%s; key.insert(key.end(), q, q + sizeof(%s)); }'
              % (p.name, p.name))
        size = ' * '.join('(long long)%s' % _extent_c(f, d, hp) for d in range(nd)) or '1'
        w('  long long n = %s;' % size)
        w('  std::vector<%s>& tab = pmx_tab_%s[key];' % (T, f.name))
        w('  if ((long long)tab.size() != n) {')
        w('    tab.assign((size_t)n, (%s)0);' % T)
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
            w(ind + '  if (%s) tab[%s] = %s;' % (' && '.join(pc) or 'true', lin or '0', val))
            w(ind + '}')
        for i in reversed(range(nd)):
            w('    ' + '  ' * i + '}')
        w('  }')
        w('  *data = tab.data(); *bytes = (size_t)n * sizeof(%s);' % T)
        w('  if (keyp) *keyp = key;')
        w('  return 0;')
        w('}')
        if host_only:
            return L
        w('static int %s(%s%svoid* stream, void** out) {' % (self.upload_fn, sig, ', ' if sig else ''))
        w('  const %s* tdata = nullptr; size_t tbytes = 0; std::vector<unsigned char> key;' % T)
        w('  int rc = %s_host(%s%s&tdata, &tbytes, &key); if (rc) return rc;'
          % (self.upload_fn, ', '.join(p.name for p in self.params), ', ' if self.params else ''))
        w('  int dev = 0; if (hipGetDevice(&dev) != hipSuccess) return -1;')
        w('  unsigned long long h = 1469598103934665603ULL;')
        w('  for (unsigned char b : key) { h ^= b; h *= 1099511628211ULL; }')
        w('  char slot[128]; snprintf(slot, sizeof slot, "%s/const/%s/%%016llx", h);' % (self.pipe.name, f.name))
        w('  rc = pmx_arena_get(slot, tbytes, out); if (rc) return rc;')
        w('  if (!pmx_tab_up_%s.count(std::make_pair(dev, key))) {' % f.name)
        w('    rc = pmx_memcpy_h2d(*out, tdata, tbytes, stream); if (rc) return rc;')
        w('    rc = pmx_stream_sync(stream); if (rc) return rc;')
        w('    pmx_tab_up_%s.insert(std::make_pair(dev, key));' % f.name)
        w('  }')
        w('  return 0;')
        w('}')
        return L


def emit_const_kernel(pipe, plan, kname, mode):
    """One thread per table point; the table reads images (e.g. a colour
    matrix interpolated from two input matrices)."""
    from .hip_emit import KernelInfo, wrap_kernel
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
    pr = Printer(dict(zip(f.variables, names)), cb, mode, lazy=True)
    args = ['const %s %s' % (c_type(p.typ), p.name) for p in params]
    for g in buffers:
        if g is f:
            args.append('%s* __restrict__ G_%s' % (c_type(g.typ), g.name))
        else:
            args.append('const %s* __restrict__ G_%s' % (c_type(g.typ), g.name))
    L = []
    w = lambda s: L.append('  ' + s)
    exts = [_extent_c(f, d, base) for d in range(nd)]
    w('long long pmx_n = %s;' % (' * '.join('(long long)%s' % e for e in exts) or '1'))
    w('long long idx = (long long)pmx_tile * 256 + threadIdx.x;')
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
    grid = '((%s + 255) / 256)' % (' * '.join('(long long)%s' % e for e in exts) or '1')
    info = KernelInfo(kname, 'const', buffers, grid, 256, 0, plan,
                      [g for g in buffers if g is not f], [f])
    return wrap_kernel(kname, params, buffers, [f], 256, 0, L), info


class _RedCase(object):
    def __init__(self, cond, red, index):
        self.cond = cond
        self.red = red
        self.index = index


class ReductionPlan(object):
    """See module docstring.  ``owner`` is True when the owner-cell kernel
    applies; ``owned`` maps each reduction variable to the target dimension
    it owns."""

    MAX_ACC = 64

    def __init__(self, pipe, group, gindex):
        self.pipe = pipe
        self.group = group
        self.gindex = gindex
        self.comp = group.comps[0]
        f = self.func = self.comp.func
        self.rvars = list(f.reductionVariables)
        self.fvars = list(f.variables)
        self.cases = []
        for i, case in enumerate(f.defn):
            if isinstance(case, Case):
                self.cases.append(_RedCase(case.condition, case.expression, i))
            else:
                self.cases.append(_RedCase(None, case, i))
        for rc in self.cases:
            if not isinstance(rc.red, Reduce):
                raise PlanError("reduction %s: definition must be Reduce" % f.name)
        self._classify()

    def _classify(self):
        f = self.func
        nd = len(self.fvars)
        owned = {}
        self.pre_size = {}
        ok = True
        self.loop_vars = []
        for j, rv in enumerate(self.rvars):
            dim = None
            hits = []
            for rc in self.cases:
                tgt = rc.red.accumulate_ref.arguments
                hit = [d for d in range(nd)
                       if classify_access(tgt[d], self.rvars).kind == 'var'
                       and classify_access(tgt[d], self.rvars).var is rv]
                hits.append(hit)
            if all(not h for h in hits) and not any(
                    rv in [v for v in _vars_of(rc.red.accumulate_ref.arguments[d])]
                    for rc in self.cases for d in range(nd)
                    if classify_access(rc.red.accumulate_ref.arguments[d], self.rvars).kind != '*'):
                # in no affine target index: every owner thread walks it
                self.loop_vars.append(rv)
                continue
            for rc, hit in zip(self.cases, hits):
                tgt = rc.red.accumulate_ref.arguments
                if len(hit) != 1:
                    ok = False
                    break
                if dim is None:
                    dim = (hit[0], str(tgt[hit[0]]))
                elif dim != (hit[0], str(tgt[hit[0]])):
                    ok = False
                    break
            if not ok or dim is None:
                ok = False
                break
            acc = classify_access(rc.red.accumulate_ref.arguments[dim[0]], self.rvars)
            if not acc.increasing:
                ok = False
                break
            owned[rv] = dim[0]
            # cells own k consecutive values of rv when the access is rv/k + c
            if acc.slope.numerator == 1 and acc.slope.denominator > 1:
                self.pre_size[rv] = int(acc.slope.denominator)
        env = self.pipe.param_env
        self.free_dims = [d for d in range(nd) if d not in owned.values()]
        nacc = 1
        for d in self.free_dims:
            lo = ceval_int(f.domain[d].lowerBound, env)
            hi = ceval_int(f.domain[d].upperBound, env)
            nacc *= (hi - lo + 1)
        self.nacc = nacc
        self.owner = ok and bool(owned) and len(owned) + len(self.loop_vars) == len(self.rvars) \
            and nacc <= self.MAX_ACC
        self.owned = owned
        self.atomic = (not self.owner) and self._atomic_ok()

    def _atomic_ok(self):
        """Integer target, integer values of the target's promoted type and
        an operator whose result is independent of the update order."""
        from .expression import getType
        from .expr_types import Int, UInt, Char, UChar, Short, UShort
        f = self.func
        if f.typ not in (Int, UInt):
            return False
        for rc in self.cases:
            if rc.red.op_type not in (Op.Sum, Op.Mul, Op.Min, Op.Max):
                return False
            vt = getType(rc.red.expression)
            if vt in (Char, UChar, Short, UShort):
                vt = Int
            if vt is not f.typ:
                return False
        return True

    def describe(self):
        kind = 'owner cells' if self.owner else ('atomic' if self.atomic else 'serial')
        return 'group %d: reduction %s (%s, %d accumulators per cell)' % (
            self.gindex, self.func.name, kind, self.nacc)


def _vars_of(e):
    return e.collect(Variable) if hasattr(e, 'collect') else []


def _prod_int(xs):
    n = 1
    for x in xs:
        n *= x
    return n


def _reads_of_reduction(plan):
    reads = []
    for rc in plan.cases:
        refs = rc.red.expression.collect(Reference)
        refs += [r for r in rc.red.accumulate_ref.collect(Reference) if r is not rc.red.accumulate_ref]
        if rc.cond is not None:
            refs += rc.cond.collect(Reference)
        for r in refs:
            if r.objectRef is not plan.func and r.objectRef not in reads:
                reads.append(r.objectRef)
    return reads


def emit_reduction_kernel(pipe, plan, kname, mode):
    from .hip_emit import KernelInfo, wrap_kernel
    f = plan.func
    params = pipe.sorted_parameters()
    reads = _reads_of_reduction(plan)
    buffers = sorted(reads, key=lambda g: g.name) + [f]
    base = Printer({}, None, mode)
    T = c_type(f.typ)
    nd = len(plan.fvars)
    rnames = ['r%d' % j for j in range(len(plan.rvars))]

    def cb(ref, pr, pro):
        g = ref.objectRef
        lin = None
        for d, arg in enumerate(ref.arguments):
            q = pr.expr(arg, pro)
            acc = classify_access(arg, plan.rvars)
            if acc.kind == '*':
                q = 'pmx_clampi(%s, %s, %s)' % (q, base.expr(g.domain[d].lowerBound),
                                                base.expr(g.domain[d].upperBound))
            t = '(%s - (%s))' % (q, base.expr(g.domain[d].lowerBound))
            lin = t if lin is None else '(%s * %s + %s)' % (lin, _extent_c(g, d, base), t)
        return 'G_%s[%s]' % (g.name, lin or '0')

    args = ['const %s %s' % (c_type(p.typ), p.name) for p in params]
    for g in buffers:
        if g is f:
            args.append('%s* __restrict__ G_%s' % (T, g.name))
        else:
            args.append('const %s* __restrict__ G_%s' % (c_type(g.typ), g.name))
    L = []
    w = lambda s, ind=1: L.append('  ' * ind + s)
    exts = [_extent_c(f, d, base) for d in range(nd)]
    lbs = [base.expr(f.domain[d].lowerBound) for d in range(nd)]

    def red_body(ind, target_ok, update=None):
        """Statements processing the current reduction point (names
        r0..rk bound).  ``target_ok(dimvals)`` filters owned cells;
        ``update(op, val)`` gives the statement applying one value (by
        default the C update of ``acc_slot()``)."""
        pr = Printer(dict(zip(plan.rvars, rnames)), cb, mode, lazy=True)
        for rc in plan.cases:
            red = rc.red
            tgt = red.accumulate_ref.arguments
            w('{', ind)
            pro = []
            tnames = []
            for d in range(nd):
                q = pr.expr(tgt[d], pro)
                if classify_access(tgt[d], plan.rvars).kind == '*':
                    q = 'pmx_clampi(%s, %s, %s)' % (q, lbs[d], base.expr(f.domain[d].upperBound))
                tnames.append(q)
            val = pr.expr(red.expression, pro)
            for l in pro:
                w(l, ind + 1)
            for d in range(nd):
                w('const int q%d = %s;' % (d, tnames[d]), ind + 1)
            conds = []
            if rc.cond is not None:
                vn = dict(zip(plan.rvars, rnames))
                for v, d in zip(plan.fvars, range(nd)):
                    if v not in plan.rvars:
                        vn[v] = 'q%d' % d
                cpr = Printer(vn, cb, mode, lazy=True)
                cpro = []
                cc = cpr.cond(rc.cond, cpro)
                for l in cpro:
                    w(l, ind + 1)
                conds.append(cc)
            extra = target_ok()
            if extra:
                conds.append(extra)
            op = red.op_type
            if update is not None:
                upd = update(op, val)
            else:
                slot = acc_slot()
                if op == Op.Sum:
                    upd = '%s = (%s + %s);' % (slot, slot, val)
                elif op == Op.Mul:
                    upd = '%s = (%s * %s);' % (slot, slot, val)
                elif op == Op.Min:
                    upd = '%s = ((%s < %s)? (%s)(%s): %s);' % (slot, val, slot, T, val, slot)
                else:
                    upd = '%s = ((%s > %s)? (%s)(%s): %s);' % (slot, val, slot, T, val, slot)
            w('if (%s) { %s }' % (' && '.join(conds) or 'true', upd), ind + 1)
            w('}', ind)

    if plan.owner:
        # thread per owned cell; accumulators for the free dims in LDS
        od = sorted(plan.owned.values())
        free = plan.free_dims
        nthreads = 256
        w_ = None
        w('%s* pmx_acc = (%s*)pmx_lds;' % (T, T))
        w('%s* acc = pmx_acc + threadIdx.x;' % T)
        ncell = ' * '.join('(long long)%s' % exts[d] for d in od)
        w('const long long pmx_ncell = %s;' % ncell)
        w('long long cell = (long long)pmx_tile * %d + threadIdx.x;' % nthreads)
        w('if (cell >= pmx_ncell) return;')
        for d in reversed(od):
            w('const int c%d = (int)(cell %% %s) + (%s); cell /= %s;' % (d, exts[d], lbs[d], exts[d]))
        # free-dim strides in the accumulator array (stride nthreads)
        fext = [exts[d] for d in free]

        def acc_index(names):
            e = None
            for k, d in enumerate(free):
                t = '(%s - (%s))' % (names[d], lbs[d])
                e = t if e is None else '(%s * %s + %s)' % (e, fext[k], t)
            return '(%s) * %d' % (e or '0', nthreads)
        # init with the default value
        dvn = {}
        w('for (int a = 0; a < %d; ++a) {' % plan.nacc)
        rem = 'a'
        fn = {}
        for k in reversed(range(len(free))):
            d = free[k]
            w('  const int f%d = (%s) %% %s + (%s);' % (d, rem, fext[k], lbs[d]))
            w('  const int a%d = (%s) / %s;' % (d, rem, fext[k]))
            rem = 'a%d' % d
            fn[d] = 'f%d' % d
        names = dict((d, 'c%d' % d) for d in od)
        names.update(fn)
        dpr = Printer(dict((plan.fvars[d], names[d]) for d in range(nd)), cb, mode, lazy=True)
        w('  acc[a * %d] = %s;' % (nthreads, dpr.expr(f.default)))
        w('}')
        # pre-image of the cell along every owned reduction variable; the
        # other reduction variables are walked over their whole range
        for j, rv in enumerate(plan.rvars):
            if rv not in plan.owned:
                w('const int lo%d = %s, hi%d = %s;' % (
                    j, base.expr(f.reductionDomain[j].lowerBound), j,
                    base.expr(f.reductionDomain[j].upperBound)))
                continue
            d = plan.owned[rv]
            arg = plan.cases[0].red.accumulate_ref.arguments[d]
            rlo = base.expr(f.reductionDomain[j].lowerBound)
            rhi = base.expr(f.reductionDomain[j].upperBound)
            fpr = Printer({rv: 'm'}, cb, mode, lazy=True)
            fx = fpr.expr(arg)
            # lo_j = first r with f(r) >= c ; hi_j = last r with f(r) <= c
            w('int lo%d, hi%d;' % (j, j))
            w('{ int a = %s, b = %s + 1;' % (rlo, rhi))
            w('  while (a < b) { const int m = a + (b - a) / 2; if ((%s) >= c%d) b = m; else a = m + 1; }' % (fx, d))
            w('  lo%d = a; }' % j)
            w('{ int a = %s - 1, b = %s;' % (rlo, rhi))
            w('  while (a < b) { const int m = a + (b - a + 1) / 2; if ((%s) <= c%d) a = m; else b = m - 1; }' % (fx, d))
            w('  hi%d = a; }' % j)
        def acc_slot_owner():
            nm = dict((d, 'q%d' % d) for d in free)
            return 'acc[%s]' % acc_index(nm)
        acc_slot = acc_slot_owner

        def owned_ok():
            return ' && '.join('q%d == c%d' % (d, d) for d in od)

        ks = [plan.pre_size.get(rv) for rv in plan.rvars]
        fast = all(ks) and _prod_int(ks) <= 256
        if fast:
            # interior cells own a full k0 x k1 x ... pre-image: constant trip
            # counts, fully unrolled, so all its loads issue up front (the
            # order of the accumulations, hence every sum, is unchanged)
            w('if (%s) {' % ' && '.join('hi%d - lo%d + 1 == %d' % (j, j, k) for j, k in enumerate(ks)))
            ind = 2
            for j, k in enumerate(ks):
                w('#pragma unroll', ind)
                w('for (int i%d = 0; i%d < %d; ++i%d) { const int r%d = lo%d + i%d;' % (j, j, k, j, j, j, j), ind)
                ind += 1
            red_body(ind, lambda: '')
            for j in reversed(range(len(plan.rvars))):
                ind -= 1
                w('}', ind)
            w('} else {')
        ind = 2 if fast else 1
        for j in range(len(plan.rvars)):
            w('for (int r%d = lo%d; r%d <= hi%d; ++r%d) {' % (j, j, j, j, j), ind)
            ind += 1
        red_body(ind, owned_ok)
        for j in reversed(range(len(plan.rvars))):
            ind -= 1
            w('}', ind)
        if fast:
            w('}')
        # write out the cell's slice of the grid
        w('for (int a = 0; a < %d; ++a) {' % plan.nacc)
        rem = 'a'
        for k in reversed(range(len(free))):
            d = free[k]
            w('  const int f%d = (%s) %% %s + (%s);' % (d, rem, fext[k], lbs[d]))
            w('  const int a%d = (%s) / %s;' % (d, rem, fext[k]))
            rem = 'a%d' % d
        lin = None
        for d in range(nd):
            t = '(long long)(%s - (%s))' % (names[d], lbs[d])
            lin = t if lin is None else '(%s * %s + %s)' % (lin, exts[d], t)
        w('  G_%s[%s] = acc[a * %d];' % (f.name, lin, nthreads))
        w('}')
        grid = '((%s + %d) / %d)' % (ncell, nthreads - 1, nthreads)
        lds = nthreads * plan.nacc * f.typ.nbytes()
        info = KernelInfo(kname, 'reduction', buffers, grid, nthreads, lds, plan, reads, [f])
        return [(wrap_kernel(kname, params, buffers, [f], nthreads, lds, L), info)]

    if plan.atomic:
        return _emit_atomic(pipe, plan, kname, mode, params, reads, buffers, exts, lbs, base,
                            red_body, L)

    # serial fallback: one thread, the reference loop order
    w('if (threadIdx.x != 0) return;')
    lin = None
    ind = 1
    for d in range(nd):
        w('for (int f%d = %s; f%d <= %s; ++f%d) {' % (d, lbs[d], d, base.expr(f.domain[d].upperBound), d), ind)
        ind += 1
        t = '(long long)(f%d - (%s))' % (d, lbs[d])
        lin = t if lin is None else '(%s * %s + %s)' % (lin, exts[d], t)
    dpr = Printer(dict((plan.fvars[d], 'f%d' % d) for d in range(nd)), cb, mode, lazy=True)
    w('G_%s[%s] = %s;' % (f.name, lin or '0', dpr.expr(f.default)), ind)
    for d in reversed(range(nd)):
        ind -= 1
        w('}', ind)
    for j in range(len(plan.rvars)):
        w('for (int r%d = %s; r%d <= %s; ++r%d) {' % (
            j, base.expr(f.reductionDomain[j].lowerBound), j,
            base.expr(f.reductionDomain[j].upperBound), j), ind)
        ind += 1
    tl = None
    for d in range(nd):
        t = '(long long)(q%d - (%s))' % (d, lbs[d])
        tl = t if tl is None else '(%s * %s + %s)' % (tl, exts[d], t)

    def acc_slot():
        return 'G_%s[%s]' % (f.name, tl or '0')
    red_body(ind, lambda: None)
    for j in reversed(range(len(plan.rvars))):
        ind -= 1
        w('}', ind)
    info = KernelInfo(kname, 'reduction', buffers, '1', 64, 0, plan, reads, [f])
    return [(wrap_kernel(kname, params, buffers, [f], 64, 0, L), info)]


_ATOMIC = {Op.Sum: 'atomicAdd', Op.Min: 'atomicMin', Op.Max: 'atomicMax'}


def _emit_atomic(pipe, plan, kname, mode, params, reads, buffers, exts, lbs, base, red_body, L):
    """Two kernels: ``<k>_init`` writes the default value at every target
    point, ``<k>`` runs one thread per reduction point (the domain
    flattened in its own order) and applies its value with a device
    atomic (a compare-and-swap loop for products)."""
    from .hip_emit import KernelInfo, wrap_kernel
    f = plan.func
    T = c_type(f.typ)
    nd = len(plan.fvars)
    nthreads = 256
    # init: one thread per target point
    I = []
    wi = lambda t, ind=1: I.append('  ' * ind + t)
    npts = ' * '.join('(long long)%s' % e for e in exts) or '1'
    wi('long long p = (long long)pmx_tile * %d + threadIdx.x;' % nthreads)
    wi('if (p >= %s) return;' % npts)
    wi('const long long lin = p;')
    for d in reversed(range(nd)):
        wi('const int f%d = (int)(p %% %s) + (%s); p /= %s;' % (d, exts[d], lbs[d], exts[d]))

    def cb(ref, pr, pro):
        g = ref.objectRef
        lin = None
        for d, arg in enumerate(ref.arguments):
            q = pr.expr(arg, pro)
            t = '(%s - (%s))' % (q, base.expr(g.domain[d].lowerBound))
            lin = t if lin is None else '(%s * %s + %s)' % (lin, _extent_c(g, d, base), t)
        return 'G_%s[%s]' % (g.name, lin or '0')
    dpr = Printer(dict((plan.fvars[d], 'f%d' % d) for d in range(nd)), cb, mode, lazy=True)
    pro = []
    dv = dpr.expr(f.default, pro)
    for l in pro:
        wi(l)
    wi('G_%s[lin] = (%s)(%s);' % (f.name, T, dv))
    init_name = kname + '_init'
    init_info = KernelInfo(init_name, 'reduction', [f], '((%s + %d) / %d)' % (npts, nthreads - 1, nthreads),
                           nthreads, 0, plan, [], [f])
    out = [(wrap_kernel(init_name, params, [f], [f], nthreads, 0, I), init_info)]
    # accumulation: one thread per reduction point
    del L[:]
    w = lambda t, ind=1: L.append('  ' * ind + t)
    rext = ['((%s) - (%s) + 1)' % (base.expr(iv.upperBound), base.expr(iv.lowerBound))
            for iv in f.reductionDomain]
    nred = ' * '.join('(long long)%s' % e for e in rext) or '1'
    w('long long p = (long long)pmx_tile * %d + threadIdx.x;' % nthreads)
    w('if (p >= %s) return;' % nred)
    for j in reversed(range(len(plan.rvars))):
        w('const int r%d = (int)(p %% %s) + (%s); p /= %s;' % (
            j, rext[j], base.expr(f.reductionDomain[j].lowerBound), rext[j]))
    tl = None
    for d in range(nd):
        t = '(long long)(q%d - (%s))' % (d, lbs[d])
        tl = t if tl is None else '(%s * %s + %s)' % (tl, exts[d], t)
    slot = 'G_%s[%s]' % (f.name, tl or '0')

    def update(op, val):
        v = '(%s)(%s)' % (T, val)
        if op in _ATOMIC:
            return '%s(&%s, %s);' % (_ATOMIC[op], slot, v)
        # product: compare-and-swap until no other thread intervened
        return ('{ %s* pa = &%s; const %s m = %s; %s old = *pa, seen; '
                'do { seen = old; old = atomicCAS(pa, seen, (%s)(seen * m)); } while (old != seen); }'
                % (T, slot, T, v, T, T))
    red_body(1, lambda: None, update)
    info = KernelInfo(kname, 'reduction', buffers, '((%s + %d) / %d)' % (nred, nthreads - 1, nthreads),
                      nthreads, 0, plan, reads, [f])
    out.append((wrap_kernel(kname, params, buffers, [f], nthreads, 0, L), info))
    return out
