"""C++/OpenMP backend: the reference's code generator, restated for the
host, from the same plans the HIP emitter uses.

Reference: sandbox/codegen.py:411-551 (``generate_c_naive_from_isl_ast``:
an OpenMP-parallel loop over the overlapped tiles of a group, per-thread
scratchpads declared inside it, one loop nest per stage in level order),
codegen.py:268-372 and 553-715 (point and expression printing),
codegen.py:844-1067 (live-out arrays, the group sequence, the signature)
and targetc.py (types, literals).  The options of compiler.py:59-68 map as
follows: ``optimize_storage`` and ``early_free`` are what the group plans
do anyway (scratch slots reused by (type, rank) class across phases, as on
the GPU; live-outs freed after their last reading group); ``pool_alloc``
keeps the live-out arrays between calls in a size-keyed pool, like
sandbox/memory_allocation/simple_pool_allocator.cpp; ``multipar`` and
``flatten_scratchpad`` need no switch (tiles are flattened into one
parallel loop; scratch is one flat per-thread buffer).

Per group the generated function runs

    #pragma omp parallel            (one scratch buffer per thread)
    #pragma omp for schedule(static)
    for every tile of the group's anchor:
        the tile's region program (the GPU kernels' prologue)
        for every phase, for every stored stage of it:
            for every point of its region (innermost loop ``omp simd``):
                value (inlined producers substituted), scratch and/or
                live-out store

Reductions run in the reference's loop order (one thread; the operator is
honoured), constant tables are built on the host like the GPU stub's, and
"const" functions that read images are plain loops.  Values are those of
the reference's C semantics, like the HIP path, because the same printer
emits both and the library is built without floating-point contraction.

The library exports ``void pipeline_<name>(params, void* inputs, void*
outputs)`` (the reference's entry point) and ``int pipeline_<name>_host``
(the same, returning 0 or an error code; message in
``pipeline_<name>_error()``).
"""

from .analysis import classify_access, unconditional_refs
from .constructs import Reference
from .cprint import Printer, c_type
from .group_plan import GroupPlan
from .reduction_plan import ReductionPlan, ConstPlan
from .region_ir import bound_c

CPU_FLAGS = ['-O3', '-fopenmp', '-march=native', '-ffp-contract=off', '-fno-fast-math',
             '-std=c++17', '-fPIC', '-shared', '-w']

HOST_PRELUDE = r'''
#include <algorithm>
#include <cmath>
#include <math.h>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <set>
#include <string>
#include <vector>
#include <omp.h>
#define PMX_HOST_MATH
static inline int pmx_abs(int v) { return std::abs(v); }
static inline long long pmx_abs(long long v) { return std::llabs(v); }
static inline float pmx_abs(float v) { return std::abs(v); }
static inline double pmx_abs(double v) { return std::abs(v); }
static inline float pmx_exp(float v) { return std::exp(v); }
static inline double pmx_exp(double v) { return std::exp(v); }
static inline double pmx_exp(int v) { return std::exp((double)v); }
static inline float pmx_sqrt(float v) { return std::sqrt(v); }
static inline double pmx_sqrt(double v) { return std::sqrt(v); }
static inline double pmx_sqrt(int v) { return std::sqrt((double)v); }
static inline float pmx_sin(float v) { return std::sin(v); }
static inline double pmx_sin(double v) { return std::sin(v); }
static inline double pmx_sin(int v) { return std::sin((double)v); }
static inline float pmx_cos(float v) { return std::cos(v); }
static inline double pmx_cos(double v) { return std::cos(v); }
static inline double pmx_cos(int v) { return std::cos((double)v); }
static inline int pmx_imin(int a, int b) { return a < b ? a : b; }
static inline int pmx_imax(int a, int b) { return a > b ? a : b; }
static inline int pmx_clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
static inline int pmx_fdiv(int a, int k) { return a >= 0 ? a / k : -((-a + k - 1) / k); }
static inline double pmx_divc(double n, double r, double d) {
  const double q = n * r;
  return std::fma(std::fma(-q, d, n), r, q);
}
#define PMX_BIG 536870912
'''


def _prod(xs):
    n = 1
    for x in xs:
        n *= x
    return n


def _extent_c(f, d, pr):
    return '(%s - %s + 1)' % (pr.expr(f.domain[d].upperBound), pr.expr(f.domain[d].lowerBound))


def _size_c(f, pr):
    return ' * '.join('(long long)%s' % _extent_c(f, d, pr) for d in range(len(f.domain))) or '1LL'


def _gindex(f, idx, pr):
    e = None
    for d, q in enumerate(idx):
        t = '(long long)((%s) - (%s))' % (q, pr.expr(f.domain[d].lowerBound))
        e = t if e is None else '(%s * %s + %s)' % (e, _extent_c(f, d, pr), t)
    return e or '0'


class _Emitter(object):
    def __init__(self, pipe):
        self.pipe = pipe
        self.mode = pipe.literal_mode
        self.base = Printer({}, None, self.mode)
        self.L = []

    def w(self, s, ind=1):
        self.L.append('  ' * ind + s)

    # -- one stencil group ---------------------------------------------------
    def group(self, plan, ind):
        pipe, mode, base = self.pipe, self.mode, self.base
        w = lambda s, i=0: self.w(s, ind + i)
        stages = plan.stages
        ir = plan.ir
        R = lambda i: 'r%d' % i
        box_ids = {}
        for s in stages:
            for pi, part in enumerate(s.parts):
                for i, (blo, bhi) in enumerate(part.box):
                    box_ids[(s, pi, i)] = (ir.expr(blo) if blo is not None else None,
                                           ir.expr(bhi) if bhi is not None else None)
        prl = ir.to_c(lambda vn: Printer(vn, None, mode))
        uncond = set()
        for s in stages:
            for part in s.parts:
                keys = set()
                for r in unconditional_refs(part.expr):
                    uncond.add(id(r))
                    keys.add((id(r.objectRef), tuple(str(a) for a in r.arguments)))
                for c in part.pred:
                    for r in c.collect(Reference):
                        uncond.add(id(r))
                for r in part.expr.collect(Reference):
                    if (id(r.objectRef), tuple(str(a) for a in r.arguments)) in keys:
                        uncond.add(id(r))

        def lds_index(p, idx):
            ext = plan.lds_ext(p)
            e = None
            for d, q in enumerate(idx):
                t = '((%s) - %s)' % (q, R(p.region[d][0]))
                e = t if e is None else '(%s * %d + %s)' % (e, ext[d], t)
            return e or '0'

        def make_cb(cur):
            def cb(ref, pr, pro):
                f = ref.objectRef
                cvars = cur.func.variables
                idx = []
                for d, arg in enumerate(ref.arguments):
                    acc = classify_access(arg, cvars)
                    q = pr.expr(arg, pro)
                    if acc.kind == '*' or (id(ref), d) in plan.clamp_refs or \
                            (id(ref) not in uncond and not plan.in_group(f)):
                        if plan.in_group(f):
                            p = plan.stage_of(f)
                            q = 'pmx_clampi(%s, %s, %s)' % (q, R(p.region[d][0]), R(p.region[d][1]))
                        else:
                            q = 'pmx_clampi(%s, %s, %s)' % (q, base.expr(f.domain[d].lowerBound),
                                                            base.expr(f.domain[d].upperBound))
                    idx.append(q)
                if plan.in_group(f):
                    p = plan.stage_of(f)
                    if p.inline:
                        return 'v_' + p.name
                    return 'L_%s[%s]' % (p.name, lds_index(p, idx))
                return 'G_%s[%s]' % (f.name, _gindex(f, idx, base))
            return cb

        def inline_closure(s):
            out, seen = [], set()

            def visit(t):
                for part in t.parts:
                    for r in plan._part_refs(part):
                        if plan.in_group(r.objectRef):
                            p = plan.stage_of(r.objectRef)
                            if p.inline and p not in seen:
                                seen.add(p)
                                visit(p)
                                out.append(p)
            visit(s)
            return out

        def stage_value(s, target, pnames, i0, skip_single):
            vn = dict(zip(s.func.variables, pnames))
            pr = Printer(vn, make_cb(s), mode, lazy=True)
            single = len(s.parts) == 1
            for pi, part in enumerate(s.parts):
                conds = []
                if not (skip_single and single):
                    for i in range(s.nd):
                        lo_id, hi_id = box_ids[(s, pi, i)]
                        if lo_id is not None:
                            conds.append('%s >= %s' % (pnames[i], R(lo_id)))
                        if hi_id is not None:
                            conds.append('%s <= %s' % (pnames[i], R(hi_id)))
                w('if (%s) {' % (' && '.join(conds) or 'true'), i0)
                inner = i0 + 1
                if part.pred:
                    pro = []
                    pc = ' && '.join(pr.cond(c, pro) for c in part.pred)
                    for l in pro:
                        w(l, inner)
                    w('if (%s) {' % pc, inner)
                    inner += 1
                pro = []
                val = pr.expr(part.expr, pro)
                for l in pro:
                    w(l, inner)
                w('%s = %s; pmx_w = true;' % (target, val), inner)
                if part.pred:
                    w('}', inner - 1)
                w('}', i0)

        # global array shapes
        anchor = plan.anchor
        grid = plan.grid_dims
        w('{ // group %d (%s)' % (plan.gindex, ', '.join(s.name for s in stages)))
        for (d, T) in grid:
            w('const int n%d = ((%s) + %d) / %d;' % (d, _extent_c(anchor.func, d, base), T - 1, T), 1)
        ntiles = ' * '.join('(long long)n%d' % d for (d, T) in grid) or '1LL'
        lds = max(64, (plan.lds_bytes + 63) // 64 * 64)
        w('#pragma omp parallel', 1)
        w('{', 1)
        w('unsigned char* pmx_lds = (unsigned char*)aligned_alloc(64, %d);' % lds, 2)
        w('#pragma omp for schedule(static)', 2)
        w('for (long long pmx_t = 0; pmx_t < %s; ++pmx_t) {' % ntiles, 2)
        w('long long pmx_b = pmx_t;', 3)
        for (d, T) in reversed(grid):
            w('const int t%d = (int)(pmx_b %% n%d); pmx_b /= n%d;' % (d, d, d), 3)
        for l in prl:
            w(l, 3)
        over = []
        for s in stages:
            for i in range(s.nd):
                over.append('(%s - %s + 1 > %d)' % (R(s.region[i][1]), R(s.region[i][0]), s.static_ext[i]))
        if over:
            w('if (%s) { __atomic_store_n(&pmx_overflow, 1, __ATOMIC_RELAXED); continue; }'
              % ' || '.join(over), 3)
        for s in stages:
            if s.needs_lds:
                w('%s* L_%s = (%s*)(pmx_lds + %d);' % (c_type(s.func.typ), s.name, c_type(s.func.typ),
                                                      s.lds_offset), 3)
        for ph in range(plan.nphases):
            for s in stages:
                if s.inline or s.phase != ph:
                    continue
                i0 = 3
                qn = ['q%d' % i for i in range(s.nd)]
                for i in range(s.nd):
                    if i == s.nd - 1:
                        w('#pragma omp simd', i0)
                    w('for (int q%d = %s; q%d <= %s; ++q%d) {' % (i, R(s.region[i][0]), i,
                                                                 R(s.region[i][1]), i), i0)
                    i0 += 1
                w('bool pmx_w = false;', i0)
                for t in inline_closure(s):
                    T = c_type(t.func.typ)
                    w('%s v_%s = (%s)0;' % (T, t.name, T), i0)
                    w('{ bool pmx_w = false;', i0)
                    stage_value(t, 'v_' + t.name, qn, i0 + 1, False)
                    w('(void)pmx_w; }', i0)
                T = c_type(s.func.typ)
                w('%s v = (%s)0;' % (T, T), i0)
                stage_value(s, 'v', qn, i0, True)
                w('if (pmx_w) {', i0)
                if s.needs_lds:
                    w('L_%s[%s] = v;' % (s.name, lds_index(s, qn)), i0 + 1)
                if s.liveout:
                    st = 'G_%s[%s] = v;' % (s.func.name, _gindex(s.func, qn, base))
                    if s is not anchor:
                        oc = ' && '.join('q%d >= %s && q%d <= %s' % (i, R(lo), i, R(hi))
                                         for i, (lo, hi) in enumerate(s.own))
                        w('if (%s) %s' % (oc, st), i0 + 1)
                    else:
                        w(st, i0 + 1)
                w('}', i0)
                for i in range(s.nd):
                    i0 -= 1
                    w('}', i0)
        w('}', 2)
        w('free(pmx_lds);', 2)
        w('}', 1)
        w('}')

    # -- reductions: the reference's loop order --------------------------------
    def reduction(self, plan, ind):
        from .constructs import Op
        f = plan.func
        base, mode = self.base, self.mode
        w = lambda s, i=0: self.w(s, ind + i)
        nd = len(plan.fvars)
        T = c_type(f.typ)

        def cb(ref, pr, pro):
            g = ref.objectRef
            idx = []
            for d, arg in enumerate(ref.arguments):
                q = pr.expr(arg, pro)
                if classify_access(arg, plan.rvars).kind == '*':
                    q = 'pmx_clampi(%s, %s, %s)' % (q, base.expr(g.domain[d].lowerBound),
                                                    base.expr(g.domain[d].upperBound))
                idx.append(q)
            return 'G_%s[%s]' % (g.name, _gindex(g, idx, base))
        w('{ // reduction %s' % f.name)
        i0 = 1
        for d in range(nd):
            w('for (int f%d = %s; f%d <= %s; ++f%d) {' % (d, base.expr(f.domain[d].lowerBound), d,
                                                         base.expr(f.domain[d].upperBound), d), i0)
            i0 += 1
        dpr = Printer(dict((plan.fvars[d], 'f%d' % d) for d in range(nd)), cb, mode, lazy=True)
        pro = []
        dv = dpr.expr(f.default, pro)
        for l in pro:
            w(l, i0)
        w('G_%s[%s] = %s;' % (f.name, _gindex(f, ['f%d' % d for d in range(nd)], base), dv), i0)
        for d in range(nd):
            i0 -= 1
            w('}', i0)
        rnames = ['r%d' % j for j in range(len(plan.rvars))]
        for j in range(len(plan.rvars)):
            w('for (int r%d = %s; r%d <= %s; ++r%d) {' % (
                j, base.expr(f.reductionDomain[j].lowerBound), j,
                base.expr(f.reductionDomain[j].upperBound), j), i0)
            i0 += 1
        pr = Printer(dict(zip(plan.rvars, rnames)), cb, mode, lazy=True)
        for rc in plan.cases:
            red = rc.red
            w('{', i0)
            pro = []
            tn = []
            for d in range(nd):
                q = pr.expr(red.accumulate_ref.arguments[d], pro)
                if classify_access(red.accumulate_ref.arguments[d], plan.rvars).kind == '*':
                    q = 'pmx_clampi(%s, %s, %s)' % (q, base.expr(f.domain[d].lowerBound),
                                                    base.expr(f.domain[d].upperBound))
                tn.append(q)
            val = pr.expr(red.expression, pro)
            for l in pro:
                w(l, i0 + 1)
            cond = 'true'
            if rc.cond is not None:
                vn = dict(zip(plan.rvars, rnames))
                for v, d in zip(plan.fvars, range(nd)):
                    if v not in plan.rvars:
                        vn[v] = '(%s)' % tn[d]
                cpro = []
                cond = Printer(vn, cb, mode, lazy=True).cond(rc.cond, cpro)
                for l in cpro:
                    w(l, i0 + 1)
            slot = 'G_%s[%s]' % (f.name, _gindex(f, tn, base))
            op = red.op_type
            if op == Op.Sum:
                upd = '%s = (%s + %s);' % (slot, slot, val)
            elif op == Op.Mul:
                upd = '%s = (%s * %s);' % (slot, slot, val)
            elif op == Op.Min:
                upd = '%s = ((%s < %s)? (%s)(%s): %s);' % (slot, val, slot, T, val, slot)
            else:
                upd = '%s = ((%s > %s)? (%s)(%s): %s);' % (slot, val, slot, T, val, slot)
            w('if (%s) { %s }' % (cond, upd), i0 + 1)
            w('}', i0)
        for j in range(len(plan.rvars)):
            i0 -= 1
            w('}', i0)
        w('}')

    # -- "const" functions that read images: one loop nest ---------------------
    def const_loops(self, plan, ind):
        f = plan.func
        base, mode = self.base, self.mode
        w = lambda s, i=0: self.w(s, ind + i)
        nd = len(f.variables)
        names = ['p%d' % i for i in range(nd)]

        def cb(ref, pr, pro):
            g = ref.objectRef
            idx = [pr.expr(a, pro) for a in ref.arguments]
            return 'G_%s[%s]' % (g.name, _gindex(g, idx, base))
        pr = Printer(dict(zip(f.variables, names)), cb, mode, lazy=True)
        w('{ // const %s' % f.name)
        i0 = 1
        for i in range(nd):
            w('for (int %s = %s; %s <= %s; ++%s) {' % (names[i], base.expr(f.domain[i].lowerBound), names[i],
                                                      base.expr(f.domain[i].upperBound), names[i]), i0)
            i0 += 1
        for part in plan.parts:
            conds = []
            for i, (blo, bhi) in enumerate(part.box):
                if blo is not None:
                    conds.append('%s >= %s' % (names[i], bound_c(blo, base)))
                if bhi is not None:
                    conds.append('%s <= %s' % (names[i], bound_c(bhi, base)))
            pro = []
            pc = [pr.cond(c, pro) for c in part.pred]
            val = pr.expr(part.expr, pro)
            w('if (%s) {' % (' && '.join(conds) or 'true'), i0)
            for l in pro:
                w(l, i0 + 1)
            w('if (%s) G_%s[%s] = %s;' % (' && '.join(pc) or 'true', f.name,
                                          _gindex(f, names, base), val), i0 + 1)
            w('}', i0)
        for i in range(nd):
            i0 -= 1
            w('}', i0)
        w('}')


def emit_cpu(pipe):
    """C++/OpenMP source of ``pipe`` (see the module docstring)."""
    from .hip_emit import plan_pipeline
    from .host_stub import _validation_lines
    plans = plan_pipeline(pipe)
    em = _Emitter(pipe)
    base = em.base
    name = pipe.name
    params = pipe.sorted_parameters()
    inputs = pipe.sorted_inputs()
    outputs = pipe.sorted_outputs()
    head = ['// C++/OpenMP code of pipeline %s (generated by polymage_amd.cpu_emit)' % name, HOST_PRELUDE]
    head.append('static thread_local std::string pmx_err_msg;')
    head.append('static void pmx_set_error(const char* m) { pmx_err_msg = m ? m : ""; }')
    head.append('extern "C" const char* pipeline_%s_error(void) { return pmx_err_msg.c_str(); }' % name)
    # reference pool (simple_pool_allocator.cpp:27-56): arrays between groups
    # kept between calls, matched by size
    head.append('static std::mutex pmx_pool_mu;')
    head.append('static std::multimap<size_t, void*> pmx_pool;')
    head.append('static void* pmx_pool_get(size_t n) { std::lock_guard<std::mutex> g(pmx_pool_mu); '
                'auto it = pmx_pool.find(n); if (it != pmx_pool.end()) { void* p = it->second; '
                'pmx_pool.erase(it); return p; } return aligned_alloc(64, (n + 63) / 64 * 64); }')
    head.append('static void pmx_pool_put(size_t n, void* p) { std::lock_guard<std::mutex> g(pmx_pool_mu); '
                'pmx_pool.insert(std::make_pair(n, p)); }')
    head.append('extern "C" void pool_init(void) {}')
    head.append('extern "C" void pool_destroy(void) { std::lock_guard<std::mutex> g(pmx_pool_mu); '
                'for (auto& kv : pmx_pool) free(kv.second); pmx_pool.clear(); }')
    for plan in plans:
        if isinstance(plan, ConstPlan) and plan.on_host:
            head.extend(plan.host_code(base, host_only=True))
    sig = ['%s %s' % (c_type(p.typ), p.name) for p in params] + \
        ['void* %s_void_arg' % f.name for f in inputs + outputs]
    em.L.append('extern "C" int pipeline_%s_host(%s) {' % (name, ', '.join(sig)))
    for l in _validation_lines(pipe, base):
        em.w(l)
    em.w('int pmx_overflow = 0;')
    # every array a group reads or writes
    for f in inputs + outputs:
        em.w('%s* G_%s = (%s*)%s_void_arg;' % (c_type(f.typ), f.name, c_type(f.typ), f.name))
    have = set(f.name for f in inputs + outputs)
    temps = []
    for plan in plans:
        f = plan.func if not isinstance(plan, GroupPlan) else None
        funcs = [f] if f is not None else [s.func for s in plan.stages if s.liveout]
        for g in funcs:
            if g.name in have:
                continue
            have.add(g.name)
            T = c_type(g.typ)
            if isinstance(plan, ConstPlan) and plan.on_host:
                em.w('const %s* h_%s = nullptr; size_t n_%s = 0;' % (T, g.name, g.name))
                em.w('{ int rc = %s_host(%s%s&h_%s, &n_%s, nullptr); if (rc) return rc; }' % (
                    plan.upload_fn, ', '.join(p.name for p in plan.params), ', ' if plan.params else '',
                    g.name, g.name))
                em.w('const %s* G_%s = h_%s;' % (T, g.name, g.name))
                continue
            em.w('const size_t z_%s = (size_t)(%s) * sizeof(%s);' % (g.name, _size_c(g, base), T))
            em.w('%s* G_%s = (%s*)pmx_pool_get(z_%s);' % (T, g.name, T, g.name))
            temps.append(g)
    for plan in plans:
        if isinstance(plan, GroupPlan):
            em.group(plan, 1)
        elif isinstance(plan, ReductionPlan):
            em.reduction(plan, 1)
        elif not plan.on_host:
            em.const_loops(plan, 1)
    for g in temps:
        em.w('pmx_pool_put(z_%s, G_%s);' % (g.name, g.name))
    em.w('if (pmx_overflow) { pmx_set_error("%s: a tile exceeded its planned scratch box; '
         'rebuild with estimates closer to these sizes"); return -34; }' % name)
    em.w('return 0;')
    em.L.append('}')
    em.L.append('extern "C" void pipeline_%s(%s) {' % (name, ', '.join(sig)))
    em.w('int rc = pipeline_%s_host(%s);' % (name, ', '.join(
        [p.name for p in params] + ['%s_void_arg' % f.name for f in inputs + outputs])))
    em.w('if (rc) fprintf(stderr, "pipeline_%s: %%s\\n", pmx_err_msg.c_str());' % name)
    em.L.append('}')
    return '\n'.join(head + em.L) + '\n'
