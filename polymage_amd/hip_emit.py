"""HIP emission: one kernel per group, one workgroup per overlapped tile,
plus the C++ host stub exporting the reference ABI.

Reference counterpart: generate_code_for_pipeline / generate_code_for_group
(sandbox/codegen.py:844-1067) emit one OpenMP-parallel tile loop per group
with per-thread scratchpads.  Here a group is a ``__global__`` function:

* the workgroup's tile comes from ``blockIdx.x`` (optionally remapped so the
  tiles of one XCD are contiguous and share halo lines in that XCD's L2);
* a scalar prologue computes every stage's region for the tile
  (``region_ir``);
* each materialised stage is one sweep over its region by all threads,
  consecutive threads on consecutive innermost points (coalesced HBM
  reads, conflict-free LDS rows), writing LDS and/or its global live-out;
  inlined stages are computed in registers inside their consumers;
* workgroup barriers separate phases.

The host stub (compiled by g++, embedding the hiprtc code object) exports
``pipeline_<name>(params, inputs, outputs)`` with host pointers, like the
reference, and ``pipeline_<name>_dev(..., stream)`` on device pointers.
"""

import hashlib
import re

from .analysis import PlanError, classify_access, eval_bexpr, unconditional_refs, shift_access, scaled_access
from .constructs import Reference, Image, Case, Reduction
from .cprint import Printer, DEVICE_PRELUDE, c_type
from .group_plan import GroupPlan
from .region_ir import REGION_PRELUDE, bound_c
from .reduction_plan import ReductionPlan, emit_reduction_kernel, ConstPlan, emit_const_kernel


def _prod(xs):
    n = 1
    for x in xs:
        n *= x
    return n


# separators of the two parts in the printed source (runtime.split_source)
HIP_MARK = '// ===== HIP kernels (hiprtc, gfx950) =====\n'
HOST_MARK = '\n// ===== host stub (g++) =====\n'


class GeneratedCode(object):
    """What ``Pipeline.generate_code`` returns; ``str()`` is the source
    (the HIP kernels followed by the host stub), as in pipe.py:934-958."""

    def __init__(self, name, hip_src, host_src, kernels, meta):
        self.name = name
        self.hip_src = hip_src
        self.host_src = host_src
        self.kernels = kernels
        self.meta = meta

    def __str__(self):
        return HIP_MARK + self.hip_src + HOST_MARK + self.host_src

    @property
    def key(self):
        return hashlib.sha256((self.hip_src + '\0' + self.host_src).encode()).hexdigest()[:20]


class KernelInfo(object):
    def __init__(self, name, kind, buffers, grid_c, threads, lds, plan, reads, writes):
        self.name = name
        self.kind = kind            # 'group' | 'reduction' | 'const'
        self.buffers = buffers      # [func] in argument order
        self.grid_c = grid_c        # C expression (host, params in scope) -> #workgroups
        self.threads = threads
        self.lds = lds
        self.plan = plan
        self.reads = reads
        self.writes = writes


# ---------------------------------------------------------------------------
# global array helpers
# ---------------------------------------------------------------------------

def _extent_c(f, d, pr):
    return '(%s - %s + 1)' % (pr.expr(f.domain[d].upperBound), pr.expr(f.domain[d].lowerBound))


def _gname(f):
    return 'G_' + f.name


class _KernelWriter(object):
    def __init__(self, pipe, mode, index_t='int'):
        self.pipe = pipe
        self.mode = mode
        self.index_t = index_t
        self.lines = []

    def w(self, s, ind=1):
        self.lines.append('  ' * ind + s)


def emit_group_kernel(pipe, plan, kname, threads, mode, xcd=True):
    """HIP source of one stencil-group kernel."""
    params = pipe.sorted_parameters()
    stages = plan.stages
    # global buffers: out-of-group producers read, live-outs written
    reads = []
    for s in stages:
        for part in s.parts:
            for r in plan._part_refs(part):
                f = r.objectRef
                if not plan.in_group(f) and f not in reads:
                    reads.append(f)
    writes = [s.func for s in stages if s.liveout]
    buffers = sorted(reads, key=lambda f: f.name) + sorted(writes, key=lambda f: f.name)
    index_t = 'long long' if plan.opts.get('index64') else 'int'

    base_pr = Printer({}, None, mode)
    L = []
    w = lambda s, ind=1: L.append('  ' * ind + s)

    # tile indices
    total = plan.total_tiles
    w('int pmx_b = pmx_tile;')
    a = plan.anchor
    grid = plan.grid_dims
    # tile counts (runtime).  With rolling (pmx_rr > 1) a workgroup takes
    # pmx_rr consecutive tiles along the first tiled dimension, one per
    # call (pmx_rk = 0..pmx_rr-1): the halo rows a tile shares with the
    # previous one, and its own inputs read again by later phases, are
    # then recent L2 lines.
    roll_d = next((d for (d, T) in grid if T > 1), None)
    for (d, T) in grid:
        w('const int n%d = ((%s) + %d) / %d;' % (d, _extent_c(a.func, d, base_pr), T - 1, T))
    nvar = dict((d, 'n%d' % d) for (d, T) in grid)
    if roll_d is not None:
        w('const int nr%d = (n%d + pmx_rr - 1) / pmx_rr;' % (roll_d, roll_d))
        nvar[roll_d] = 'nr%d' % roll_d
    ntot = ' * '.join(nvar[d] for (d, T) in grid) or '1'
    w('const int pmx_ntiles = %s;' % ntot)
    if xcd:
        # blocks are dealt round-robin to the 8 XCDs; give each XCD a
        # contiguous run of tiles so neighbouring halos hit the same L2
        w('if (pmx_remap) { const int per = pmx_ntiles / 8, rem = pmx_ntiles % 8;')
        w('  const int xcd = pmx_b % 8, k = pmx_b / 8;')
        w('  pmx_b = (xcd < rem ? xcd * (per + 1) : rem * (per + 1) + (xcd - rem) * per) + k; }')
    fast = [g for g in grid if g[0] in getattr(plan, 'split_fast', ())]
    for (d, T) in fast + [g for g in reversed(grid) if g not in fast]:
        if d == roll_d:
            w('const int t%d = (pmx_b %% nr%d) * pmx_rr + pmx_rk; pmx_b /= nr%d;' % (d, d, d))
            w('if (t%d >= n%d) return;' % (d, d))
        else:
            w('const int t%d = pmx_b %% n%d; pmx_b /= n%d;' % (d, d, d))

    # extra prologue values: part boxes of every stage
    ir = plan.ir
    box_ids = {}
    for s in stages:
        for pi, part in enumerate(s.parts):
            for i, (blo, bhi) in enumerate(part.box):
                box_ids[(s, pi, i)] = (ir.expr(blo) if blo is not None else None,
                                       ir.expr(bhi) if bhi is not None else None)
    prl = ir.to_c(lambda vn: Printer(vn, None, mode))
    for l in prl:
        w(l)
    R = lambda i: 'r%d' % i

    # global strides
    gfun = buffers
    for f in gfun:
        nd = len(f.domain)
        for d in range(nd):
            w('const %s E_%s_%d = %s;' % (index_t, f.name, d, _extent_c(f, d, base_pr)))
            w('const int LB_%s_%d = %s;' % (f.name, d, base_pr.expr(f.domain[d].lowerBound)))
        for d in range(nd - 1):
            w('const %s S_%s_%d = %s;' % (index_t, f.name, d, ' * '.join(
                'E_%s_%d' % (f.name, j) for j in range(d + 1, nd))))

    def gindex(f, idx_strs):
        e = None
        for d, q in enumerate(idx_strs):
            term = '(%s)(%s - LB_%s_%d)' % (index_t, q, f.name, d)
            e = term if e is None else '(%s * E_%s_%d + %s)' % (e, f.name, d, term)
        return e if e is not None else '0'

    def lds_index(p, idx_strs):
        e = None
        for d, q in enumerate(idx_strs):
            term = '(%s - %s)' % (q, R(p.region[d][0]))
            e = term if e is None else '(__mul24(%s, %d) + %s)' % (e, plan.lds_ext(p)[d], term)
        return e if e is not None else '0'

    # references that only run inside a Select branch: their global reads
    # are clamped to the producer's domain (the run-time bounds check in
    # the host stub covers unconditional references only)
    # A branch reference naming the same point as an unconditional
    # reference of the same part reads an address that is read anyway, so
    # it needs no clamp either.
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

    def make_ref_cb(cur_stage, item=None):
        def cb(ref, pr, pro):
            f = ref.objectRef
            cvars = cur_stage.func.variables
            # a reference merged from the two branches of a select
            # (SelectUnifier) is clamped where either of them is
            rids = [id(o) for o in getattr(ref, 'unified_from', (ref,))]
            cl = lambda d: any((i, d) in plan.clamp_refs for i in rids)
            clamped = any(cl(d) for d in range(len(ref.arguments)))
            if item is not None and not clamped and \
                    not (plan.in_group(f) and plan.stage_of(f).inline) and \
                    (plan.in_group(f) or id(ref) in uncond):
                sh = [shift_access(arg, cvars) for arg in ref.arguments]
                if all(x is not None for x in sh):
                    return item.ref(f, sh, pr)
                if plan.in_group(f) and not plan.stage_of(f).lane:
                    # downsampling reads of an LDS stage (f(2 * x + 1)):
                    # per-item base, constant offsets per point
                    sc = [scaled_access(arg, cvars) for arg in ref.arguments]
                    if all(x is not None for x in sc):
                        return item.ref_scaled(f, sc, pr)
            idx = []
            for d, arg in enumerate(ref.arguments):
                acc = classify_access(arg, cvars)
                q = pr.expr(arg, pro)
                if acc.kind == '*' or cl(d) or \
                        (id(ref) not in uncond and not plan.in_group(f)):
                    if plan.in_group(f):
                        p = plan.stage_of(f)
                        q = 'pmx_clampi(%s, %s, %s)' % (q, R(p.region[d][0]), R(p.region[d][1]))
                    else:
                        q = 'pmx_clampi(%s, LB_%s_%d, LB_%s_%d + (int)E_%s_%d - 1)' % (
                            q, f.name, d, f.name, d, f.name, d)
                idx.append(q)
            if plan.in_group(f):
                p = plan.stage_of(f)
                if p.inline:
                    return 'v_' + p.name
                return 'L_%s[%s]' % (p.name, lds_index(p, idx))
            return '%s[%s]' % (_gname(f), gindex(f, idx))
        return cb

    class Item(object):
        """Addressing of one per-thread block: point = item base ``pb`` +
        offset ``d`` per stage dimension.  A reference whose arguments are
        consumer variables plus parameter/constant offsets becomes a per-item
        base (pointer or LDS index, computed once) plus an offset that is a
        compile-time constant or uniform, so neighbouring points share the
        address arithmetic and their loads."""

        def __init__(self, pb, d):
            self.pb = pb
            self.d = d
            self.decls = []
            self.names = {}
            # load forwarding between aligned sweeps: global reads at a
            # compile-time point of the block are recorded (key -> load)
            # or, for a later sweep, replaced by the register holding them
            self.record = None
            self.forward = None
            self.lane_keys = None       # lane-fused reads this item makes
            # 128-bit LDS row loads of a 4-wide block: a dry run collects the
            # columns read per producer row (vprobe); the emitted fast path
            # then reads components of the row registers (vmap)
            self.vcd = None
            self.vprobe = None
            self.vmap = None

        def _base(self, key, make):
            if key not in self.names:
                nm = 'ib%d' % len(self.names)
                self.names[key] = nm
                self.decls.append(make(nm))
            return self.names[key]

        def _off(self, parts, pr):
            """Sum of (offset, stride) pairs as C text; integer offsets with
            integer strides fold into one constant."""
            terms = []
            const = 0
            for (o, stride) in parts:
                if isinstance(o, int):
                    if o == 0:
                        continue
                    if isinstance(stride, int):
                        const += o * stride
                    else:
                        terms.append('(%d * %s)' % (o, stride))
                elif stride == 1:
                    terms.append('(%s)' % o)
                elif isinstance(stride, int):
                    # LDS strides and offsets are small: a 24-bit multiply
                    # (full rate; a 32-bit one issues at a quarter)
                    terms.append('__mul24(%s, %d)' % (o, stride))
                else:
                    terms.append('((%s) * %s)' % (o, stride))
            if const or not terms:
                terms.append('%d' % const)
            return ' + '.join(terms)

        def _dsum(self, k, off, pr):
            d = self.d[k]
            if isinstance(off, int):
                if isinstance(d, int):
                    return d + off
                return '%s + %d' % (d, off) if off else d
            o = pr.expr(off)
            return '%s + %s' % (d, o) if not (isinstance(d, int) and d == 0) else o

        def ref(self, f, sh, pr):
            kinds = tuple(x[0] for x in sh)
            parts = []
            if plan.in_group(f) and plan.stage_of(f).lane:
                # a lane-fused producer: this thread's registers (its column)
                # or a neighbouring lane's (other columns)
                p = plan.stage_of(f)
                lp = plan.lane_pair
                rd_, cd_ = p.bdims
                row = self.d[rd_] + sh[rd_][1] + lp[2]
                dc = sh[cd_][1]
                nm = 'lv_%s_%d' % (p.name, row)
                if dc:
                    nm = 'lx_%s_%d_%s%d' % (p.name, row, 'm' if dc < 0 else 'p', abs(dc))
                if self.lane_keys is not None:
                    self.lane_keys.add((p.name, row, dc))
                return nm
            if plan.in_group(f):
                p = plan.stage_of(f)
                ext = plan.lds_ext(p)
                W = [_prod(ext[j + 1:]) for j in range(len(ext))]
                nm = self.lds_base(p, kinds)
                rows = []
                for j, (k, off) in enumerate(sh):
                    o = self._dsum(k, off, pr) if k != 'const' else \
                        (off if isinstance(off, int) else pr.expr(off))
                    parts.append((o, W[j]))
                    rows.append(o)
                jc = len(sh) - 1
                if p.vread and (self.vprobe is not None or self.vmap is not None) and \
                        sh[jc][0] == self.vcd and isinstance(rows[jc], int):
                    key = (p.name, kinds, tuple(rows[:jc]))
                    if self.vprobe is not None:
                        self.vprobe.setdefault(key, (p, set()))[1].add(rows[jc])
                    elif key in self.vmap:
                        cmin, prefix = self.vmap[key][:2]
                        i = rows[jc] - cmin
                        return '%s_%d.%s' % (prefix, i // 4, 'xyzw'[i % 4])
                return 'L_%s[%s + %s]' % (p.name, nm, self._off(parts, pr))

            nd = len(sh)

            def mk(nm):
                terms = []
                for j, (k, off) in enumerate(sh):
                    pbk = self.pb[k] if k != 'const' else '0'
                    terms.append('(%s)(%s - LB_%s_%d)%s' % (index_t, pbk, f.name, j,
                                                             ' * S_%s_%d' % (f.name, j) if j < nd - 1 else ''))
                return 'const %s* __restrict__ %s = %s + (%s);' % (
                    c_type(f.typ), nm, _gname(f), ' + '.join(terms) if terms else '0')
            nm = self._base(('g', f.name, kinds), mk)
            absd = []
            for j, (k, off) in enumerate(sh):
                o = self._dsum(k, off, pr) if k != 'const' else \
                    (off if isinstance(off, int) else pr.expr(off))
                absd.append(o)
                parts.append((o, 'S_%s_%d' % (f.name, j) if j < nd - 1 else 1))
            load = '%s[%s]' % (nm, self._off(parts, pr))
            if all(isinstance(o, int) for o in absd):
                key = (f.name, kinds, tuple(absd))
                if self.forward is not None and key in self.forward:
                    return self.forward[key]
                if self.record is not None:
                    self.record.setdefault(key, load)
            return load

        def ref_scaled(self, f, sc, pr):
            """An LDS read whose arguments are ``a * variable + c`` (a >= 1)
            or constants: the per-item base is at ``a`` times the item's base
            point, each point adds ``a * d + c`` per dimension."""
            p = plan.stage_of(f)
            ext = plan.lds_ext(p)
            W = [_prod(ext[j + 1:]) for j in range(len(ext))]
            kinds = tuple('const' if k == 'const' else (k if a == 1 else (k, a))
                          for (k, a, c) in sc)
            nm = self.lds_base(p, kinds)
            parts = []
            rows = []
            for j, (k, a, c) in enumerate(sc):
                if k == 'const':
                    o = c if isinstance(c, int) else pr.expr(c)
                else:
                    d = self.d[k]
                    o = a * d + c if isinstance(d, int) else '(%d * (%s) + %d)' % (a, d, c)
                parts.append((o, W[j]))
                rows.append(o)
            jc = len(sc) - 1
            if p.vread and (self.vprobe is not None or self.vmap is not None) and \
                    sc[jc][0] == self.vcd and isinstance(rows[jc], int):
                key = (p.name, kinds, tuple(rows[:jc]))
                if self.vprobe is not None:
                    self.vprobe.setdefault(key, (p, set()))[1].add(rows[jc])
                elif key in self.vmap:
                    cmin, prefix = self.vmap[key][:2]
                    i = rows[jc] - cmin
                    return '%s_%d.%s' % (prefix, i // 4, 'xyzw'[i % 4])
            return 'L_%s[%s + %s]' % (p.name, nm, self._off(parts, pr))

        def lds_base(self, p, kinds):
            """Name of the per-item LDS index of producer p's element at
            the item's base point: consumer dimension ``kinds[j]`` (or
            ``(k, a)``: a times dimension k) for p's dimension j, 'const'
            for a constant argument."""
            ext = plan.lds_ext(p)
            W = [_prod(ext[j + 1:]) for j in range(len(ext))]

            def mk(nm):
                terms = []
                for j, k in enumerate(kinds):
                    if isinstance(k, tuple):
                        pbk = '%d * %s' % (k[1], self.pb[k[0]])
                    else:
                        pbk = self.pb[k] if k != 'const' else '0'
                    terms.append('__mul24(%s - %s, %d)' % (pbk, R(p.region[j][0]), W[j]))
                return 'const int %s = %s;' % (nm, ' + '.join(terms) if terms else '0')
            return self._base(('l', p.name, kinds), mk)

        def vrow_index(self, p, kinds, rows, col):
            """LDS index of producer p's element at the given offsets from
            the item's base point (``rows`` for every dimension but the
            innermost, ``col`` for the innermost)."""
            ext = plan.lds_ext(p)
            W = [_prod(ext[j + 1:]) for j in range(len(ext))]
            nm = self.lds_base(p, kinds)
            parts = [(o, W[j]) for j, o in enumerate(rows)] + [(col, 1)]
            return '%s + %s' % (nm, self._off(parts, None))

    # a tile whose region exceeds the planned static box (possible only for
    # sizes far from the planning estimates) is skipped and reported
    over = []
    for s_ in stages:
        for i in range(s_.nd):
            over.append('(%s - %s + 1 > %d)' % (R(s_.region[i][1]), R(s_.region[i][0]), s_.static_ext[i]))
    if over:
        w('if (%s) { if (threadIdx.x == 0) atomicOr(pmx_err, 1); return; }' % ' || '.join(over))
    # LDS typed views
    for s in stages:
        if s.needs_lds:
            w('%s* __restrict__ L_%s = (%s*)(pmx_lds + %d);' % (c_type(s.func.typ), s.name,
                                                                 c_type(s.func.typ), s.lds_offset))

    def inline_closure(s):
        out = []
        seen = set()

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

    def part_cond(s, pi, part, pnames, skip_box):
        conds = []
        if not skip_box:
            for i, (blo, bhi) in enumerate(part.box):
                lo_id, hi_id = box_ids[(s, pi, i)]
                if lo_id is not None:
                    conds.append('%s >= %s' % (pnames[i], R(lo_id)))
                if hi_id is not None:
                    conds.append('%s <= %s' % (pnames[i], R(hi_id)))
        return conds

    def emit_stage_value(out, s, target, pnames, ind, skip_box_single, item):
        """Statements assigning the value of stage s at point pnames to
        ``target`` and setting ``pmx_w`` when a part applied."""
        ww = lambda t, i: out.append('  ' * i + t)
        vn = dict(zip(s.func.variables, pnames))
        cv = s.func.variables

        def cheap(r):
            # loads at constant offsets from a per-item base stay whole
            # when the branches of a select are merged
            return all(shift_access(a, cv) is not None or scaled_access(a, cv) is not None
                       for a in r.arguments)
        pr = Printer(vn, make_ref_cb(s, item), mode, lazy=True,
                     exact_fma=not plan.opts.get('no_exact_fma'),
                     unify=None if plan.opts.get('no_unify') else cheap)
        pr._tmp = tmp_counter
        single = (len(s.parts) == 1)
        for pi, part in enumerate(s.parts):
            bc = part_cond(s, pi, part, pnames, skip_box_single and single)
            ww('if (%s) {' % (' && '.join(bc) if bc else 'true'), ind)
            inner = ind + 1
            if part.pred:
                pro = []
                pc = ' && '.join(pr.cond(c, pro) for c in part.pred)
                for l in pro:
                    ww(l, inner)
                ww('if (%s) {' % pc, inner)
                inner += 1
            pro = []
            val = pr.expr(part.expr, pro)
            for l in pro:
                ww(l, inner)
            ww('%s = %s; pmx_w = true;' % (target, val), inner)
            if part.pred:
                ww('}', inner - 1)
            ww('}', ind)

    def emit_point(out, s, item, ind, vname, wname):
        """One point of stage s: ``vname``/``wname`` receive value and flag."""
        ww = lambda t, i: out.append('  ' * i + t)
        ww('{', ind)
        pnames = []
        for i in range(s.nd):
            d = item.d[i]
            ww('const int q%d = %s + %s;' % (i, item.pb[i], d), ind + 1)
            pnames.append('q%d' % i)
        ww('bool pmx_w = false;', ind + 1)
        for t in inline_closure(s):
            tn = ['q%d' % i for i in range(t.nd)]
            ww('%s v_%s = (%s)0;' % (c_type(t.func.typ), t.name, c_type(t.func.typ)), ind + 1)
            ww('{', ind + 1)
            ww('bool pmx_w = false;', ind + 2)
            emit_stage_value(out, t, 'v_' + t.name, tn, ind + 2, False, item)
            ww('(void)pmx_w;', ind + 2)
            ww('}', ind + 1)
        ww('%s v = (%s)0;' % (c_type(s.func.typ), c_type(s.func.typ)), ind + 1)
        emit_stage_value(out, s, 'v', pnames, ind + 1, True, item)
        ww('%s = v; %s = pmx_w;' % (vname, wname), ind + 1)
        ww('}', ind)

    def emit_store(out, s, item, ind, vname, wname):
        ww = lambda t, i: out.append('  ' * i + t)
        ww('if (%s) {' % wname, ind)
        if s.needs_lds:
            ext = plan.lds_ext(s)
            W = [_prod(ext[j + 1:]) for j in range(s.nd)]

            def mk(nm):
                terms = ['__mul24(%s - %s, %d)' % (item.pb[j], R(s.region[j][0]), W[j])
                         for j in range(s.nd)]
                return 'const int %s = %s;' % (nm, ' + '.join(terms) if terms else '0')
            nm = item._base(('ls', s.name), mk)
            off = item._off([(item.d[j], W[j]) for j in range(s.nd)], None)
            ww('L_%s[%s + %s] = %s;' % (s.name, nm, off, vname), ind + 1)
        if s.liveout:
            f = s.func

            def mkg(nm):
                terms = ['(%s)(%s - LB_%s_%d)%s' % (index_t, item.pb[j], f.name, j,
                                                     ' * S_%s_%d' % (f.name, j) if j < s.nd - 1 else '')
                         for j in range(s.nd)]
                return '%s* __restrict__ %s = %s + (%s);' % (
                    c_type(f.typ), nm, _gname(f), ' + '.join(terms) if terms else '0')
            nm = item._base(('gs', f.name), mkg)
            off = item._off([(item.d[j], 'S_%s_%d' % (f.name, j) if j < s.nd - 1 else 1)
                             for j in range(s.nd)], None)
            if f in final_outputs:
                # option: streaming stores for results nobody reads back
                # (measured slower on unsharp, so off by default)
                st = '__builtin_nontemporal_store(%s, &%s[%s]);' % (vname, nm, off)
            else:
                st = '%s[%s] = %s;' % (nm, off, vname)
            if s is not plan.anchor:
                oc = []
                for i in range(s.nd):
                    lo, hi = s.own[i]
                    q = '(%s + %s)' % (item.pb[i], item.d[i])
                    oc.append('%s >= %s && %s <= %s' % (q, R(lo), q, R(hi)))
                ww('if (%s) %s' % (' && '.join(oc), st), ind + 1)
            else:
                ww(st, ind + 1)
        ww('}', ind)

    tmp_counter = [0]
    # the pipeline's outputs are written once and not read back by the
    # pipeline's later kernels' tiles: non-temporal stores keep them from
    # displacing the halos and intermediates the caches serve (Harris
    # 17.7 -> 15.6 us; 0.3-0.9% on the others); gpu option nt_stores=0
    # turns them off
    final_outputs = set(pipe.outputs) if plan.opts.get('nt_stores', 1) else set()
    # union boxes of fused sweeps (equal on every sampled tile; the union
    # keeps the loop correct for any tile and is covered by the guard)
    union_lines = []
    sweep_box = {}
    over2 = []
    for si, (ph, members) in enumerate(plan.sweeps):
        s0 = members[0]
        if len(members) == 1:
            sweep_box[si] = [(R(lo), R(hi)) for (lo, hi) in s0.region]
            continue
        box = []
        for i in range(s0.nd):
            los = sorted(set(R(m.region[i][0]) for m in members))
            his = sorted(set(R(m.region[i][1]) for m in members))
            lo = los[0]
            for x in los[1:]:
                lo = 'pmx_imin(%s, %s)' % (lo, x)
            hi = his[0]
            for x in his[1:]:
                hi = 'pmx_imax(%s, %s)' % (hi, x)
            union_lines.append('const int sw%d_lo%d = %s; const int sw%d_hi%d = %s;' % (si, i, lo, si, i, hi))
            box.append(('sw%d_lo%d' % (si, i), 'sw%d_hi%d' % (si, i)))
            over2.append('(sw%d_hi%d - sw%d_lo%d + 1 > %d)' % (si, i, si, i, s0.static_ext[i]))
        sweep_box[si] = box
    # (emitted before the LDS views; see the guard above)
    for l in union_lines:
        w(l)
    if over2:
        w('if (%s) { if (threadIdx.x == 0) atomicOr(pmx_err, 1); return; }' % ' || '.join(over2))

    fwd = _forward_pair(plan, threads)
    fw_src = {}         # key -> load expression in the first sweep's fast path
    fw_need = set()     # keys the later sweep's fast path reads
    fw_vars = {}        # key -> register name
    phase_code = {}     # phase -> [(comment, nitems, lines)]
    lp = getattr(plan, 'lane_pair', None)
    for si, (ph, members) in enumerate(plan.sweeps):
        if lp is not None and si == lp[0]:
            continue            # computed inside the consumer's sweep
        if lp is not None and si == lp[1]:
            phase_code.setdefault(ph, []).append(
                _lane_sweep(plan, si, sweep_box[si], Item, emit_point, emit_store, R, threads))
            continue
        SL = []
        w3 = lambda t, ind=3: SL.append('  ' * ind + t)
        s0 = members[0]
        box = sweep_box[si]
        ext = s0.static_ext
        if fwd is not None and si == fwd[1]:
            # the later sweep walks the first one's item space (its own
            # region checks added below), so each thread's block sits on
            # the points it loaded there
            box = sweep_box[fwd[0]]
            ext = plan.sweeps[fwd[0]][1][0].static_ext
        U, V = s0.block
        nd = s0.nd
        rd, cd = s0.bdims
        step = [1] * nd
        cnt = list(ext)
        if cd is not None:
            step[cd] = V
            cnt[cd] = -(-ext[cd] // V)
        if rd is not None:
            step[rd] = U
            cnt[rd] = -(-ext[rd] // U)
        nitems = _prod(cnt)
        comment = 'stage %s (phase %d, %dx%d points per thread)' % (
            ' + '.join(m.name for m in members), ph, U, V)
        rem = 'it'
        pb = []
        for i in range(nd):
            st = _prod(cnt[i + 1:])
            if cnt[i] == 1:
                # a dimension of one point (a split level or channel): the
                # index is 0, uniform, so what depends on it is scalar work
                w3('const int j%d = 0;' % i)
            elif i < nd - 1:
                mg = _magic24(cnt[i] * st, st)
                if mg:
                    # it / st as a 24-bit multiply and a shift, exact for
                    # every dividend this sweep produces
                    w3('const int j%d = (int)(__umul24((unsigned)%s, %du) >> %d); '
                       'const int k%d = %s - __mul24(j%d, %d);' % (i, rem, mg[0], mg[1], i, rem, i, st))
                else:
                    w3('const int j%d = %s / %d; const int k%d = %s - j%d * %d;' % (
                        i, rem, st, i, rem, i, st))
                rem = 'k%d' % i
            else:
                w3('const int j%d = %s;' % (i, rem))
            w3('const int b%d = %s + j%d%s;' % (i, box[i][0], i,
                                                 (' * %d' % step[i]) if step[i] > 1 else ''))
            pb.append('b%d' % i)
        guard = ' || '.join('b%d > %s' % (i, box[i][1]) for i in range(nd))
        if guard:
            w3('if (%s) continue;' % guard)
        blocked = U * V > 1
        pts = [(u, v) for u in range(U) for v in range(V)]
        body_ind = 4 if blocked or len(members) > 1 or (fwd is not None and si == fwd[1]) else 3

        def dvec(u, v):
            d = [0] * nd
            if rd is not None:
                d[rd] = u
            if cd is not None:
                d[cd] = v
            return d

        def inside(m, lo_terms, hi_terms):
            """C condition: coordinates lo_terms[i]..hi_terms[i] (hi None:
            just lo_terms[i]) lie in m's region; checks the loop guard
            already implies (b_i <= the sweep box's upper bound, which is
            m's own) are left out."""
            c = []
            for i in range(nd):
                mlo, mhi = R(m.region[i][0]), R(m.region[i][1])
                if mlo != box[i][0]:
                    c.append('%s >= %s' % (lo_terms[i], mlo))
                top = hi_terms[i] if hi_terms[i] is not None else lo_terms[i]
                if not (mhi == box[i][1] and top == 'b%d' % i):
                    c.append('%s <= %s' % (top, mhi))
            return c

        fitem = Item(pb, dvec(0, 0))
        if fwd is not None and si == fwd[0]:
            fitem.record = fw_src
        if fwd is not None and si == fwd[1]:
            # which of the first sweep's loads this fast path reads (a dry
            # run of its points), then the same points reading registers
            probe = Item(pb, dvec(0, 0))
            probe.record = {}
            for m in members:
                for (u, v) in pts:
                    probe.d = dvec(u, v)
                    emit_point([], m, probe, body_ind, 'x', 'y')
            for key in probe.record:
                if key in fw_src:
                    fw_need.add(key)
            fitem.forward = dict((k, fw_vars[k]) for k in fw_need)
        vconds = []
        vloads = []
        if V == 4 and cd is not None and any(
                plan.in_group(r.objectRef) and plan.stage_of(r.objectRef).vread
                for m in members for (_, r) in plan.closure_refs(m)):
            # the columns each producer row is read at over the block's
            # points (a dry run), then one aligned 128-bit LDS load per four
            # columns of each row; a block whose rows are not 16-byte
            # aligned takes the per-point path
            vprobe = Item(pb, dvec(0, 0))
            vprobe.vcd = cd
            vprobe.vprobe = {}
            for m in members:
                for (u, v) in pts:
                    vprobe.d = dvec(u, v)
                    emit_point([], m, vprobe, body_ind, 'x', 'y')
            vmap = {}
            for n, key in enumerate(sorted(vprobe.vprobe, key=str)):
                p, cols = vprobe.vprobe[key]
                cmin = min(cols)
                nch = (max(cols) - cmin) // 4 + 1
                vmap[key] = (cmin, 'lq%d' % n, nch, p)
            fitem.vcd = cd
            fitem.vmap = vmap
            vtype = {'float': 'float4', 'int': 'int4', 'unsigned int': 'uint4'}
            for key in sorted(vmap, key=lambda k: int(vmap[k][1][2:])):
                cmin, prefix, nch, p = vmap[key]
                idx = fitem.vrow_index(p, key[1], list(key[2]), cmin)
                vt = vtype[c_type(p.func.typ)]
                vconds.append('((%s) & 3) == 0' % idx)
                for c in range(nch):
                    vloads.append('  ' * body_ind + 'const %s %s_%d = pmx_ld4<%s>(&L_%s[%s + %d]);' % (
                        vt, prefix, c, vt, p.name, idx, 4 * c))
        fast = list(vloads)
        for mi, m in enumerate(members):
            T = c_type(m.func.typ)
            for (u, v) in pts:
                fast.append('  ' * body_ind + '%s pv%d_%d_%d = (%s)0; bool pw%d_%d_%d = false;' % (
                    T, mi, u, v, T, mi, u, v))
        for mi, m in enumerate(members):
            for (u, v) in pts:
                fitem.d = dvec(u, v)
                emit_point(fast, m, fitem, body_ind, 'pv%d_%d_%d' % (mi, u, v), 'pw%d_%d_%d' % (mi, u, v))
        for mi, m in enumerate(members):
            for (u, v) in pts:
                fitem.d = dvec(u, v)
                emit_store(fast, m, fitem, body_ind, 'pv%d_%d_%d' % (mi, u, v), 'pw%d_%d_%d' % (mi, u, v))
        if fwd is not None and si == fwd[0]:
            # every load of this fast path a later sweep may take over, kept
            # in a register (the point code above issued the same loads)
            for key in sorted(fw_src):
                fw_vars[key] = 'fw%d' % len(fw_vars)
                fast.append('  ' * body_ind + '%s = %s;' % (fw_vars[key], fw_src[key]))
            # the later sweep takes the registers only where this fast path
            # ran for its item (its block inside this sweep's region)
            fast.append('  ' * body_ind + 'fw_ok = true;')
        slow = []
        sitem = Item(pb, dvec('pu', 'pv'))
        sitem.names = fitem.names
        sitem.decls = fitem.decls
        # (a sweep walking another's item space checks its own region)
        need_slow = blocked or len(members) > 1 or (fwd is not None and si == fwd[1])
        if need_slow:
            slow.append('  ' * 4 + 'for (int pu = 0; pu < %d; ++pu) for (int pv = 0; pv < %d; ++pv) {' % (U, V))
            pt = ['b%d%s' % (i, ' + pu' if i == rd else (' + pv' if i == cd else '')) for i in range(nd)]
            for mi, m in enumerate(members):
                conds = inside(m, pt, [None] * nd)
                T = c_type(m.func.typ)
                slow.append('  ' * 5 + 'if (%s) {' % (' && '.join(conds) if conds else 'true'))
                slow.append('  ' * 6 + '%s pv_s = (%s)0; bool pw_s = false;' % (T, T))
                emit_point(slow, m, sitem, 6, 'pv_s', 'pw_s')
                emit_store(slow, m, sitem, 6, 'pv_s', 'pw_s')
                slow.append('  ' * 5 + '}')
            slow.append('  ' * 4 + '}')
        for dcl in fitem.decls:
            w3(dcl)
        if need_slow:
            fc = []
            lo_t = ['b%d' % i for i in range(nd)]
            hi_t = ['b%d + %d' % (i, (U - 1) if i == rd else ((V - 1) if i == cd else 0))
                    if i in (rd, cd) else None for i in range(nd)]
            for m in members:
                for c in inside(m, lo_t, hi_t):
                    if c not in fc:
                        fc.append(c)
            if fwd is not None and si == fwd[1] and fw_need:
                fc.append('fw_ok')
            fc.extend(vconds)
            w3('if (%s) {' % (' && '.join(fc) if fc else 'true'))
            SL.extend(fast)
            w3('} else {')
            SL.extend(slow)
            w3('}')
        else:
            SL.extend(fast)
        phase_code.setdefault(ph, []).append((comment, nitems, SL))

    if fwd is not None:
        # registers carried from the first sweep to the later one (each
        # thread runs one item of each, on the same block)
        ftyp = dict((f.name, c_type(f.typ)) for f in buffers)
        keep = set(fw_vars[k] for k in fw_need)
        w('bool fw_ok = false; (void)fw_ok;')
        for key in sorted(fw_vars, key=lambda k: fw_vars[k]):
            if fw_vars[key] in keep:
                w('%s %s = (%s)0;' % (ftyp[key[0]], fw_vars[key], ftyp[key[0]]))
        for ph, sws in phase_code.items():
            for (_, _, SL) in sws:
                SL[:] = [l for l in SL if not (_FW_LINE.match(l.strip()) and
                                               l.strip().split(' ', 1)[0] not in keep)]
    # one loop per phase over the concatenated item spaces of its sweeps, so
    # several small sweeps share the workgroup's lanes instead of each
    # leaving most of them idle in its last iteration
    first = True
    for ph in range(plan.nphases):
        sws = phase_code.get(ph, [])
        if not sws:
            continue
        if not first:
            w('__syncthreads();')
        first = False
        total_items = sum(n for (_, n, _) in sws)
        w('{ // phase %d: %s' % (ph, '; '.join(c for (c, _, _) in sws)))
        w('for (int pmx_it = pmx_tx; pmx_it < %d; pmx_it += %d) {' % (total_items, threads), 2)
        base = 0
        for k, (c, n, SL) in enumerate(sws):
            if len(sws) == 1:
                w('{ const int it = pmx_it;', 2)
            else:
                w('%sif (pmx_it < %d) { const int it = pmx_it - %d;' % (
                    '' if k == 0 else '} else ', base + n, base), 2)
            L.extend(SL)
            base += n
        w('}', 2)
        w('}', 2)
        w('}')
    roll = max(1, int(plan.opts.get('roll', 1))) if roll_d is not None else 1
    tiles_c = ' * '.join('((%s + %d) / %d)' % (_extent_c(a.func, d, base_pr), T - 1, T)
                         for (d, T) in grid) or '1'
    grid_c = ' * '.join(('((((%s + %d) / %d) + %d) / %d)' % (_extent_c(a.func, d, base_pr), T - 1, T,
                                                              roll - 1, roll))
                        if d == roll_d else
                        '((%s + %d) / %d)' % (_extent_c(a.func, d, base_pr), T - 1, T)
                        for (d, T) in grid) or '1'
    info = KernelInfo(kname, 'group', buffers, grid_c, threads, plan.lds_bytes, plan, reads, writes)
    info.tiles_c = tiles_c
    return wrap_kernel(kname, params, buffers, writes, threads, plan.lds_bytes, L, roll=roll), info




def _lane_sweep(plan, si, box, Item, emit_point, emit_store, R, threads):
    """Code of a consumer sweep with its producer sweep lane-fused
    (group_plan._lane_fuse): (comment, items, lines).

    Items run over the consumer's region: outer dimensions, blocks of U
    rows, then columns in wavefront-wide groups of 64 lanes whose
    consecutive columns overlap by 2*reach (a wavefront advances by
    64 - 2*reach columns), so every item of a wavefront shares its outer
    index and row block.  Each item computes the producer stages for its
    column on rows -below .. U-1+above into registers, exchanges them with
    the lanes ``reach`` to each side (all lanes take part: no branch
    encloses the exchange), and lanes reach .. 63-reach compute the
    consumer's U points."""
    lp = plan.lane_pair
    ph, members = plan.sweeps[si]
    c = members[0]
    prod = plan.sweeps[lp[0]][1]
    below, above, B = lp[2], lp[3], lp[4]
    U = c.block[0]
    rd, cd = c.bdims
    nd = c.nd
    ext = c.static_ext
    Wv = 64 - 2 * B
    cnt = list(ext)
    cnt[rd] = -(-ext[rd] // U)
    cnt[cd] = -(-ext[cd] // Wv) * 64
    nitems = _prod(cnt)
    SL = []
    w3 = lambda t, ind=3: SL.append('  ' * ind + t)
    rem = 'it'
    pb = []
    for i in range(nd):
        st = _prod(cnt[i + 1:])
        if cnt[i] == 1:
            w3('const int j%d = 0;' % i)
        elif i < nd - 1:
            mg = _magic24(cnt[i] * st, st)
            if mg:
                w3('const int j%d = (int)(__umul24((unsigned)%s, %du) >> %d); '
                   'const int k%d = %s - __mul24(j%d, %d);' % (i, rem, mg[0], mg[1], i, rem, i, st))
            else:
                w3('const int j%d = %s / %d; const int k%d = %s - j%d * %d;' % (i, rem, st, i, rem, i, st))
            rem = 'k%d' % i
        else:
            w3('const int j%d = %s;' % (i, rem))
        if i == cd:
            w3('const int pmx_lane = j%d & 63;' % i)
            w3('const int b%d = %s - %d + (j%d >> 6) * %d + pmx_lane;' % (i, box[i][0], B, i, Wv))
        elif i == rd:
            w3('const int b%d = %s + j%d * %d;' % (i, box[i][0], i, U))
        else:
            w3('const int b%d = %s + j%d;' % (i, box[i][0], i))
        pb.append('b%d' % i)
    # whole wavefronts share the outer index and the row block
    og = ' || '.join('b%d > %s' % (i, box[i][1]) for i in range(nd) if i != cd)
    if og:
        w3('if (%s) continue;' % og)

    def dvec(u):
        d = [0] * nd
        d[rd] = u
        return d
    rows = list(range(-below, U + above))
    item = Item(pb, dvec(0))
    # producer values on the rows of this block, this column
    for m in prod:
        T = c_type(m.func.typ)
        for u in rows:
            w3('%s lv_%s_%d = (%s)0;' % (T, m.name, u + below, T))
    fast_c = []
    for m in prod:
        lo_r, hi_r = R(m.region[rd][0]), R(m.region[rd][1])
        lo_c, hi_c = R(m.region[cd][0]), R(m.region[cd][1])
        fast_c += ['b%d - %d >= %s' % (rd, below, lo_r), 'b%d + %d <= %s' % (rd, U - 1 + above, hi_r),
                   'b%d >= %s' % (cd, lo_c), 'b%d <= %s' % (cd, hi_c)]
        for i in range(nd):
            if i not in (rd, cd):
                fast_c += ['b%d >= %s' % (i, R(m.region[i][0])), 'b%d <= %s' % (i, R(m.region[i][1]))]
    body = []
    for m in prod:
        for u in rows:
            item.d = dvec(u)
            emit_point(body, m, item, 4, 'lv_%s_%d' % (m.name, u + below), 'pmx_lw')
    slow = []
    for m in prod:
        for u in rows:
            conds = ['b%d + %d >= %s' % (rd, u, R(m.region[rd][0])),
                     'b%d + %d <= %s' % (rd, u, R(m.region[rd][1])),
                     'b%d >= %s' % (cd, R(m.region[cd][0])), 'b%d <= %s' % (cd, R(m.region[cd][1]))]
            for i in range(nd):
                if i not in (rd, cd):
                    conds += ['b%d >= %s' % (i, R(m.region[i][0])), 'b%d <= %s' % (i, R(m.region[i][1]))]
            slow.append('  ' * 4 + 'if (%s) {' % ' && '.join(conds))
            item.d = dvec(u)
            emit_point(slow, m, item, 5, 'lv_%s_%d' % (m.name, u + below), 'pmx_lw')
            slow.append('  ' * 4 + '}')
    for dcl in item.decls:
        w3(dcl)
    w3('{ bool pmx_lw = false; (void)pmx_lw;')
    w3('if (%s) {' % ' && '.join(fast_c))
    SL.extend(body)
    w3('} else {')
    SL.extend(slow)
    w3('}')
    w3('}')
    # the consumer's points: which neighbouring values they read
    probe = Item(pb, dvec(0))
    probe.lane_keys = set()
    for u in range(U):
        probe.d = dvec(u)
        emit_point([], c, probe, 4, 'x', 'y')
    typ = dict((m.name, c_type(m.func.typ)) for m in prod)
    for (mn, row, dc) in sorted(probe.lane_keys):
        if dc:
            w3('const %s lx_%s_%d_%s%d = pmx_shfl(lv_%s_%d, pmx_lane + (%d));' % (
                typ[mn], mn, row, 'm' if dc < 0 else 'p', abs(dc), mn, row, dc))
    citem = Item(pb, dvec(0))
    citem.names = item.names        # one namespace of base addresses per item
    pts = []
    for u in range(U):
        T = c_type(c.func.typ)
        pts.append('  ' * 4 + '%s pv_%d = (%s)0; bool pw_%d = false;' % (T, u, T, u))
    fast = list(pts)
    for u in range(U):
        citem.d = dvec(u)
        emit_point(fast, c, citem, 4, 'pv_%d' % u, 'pw_%d' % u)
    for u in range(U):
        citem.d = dvec(u)
        emit_store(fast, c, citem, 4, 'pv_%d' % u, 'pw_%d' % u)
    cslow = []
    for u in range(U):
        conds = ['b%d + %d <= %s' % (rd, u, box[rd][1]), 'b%d >= %s' % (cd, box[cd][0]),
                 'b%d <= %s' % (cd, box[cd][1])]
        cslow.append('  ' * 4 + 'if (%s) {' % ' && '.join(conds))
        T = c_type(c.func.typ)
        cslow.append('  ' * 5 + '%s pv_s = (%s)0; bool pw_s = false;' % (T, T))
        citem.d = dvec(u)
        emit_point(cslow, c, citem, 5, 'pv_s', 'pw_s')
        emit_store(cslow, c, citem, 5, 'pv_s', 'pw_s')
        cslow.append('  ' * 4 + '}')
    for dcl in citem.decls:
        w3(dcl)
    w3('if (pmx_lane >= %d && pmx_lane < %d) {' % (B, 64 - B))
    w3('if (b%d + %d <= %s && b%d >= %s && b%d <= %s) {' % (rd, U - 1, box[rd][1], cd, box[cd][0],
                                                          cd, box[cd][1]), 4)
    SL.extend(fast)
    w3('} else {', 4)
    SL.extend(cslow)
    w3('}', 4)
    w3('}')
    comment = 'stage %s, with %s in registers across lanes (phase %d, %dx1 points per thread)' % (
        c.name, ' + '.join(m.name for m in prod), ph, U)
    return (comment, nitems, SL)


_FW_LINE = re.compile(r'^fw\d+ = ')


def _forward_pair(plan, threads):
    """(first, later) sweep indices whose global loads can be forwarded in
    registers, or None.  Both are the only sweep of their phase, have the
    same block shape and dimensions, need one item per thread (so each
    thread's registers from the first survive the barrier into the
    later), and the later one's region lies inside the first's on every
    sampled tile (it then walks the first's item space).  Unsharp: the
    mask reads the image at the points the blur loaded; without this the
    second read goes back to L2."""
    import numpy as np
    if plan.opts.get('no_forward') or getattr(plan, 'lane_pair', None) is not None:
        return None
    by_phase = {}
    for si, (ph, members) in enumerate(plan.sweeps):
        by_phase.setdefault(ph, []).append(si)

    def items(s):
        U, V = s.block
        rd, cd = s.bdims
        cnt = list(s.static_ext)
        if cd is not None:
            cnt[cd] = -(-cnt[cd] // V)
        if rd is not None:
            cnt[rd] = -(-cnt[rd] // U)
        return _prod(cnt)

    phs = sorted(by_phase)
    for a_i, a in enumerate(phs):
        if len(by_phase[a]) != 1:
            continue
        s1i = by_phase[a][0]
        m1 = plan.sweeps[s1i][1]
        if len(m1) != 1 or items(m1[0]) > threads:
            continue
        s1 = m1[0]
        for b in phs[a_i + 1:]:
            if len(by_phase[b]) != 1:
                continue
            s2i = by_phase[b][0]
            m2 = plan.sweeps[s2i][1]
            if len(m2) != 1:
                continue
            s2 = m2[0]
            if (s2.nd, s2.block, s2.bdims) != (s1.nd, s1.block, s1.bdims):
                continue
            if s2.block[1] == 4 and any(plan.in_group(r.objectRef) and plan.stage_of(r.objectRef).vread
                                        for (t, r) in plan.closure_refs(s2)):
                # its 128-bit LDS row loads align on its own item grid, not
                # on the first sweep's (which starts at the producer's region)
                continue
            g1 = set(r.objectRef for (t, r) in plan.closure_refs(s1) if not plan.in_group(r.objectRef))
            g2 = set(r.objectRef for (t, r) in plan.closure_refs(s2) if not plan.in_group(r.objectRef))
            if not (g1 & g2):
                continue
            inside = all(np.all(l2 >= l1) and np.all(h2 <= h1)
                         for (l1, h1), (l2, h2) in zip(plan._region_samples[s1], plan._region_samples[s2]))
            if inside and all(e2 <= e1 for e1, e2 in zip(s1.static_ext, s2.static_ext)):
                return (s1i, s2i)
    return None


def _magic24(n, d):
    """(m, s) with (x * m) >> s == x // d for every 0 <= x < n, where x * m
    stays below 2**32 and m below 2**24 (so a 24-bit multiply computes it
    exactly), or None."""
    if d <= 1 or n <= 0 or n >= (1 << 24):
        return None
    import numpy as np
    x = np.arange(n, dtype=np.uint64)
    for sh in range(1, 33):
        m = -(-(1 << sh) // d)
        if m >= (1 << 24) or (n - 1) * m >= (1 << 32):
            break
        if np.array_equal((x * np.uint64(m)) >> np.uint64(sh), x // np.uint64(d)):
            return m, sh
    return None


def kernel_args(params, buffers, writes):
    """(declarations, names, frame-struct fields) of a kernel's arguments."""
    decls, names, fields = ['int* __restrict__ pmx_err'], ['pmx_err'], ['int* pmx_err;']
    for f in buffers:
        q = '' if f in writes else 'const '
        decls.append('%s%s* __restrict__ G_%s' % (q, c_type(f.typ), f.name))
        names.append('G_' + f.name)
        fields.append('%s%s* G_%s;' % (q, c_type(f.typ), f.name))
    fields.append('long long pmx_tile0;')
    for p in params:
        decls.append('const %s %s' % (c_type(p.typ), p.name))
        names.append(p.name)
        fields.append('%s %s;' % (c_type(p.typ), p.name))
    return decls, names, fields


def wrap_kernel(kname, params, buffers, writes, threads, lds_bytes, body_lines, roll=1):
    """The kernel body as a device function plus two entry points:

    ``<k>``       one workgroup per tile (grid = number of tiles);
    ``<k>_batch`` persistent: a fixed set of workgroups pulls (frame, tile)
                  work items from 8 atomic queues (one per XCD) until all
                  are empty; frames are described by ``<k>_frame`` records
                  (buffers, first global work item, parameters).
    Every workgroup of the persistent form leaves the loop once all 8
    counters have passed their queue's end, so the grid always drains."""
    decls, names, fields = kernel_args(params, buffers, writes)
    lds = max(16, lds_bytes)
    tail = ['const int pmx_tile', 'const int pmx_remap', 'const int pmx_rk', 'const int pmx_rr',
            'unsigned char* __restrict__ pmx_lds']
    prologue = '  (void)pmx_remap; (void)pmx_lds; (void)pmx_err; (void)pmx_rk; (void)pmx_rr;'
    # plain form: kernel arguments, whose address space and aliasing the
    # compiler knows, and the thread index with its known range
    L = ['__device__ __forceinline__ void %s_body(%s) {' % (kname, ', '.join(decls + tail)),
         prologue, '  const int pmx_tx = (int)threadIdx.x; (void)pmx_tx;']
    L += body_lines
    L.append('}')
    # persistent form: the same body over pointers read from frame records.
    # Buffers the kernel writes are global-address-space pointers (global,
    # not flat, stores); buffers it only reads are constant-address-space
    # pointers (they do not change during the launch, so reads at uniform
    # addresses become scalar loads).  The thread index is opaque to the
    # optimiser, so per-thread index arithmetic is recomputed for every work
    # item instead of being hoisted out of the work loop into registers.
    ptypes = {'pmx_err': 'int'}
    for f in buffers:
        ptypes['G_' + f.name] = ('' if f in writes else 'const ') + c_type(f.typ)
    aspace = dict((n, 'PMX_CONST' if t.startswith('const ') else 'PMX_GLOBAL')
                  for n, t in ptypes.items())
    bdecls = ['%s %s* __restrict__ %s' % (ptypes[n], aspace[n], n) if n in ptypes else d
              for n, d in zip(names, decls)]
    fargs = ['(%s %s*)f.%s' % (ptypes[n], aspace[n], n) if n in ptypes else 'f.' + n
             for n in names]
    L += ['__device__ __forceinline__ void %s_bbody(%s) {' % (kname, ', '.join(bdecls + tail)),
          prologue, '  int pmx_tx = (int)threadIdx.x; PMX_OPAQUE(pmx_tx); (void)pmx_tx;']
    L += body_lines
    L.append('}')
    L.append('extern "C" __global__ __launch_bounds__(%d) void %s(%s) {' % (
        threads, kname, ', '.join(decls)))
    L.append('  __shared__ __align__(16) unsigned char pmx_lds[%d];' % lds)
    if roll > 1:
        L.append('  for (int pmx_k = 0; pmx_k < %d; ++pmx_k) {' % roll)
        L.append('    if (pmx_k) __syncthreads();')
        L.append('    %s_body(%s);' % (kname, ', '.join(names + ['(int)blockIdx.x', '1', 'pmx_k',
                                                             str(roll), 'pmx_lds'])))
        L.append('  }')
    else:
        L.append('  %s_body(%s);' % (kname, ', '.join(names + ['(int)blockIdx.x', '1', '0', '1',
                                                             'pmx_lds'])))
    L.append('}')
    L.append('struct %s_frame { %s };' % (kname, ' '.join(fields)))
    L.append('extern "C" __global__ __launch_bounds__(%d) void %s_batch(const %s_frame* __restrict__ F, '
             'const int nframes, unsigned long long* __restrict__ counter, const long long total) {'
             % (threads, kname, kname))
    L.append('  __shared__ __align__(16) unsigned char pmx_lds[%d];' % lds)
    L.append('  __shared__ long long pmx_work;')
    L.append('  // frame records do not change during the launch: read through the')
    L.append('  // constant address space they are scalar loads')
    L.append('  const %s_frame PMX_CONST* __restrict__ Fc = (const %s_frame PMX_CONST*)F;' % (kname, kname))
    L.append('  // 8 queues, one per XCD (workgroups are dealt round-robin to the')
    L.append('  // XCDs): queue q holds the contiguous items [total*q/8, total*(q+1)/8),')
    L.append('  // so neighbouring tiles share an L2; an empty home queue sends the')
    L.append('  // workgroup on to the next ones until all 8 are empty.  The next')
    L.append('  // item is fetched while the current one is processed.')
    L.append('  const int home = (int)(blockIdx.x & 7);')
    L.append('  int q = 0;')
    L.append('  auto pmx_fetch = [&]() -> long long {')
    L.append('    while (q < 8) {')
    L.append('      const int qq = (home + q) & 7;')
    L.append('      const long long lo = total * qq / 8, hi = total * (qq + 1) / 8;')
    L.append('      const long long k = (long long)atomicAdd(&counter[qq], 1ULL);')
    L.append('      if (lo + k < hi) return lo + k;')
    L.append('      ++q;')
    L.append('    }')
    L.append('    return -1;')
    L.append('  };')
    L.append('  if (threadIdx.x == 0) pmx_work = pmx_fetch();')
    L.append('  __syncthreads();')
    L.append('  for (;;) {')
    L.append('    // the work item and its frame are the same for the whole workgroup:')
    L.append('    // readfirstlane makes that visible, so the frame record (copied')
    L.append('    // word by word) and all per-tile index arithmetic stay in scalar')
    L.append('    // registers')
    L.append('    const long long tw = pmx_work;')
    L.append('    const long long t = (long long)(((unsigned long long)(unsigned)__builtin_amdgcn_readfirstlane('
             '(int)(tw >> 32)) << 32) | (unsigned long long)(unsigned)__builtin_amdgcn_readfirstlane((int)tw));')
    L.append('    if (t < 0) break;')
    L.append('    long long nxt = -1;')
    L.append('    if (threadIdx.x == 0) nxt = pmx_fetch();')
    L.append('    // the frame holding work item t: last frame whose first item is <= t')
    L.append('    int lo = 0, hi = nframes - 1;')
    L.append('    while (lo < hi) { const int mid = (lo + hi + 1) >> 1; if (Fc[mid].pmx_tile0 <= t) lo = mid; else hi = mid - 1; }')
    L.append('    lo = __builtin_amdgcn_readfirstlane(lo);')
    L.append('    %s_frame f;' % kname)
    L.append('    { int w[sizeof(f) / 4]; __builtin_memcpy(w, (const void*)&Fc[lo], sizeof(f));')
    L.append('      for (int i = 0; i < (int)(sizeof(f) / 4); ++i) w[i] = __builtin_amdgcn_readfirstlane(w[i]);')
    L.append('      __builtin_memcpy(&f, w, sizeof(f)); }')
    L.append('    %s_bbody(%s);' % (kname, ', '.join(fargs +
                                                   ['(int)(t - f.pmx_tile0)', '0', '0', '1',
                                                    'pmx_lds'])))
    L.append('    __syncthreads();')
    L.append('    if (threadIdx.x == 0) pmx_work = nxt;')
    L.append('    __syncthreads();')
    L.append('  }')
    L.append('}')
    return '\n'.join(L) + '\n'


# ---------------------------------------------------------------------------
# pipeline
# ---------------------------------------------------------------------------

def plan_pipeline(pipe):
    plans = []
    for gi, g in enumerate(pipe.groups):
        comps = g.comps
        if len(comps) == 1 and comps[0].is_reduction:
            plans.append(ReductionPlan(pipe, g, gi))
        elif len(comps) == 1 and comps[0].is_const:
            plans.append(ConstPlan(pipe, g, gi))
        else:
            plans.append(GroupPlan(pipe, g, gi))
    return plans


def emit_pipeline(pipe, is_extern_c_func=True, are_io_void_ptrs=True):
    mode = pipe.literal_mode
    opts = pipe.gpu_options
    threads = int(opts.get('threads', 256))
    xcd = bool(opts.get('xcd_swizzle', True))
    plans = plan_pipeline(pipe)
    srcs = ['// generated by polymage_amd for pipeline %s' % pipe.name,
            DEVICE_PRELUDE, REGION_PRELUDE,
            'extern "C" __global__ void pmx_zero_u64(unsigned long long* p, const int n) {\n'
            '  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);\n'
            '  if (i < n) p[i] = 0ULL;\n}\n']
    kernels = []
    host_consts = []
    for i, plan in enumerate(plans):
        kname = 'pmx_%s_g%d' % (pipe.name, i)
        if isinstance(plan, GroupPlan):
            src, info = emit_group_kernel(pipe, plan, kname, threads, mode, xcd)
        elif isinstance(plan, ReductionPlan):
            for src, info in emit_reduction_kernel(pipe, plan, kname, mode):
                srcs.append(src)
                kernels.append(info)
            continue
        else:
            if plan.on_host:
                host_consts.append(plan)
                continue
            src, info = emit_const_kernel(pipe, plan, kname, mode)
        srcs.append(src)
        kernels.append(info)
    hip_src = '\n'.join(srcs)
    from .host_stub import emit_host_stub
    host_src = emit_host_stub(pipe, kernels, host_consts)
    return GeneratedCode(pipe.name, hip_src, host_src, kernels,
                         {'plans': plans, 'threads': threads})
