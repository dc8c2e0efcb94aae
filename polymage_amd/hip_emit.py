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

from .analysis import PlanError, classify_access, eval_bexpr, unconditional_refs
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
        return ('// ===== HIP kernels (hiprtc, gfx950) =====\n' + self.hip_src +
                '\n// ===== host stub (g++) =====\n' + self.host_src)

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
    # tile counts (runtime)
    for (d, T) in grid:
        w('const int n%d = ((%s) + %d) / %d;' % (d, _extent_c(a.func, d, base_pr), T - 1, T))
    ntot = ' * '.join('n%d' % d for (d, T) in grid) or '1'
    w('const int pmx_ntiles = %s;' % ntot)
    if xcd:
        # blocks are dealt round-robin to the 8 XCDs; give each XCD a
        # contiguous run of tiles so neighbouring halos hit the same L2
        w('if (pmx_remap) { const int per = pmx_ntiles / 8, rem = pmx_ntiles % 8;')
        w('  const int xcd = pmx_b % 8, k = pmx_b / 8;')
        w('  pmx_b = (xcd < rem ? xcd * (per + 1) : rem * (per + 1) + (xcd - rem) * per) + k; }')
    for (d, T) in reversed(grid):
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
            e = term if e is None else '(%s * %d + %s)' % (e, p.static_ext[d], term)
        return e if e is not None else '0'

    # references that only run inside a Select branch: their global reads
    # are clamped to the producer's domain (the run-time bounds check in
    # the host stub covers unconditional references only)
    uncond = set()
    for s in stages:
        for part in s.parts:
            for r in unconditional_refs(part.expr):
                uncond.add(id(r))
            for c in part.pred:
                for r in c.collect(Reference):
                    uncond.add(id(r))

    def make_ref_cb(cur_stage):
        def cb(ref, pr, pro):
            f = ref.objectRef
            cvars = cur_stage.func.variables
            idx = []
            for d, arg in enumerate(ref.arguments):
                acc = classify_access(arg, cvars)
                q = pr.expr(arg, pro)
                if acc.kind == '*' or (id(ref) not in uncond and not plan.in_group(f)):
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

    def emit_stage_value(s, target, pnames, ind, skip_box_single):
        """Statements assigning the value of stage s at point pnames to
        ``target``; returns name of a bool flag set when a part applied."""
        vn = dict(zip(s.func.variables, pnames))
        pr = Printer(vn, make_ref_cb(s), mode, lazy=True)
        pr._tmp = tmp_counter
        single = (len(s.parts) == 1)
        for pi, part in enumerate(s.parts):
            bc = part_cond(s, pi, part, pnames, skip_box_single and single)
            w('if (%s) {' % (' && '.join(bc) if bc else 'true'), ind)
            inner = ind + 1
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
            w('}', ind)

    tmp_counter = [0]
    for ph in range(plan.nphases):
        for s in stages:
            if s.inline or s.phase != ph:
                continue
            ext = s.static_ext
            N = _prod(ext)
            w('{ // stage %s (phase %d)' % (s.name, ph))
            w('for (int idx = threadIdx.x; idx < %d; idx += %d) {' % (N, threads), 2)
            rem = 'idx'
            pnames = []
            for i in range(s.nd):
                st = _prod(ext[i + 1:])
                if i < s.nd - 1:
                    w('const int i%d = %s / %d; const int m%d = %s - i%d * %d;' % (
                        i, rem, st, i, rem, i, st), 3)
                    rem = 'm%d' % i
                else:
                    w('const int i%d = %s;' % (i, rem), 3)
                w('const int p%d = %s + i%d;' % (i, R(s.region[i][0]), i), 3)
                pnames.append('p%d' % i)
            guard = ' || '.join('p%d > %s' % (i, R(s.region[i][1])) for i in range(s.nd))
            if guard:
                w('if (%s) continue;' % guard, 3)
            w('bool pmx_w = false;', 3)
            for t in inline_closure(s):
                tn = ['p%d' % i for i in range(t.nd)]
                w('%s v_%s = (%s)0;' % (c_type(t.func.typ), t.name, c_type(t.func.typ)), 3)
                w('{', 3)
                w('bool pmx_w = false;', 4)
                emit_stage_value(t, 'v_' + t.name, tn, 4, False)
                w('(void)pmx_w;', 4)
                w('}', 3)
            w('%s v = (%s)0;' % (c_type(s.func.typ), c_type(s.func.typ)), 3)
            emit_stage_value(s, 'v', pnames, 3, True)
            w('if (pmx_w) {', 3)
            if s.needs_lds:
                w('L_%s[idx] = v;' % s.name, 4)
            if s.liveout:
                oc = []
                if s is not plan.anchor:
                    for i in range(s.nd):
                        lo, hi = s.own[i]
                        oc.append('p%d >= %s && p%d <= %s' % (i, R(lo), i, R(hi)))
                st = '%s[%s] = v;' % (_gname(s.func), gindex(s.func, pnames))
                if oc:
                    w('if (%s) %s' % (' && '.join(oc), st), 4)
                else:
                    w(st, 4)
            w('}', 3)
            w('}', 2)
            w('}')
        if ph + 1 < plan.nphases:
            w('__syncthreads();')
    grid_c = ' * '.join('((%s + %d) / %d)' % (_extent_c(a.func, d, base_pr), T - 1, T)
                        for (d, T) in grid) or '1'
    info = KernelInfo(kname, 'group', buffers, grid_c, threads, plan.lds_bytes, plan, reads, writes)
    return wrap_kernel(kname, params, buffers, writes, threads, plan.lds_bytes, L), info




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


def wrap_kernel(kname, params, buffers, writes, threads, lds_bytes, body_lines):
    """The kernel body as a device function plus two entry points:

    ``<k>``       one workgroup per tile (grid = number of tiles);
    ``<k>_batch`` persistent: a fixed set of workgroups pulls (frame, tile)
                  work items from an atomic counter until the queue over all
                  frames is empty; frames are described by ``<k>_frame``
                  records (buffers, first global work item, parameters).
    Every workgroup of the persistent form leaves the loop once the counter
    passes ``total``, so the grid always drains."""
    decls, names, fields = kernel_args(params, buffers, writes)
    lds = max(16, lds_bytes)
    L = ['__device__ __forceinline__ void %s_body(%s) {' % (
        kname, ', '.join(decls + ['const int pmx_tile', 'const int pmx_remap',
                                  'unsigned char* __restrict__ pmx_lds']))]
    L += ['  (void)pmx_remap; (void)pmx_lds; (void)pmx_err;']
    L += body_lines
    L.append('}')
    L.append('extern "C" __global__ __launch_bounds__(%d) void %s(%s) {' % (
        threads, kname, ', '.join(decls)))
    L.append('  __shared__ __align__(16) unsigned char pmx_lds[%d];' % lds)
    L.append('  %s_body(%s);' % (kname, ', '.join(names + ['(int)blockIdx.x', '1', 'pmx_lds'])))
    L.append('}')
    L.append('struct %s_frame { %s };' % (kname, ' '.join(fields)))
    L.append('extern "C" __global__ __launch_bounds__(%d) void %s_batch(const %s_frame* __restrict__ F, '
             'const int nframes, unsigned long long* __restrict__ counter, const long long total) {'
             % (threads, kname, kname))
    L.append('  __shared__ __align__(16) unsigned char pmx_lds[%d];' % lds)
    L.append('  __shared__ long long pmx_work;')
    L.append('  for (;;) {')
    L.append('    __syncthreads();')
    L.append('    if (threadIdx.x == 0) pmx_work = (long long)atomicAdd(counter, 1ULL);')
    L.append('    __syncthreads();')
    L.append('    const long long t = pmx_work;')
    L.append('    if (t >= total) break;')
    L.append('    int lo = 0, hi = nframes - 1;')
    L.append('    while (lo < hi) { const int mid = (lo + hi + 1) >> 1; if (F[mid].pmx_tile0 <= t) lo = mid; else hi = mid - 1; }')
    L.append('    const %s_frame& f = F[lo];' % kname)
    L.append('    %s_body(%s);' % (kname, ', '.join(['f.' + n for n in names] +
                                                   ['(int)(t - f.pmx_tile0)', '0', 'pmx_lds'])))
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
            src, info = emit_reduction_kernel(pipe, plan, kname, mode)
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
