"""Planning of one fused stencil group for the GPU: tile, regions, storage.

One group becomes one HIP kernel and one workgroup per overlapped tile
(the GPU form of the reference's OpenMP tile loop, codegen.py:411-551 and
poly_schedule.py:429-536).  Planning steps:

1. anchor: the group's last live-out; the tile is a box of the anchor's
   domain.  ``tile_sizes`` are consumed by anchor dimensions longer than
   the tile, in order, exactly as overlap_tile does (poly_schedule.py:
   433-451); untiled dimensions along which every in-group dependence is
   the identity (channels) are split one point per workgroup.
2. alignment/scale: every stage dimension that is a monotone image of an
   anchor dimension inherits its alignment and rational scale (the role of
   align_scale.py:93-854).  Live-outs other than the anchor own the points
   whose scaled anchor coordinate falls in the tile, so tiles partition
   every live-out and each point is written once.
3. regions: exact per-tile boxes back-propagated through the references of
   each Case part (``region_ir``), clipped to the domain and to the hull of
   the stage's Case boxes.  No stage reads an uncomputed producer point,
   unlike the reference trapezoid (its edge points read scratch index -1).
4. inlining: stages read only through identity accesses and not live-out
   are evaluated in registers inside their consumers (cast to their own
   type, so rounding is unchanged); the others are materialised in LDS.
5. phases: materialised stages are ordered in phases separated by
   workgroup barriers; LDS buffers are reused across phases by storage
   class and liveness, like storage_mapping.py:209-682 reuses scratchpads.
"""

from fractions import Fraction
import itertools
import math

from .analysis import (PlanError, classify_access, case_parts, ceval_int,
                       eval_bexpr, shift_access, scaled_access)
from .constructs import Reference, Reduction, Variable
from .region_ir import RegionIR, BIG

LDS_LIMIT_BYTES = 160 * 1024
CHEAP_OPS = 4          # operations of a pointwise stage recomputed per tap


# longest leading dimension kept whole in a tile (gpu option kblock)
KBLOCK_MAX = 16


class StageInfo(object):
    def __init__(self, comp, index):
        self.comp = comp
        self.func = comp.func
        self.index = index
        self.nd = len(comp.func.variables)
        self.parts = comp.parts
        self.liveout = False
        self.inline = False
        self.consumers = []        # [(StageInfo consumer, part, Reference)]
        self.align = [None] * self.nd   # (anchor dim, Fraction scale)
        self.region = None              # [(lo_id, hi_id)]
        self.own = None
        self.static_ext = None
        self.phase = None
        self.lds_offset = None
        self.lds_class = None
        self.needs_lds = False
        self.lane = False               # kept in registers, read across lanes
        self.vread = False              # read by a 4-wide block in 128-bit rows

    def reset(self):
        """Forget the analysis results (parts and live-out flag stay)."""
        self.inline = False
        self.consumers = []
        self.align = [None] * self.nd
        self.region = None
        self.own = None
        self.static_ext = None
        self.phase = None
        self.lds_offset = None
        self.lds_class = None
        self.needs_lds = False
        self.lane = False
        self.vread = False

    @property
    def name(self):
        return self.func.name

    def __repr__(self):
        return 'Stage(%s)' % self.name


class _VirtualComp(object):
    """Stand-in compute object of a staging copy stage (not in the
    pipeline's graph; it exists only inside one group kernel)."""

    def __init__(self, func, parts):
        self.func = func
        self.parts = parts
        self.level = -1
        self.is_reduction = False
        self.is_const = False


class GroupPlan(object):
    def __init__(self, pipe, group, gindex, tile_sizes=None, opts=None):
        self.pipe = pipe
        self.group = group
        self.gindex = gindex
        self.opts = dict(pipe.gpu_options)
        if opts:
            self.opts.update(opts)
        if tile_sizes is None:
            # gpu_options['group_tiles']: {function name: tile sizes} for the
            # group holding that function (a grid-resolution group next to
            # pixel-resolution ones); other groups use the pipeline's
            own = self.opts.get('group_tiles') or {}
            picks = dict((c.func.name, tuple(own[c.func.name])) for c in group.comps
                         if c.func.name in own)
            if len(set(picks.values())) > 1:
                raise ValueError("gpu_options['group_tiles']: functions of one group ask for "
                                 "different tiles: %s" % sorted(picks.items()))
            if picks:
                tile_sizes = list(next(iter(picks.values())))
        self.tile_sizes = list(tile_sizes if tile_sizes is not None else pipe.tile_sizes)
        self.env = pipe.param_env
        comps = group.comps
        if any(c.is_reduction for c in comps) or any(c.is_const for c in comps):
            raise PlanError("not a stencil group")
        self.staged = []
        self._virtual = {}
        # staging trades LDS (occupancy) for fewer L2/HBM re-reads; measured
        # slower on unsharp at 32x256, so it is an option the tuner can try
        self._build(stage=bool(self.opts.get('stage')))

    def _build(self, stage):
        pipe, group = self.pipe, self.group
        comps = group.comps
        self.stages = [StageInfo(c, i) for i, c in enumerate(comps)]
        self.by_comp = dict((s.comp, s) for s in self.stages)
        self._virtual = {}
        self.staged = []
        liveouts = pipe.liveouts(group)
        for s in self.stages:
            s.liveout = s.comp in liveouts
        self._offset_inline()
        self._analyse()
        if stage and self._stage_inputs():
            for s in self.stages:
                s.reset()
            try:
                self._analyse()
                self._blocks()
                self._storage()
                return
            except PlanError:
                # staging does not fit: plan without it
                self._build(stage=False)
                return
        self._blocks()
        self._lane_fuse()
        self._storage()

    def _analyse(self):
        self._consumers()
        self._anchor()
        self._align()
        self._grid()
        self._inline()
        self._regions()
        self._sizes()
        self._phases()

    def _stage_inputs(self):
        """Stage in LDS every out-of-group array that stages of different
        phases read: a copy stage, computed first, holds the union of the
        points the group reads (128-bit coalesced loads, each element read
        from HBM once per tile) and the readers take them from LDS.
        Unsharp's blur reads the image in the first phase and the mask
        reads it again in the last one; without staging that second read
        goes back to L2/HBM.  Returns True when stages were added."""
        from .analysis import Part
        from .constructs import Function, Interval, Variable
        from .expr_ast import Reference as Ref
        from .expr_types import Int
        from .expression import substitute_refs, _subst_cond
        phases = {}
        for s in self.stages:
            if s.inline:
                continue
            for (t, r) in self.closure_refs(s):
                f = r.objectRef
                if self.in_group(f) or not f.variables:
                    continue
                phases.setdefault(f, set()).add(s.phase)
        cands = [f for f, ph in phases.items() if len(ph) >= 2]
        if not cands:
            return False
        for g in sorted(cands, key=lambda f: f.name):
            vs = [Variable(Int, '_st%d' % i) for i in range(len(g.variables))]
            dom = [Interval(Int, d.lowerBound, d.upperBound) for d in g.domain]
            vf = Function((vs, dom), g.typ, 'stg_' + g.name)
            comp = _VirtualComp(vf, [Part([(None, None)] * len(vs), [], Ref(g, list(vs)), 0)])
            st = StageInfo(comp, len(self.stages))
            self._virtual[vf] = st
            for s in self.stages:
                new_parts = []
                for part in s.parts:
                    m = {}
                    for r in self._part_refs(part):
                        if r.objectRef is g:
                            m[r] = Ref(vf, list(r.arguments))
                    if not m:
                        new_parts.append(part)
                        continue
                    new_parts.append(Part(part.box,
                                          [_subst_cond(c, lambda e: substitute_refs(e, m))
                                           for c in part.pred],
                                          substitute_refs(part.expr, m), part.case_index))
                s.parts = new_parts
            self.stages.insert(0, st)
            self.staged.append(g.name)
        for i, s in enumerate(self.stages):
            s.index = i
        self.by_comp = dict((s.comp, s) for s in self.stages)
        return True

    # -- structure -------------------------------------------------------------
    def in_group(self, func):
        if func in self._virtual:
            return True
        c = self.pipe.func_map.get(func)
        return c is not None and c in self.by_comp

    def stage_of(self, func):
        if func in self._virtual:
            return self._virtual[func]
        return self.by_comp[self.pipe.func_map[func]]

    def _part_refs(self, part):
        refs = part.expr.collect(Reference)
        for c in part.pred:
            refs += [r for r in c.collect(Reference) if r not in refs]
        return refs

    def _offset_inline(self):
        """Recompute cheap pointwise stages at every tap of their stencil
        consumers instead of storing them.

        A stage that is not a live-out, has one unconditional definition of
        at most CHEAP_OPS operations and reads its producers only at its own
        point (``Ixx = Ix * Ix``; a plane copy ``r_r(x, y) =
        deinterleaved(1, x, y)`` too) is substituted into its in-group consumers
        when one of them reads it with an offset (``Sxx`` sums Ixx over a
        3x3 window): the consumer then reads the producers at the shifted
        points.  The value is unchanged (the substituted expression is cast
        to the stage's type), the stage disappears
        from the group, and its producers, now read with offsets, are the
        ones kept in LDS: Harris stores Ix and Iy instead of the three
        products, as the reference's generated code does
        (ipp_bench/harris/harris_function.cpp:17-21)."""
        if self.opts.get('no_offset_inline'):
            return
        from .analysis import Part
        from .expression import substitute_vars, substitute_refs
        from .expr_ast import (AbstractBinaryOpNode, AbstractUnaryOpNode, Select,
                               InbuiltFunction, Cast)
        changed = True
        while changed:
            changed = False
            for p in self.stages:
                if p.liveout or len(p.parts) != 1 or p.parts[0].pred:
                    continue
                part = p.parts[0]
                e = part.expr
                if e.has(Select) or e.has(InbuiltFunction):
                    continue
                nops = 0
                stack = [e]
                while stack:
                    x = stack.pop()
                    if isinstance(x, (AbstractBinaryOpNode, AbstractUnaryOpNode, Cast)):
                        nops += 1
                    if hasattr(x, 'children'):
                        stack.extend(x.children())
                if nops > CHEAP_OPS:
                    continue
                refs = e.collect(Reference)
                if not refs or not all(self._own_point(r, p) for r in refs):
                    continue
                # its consumers in the group; at least one reads with an offset
                cons = []
                offset = False
                for c in self.stages:
                    for cp in c.parts:
                        for r in self._part_refs(cp):
                            if r.objectRef is p.func:
                                cons.append(c)
                                if len(r.arguments) != c.nd or \
                                        any(a is not v for a, v in zip(r.arguments, c.func.variables)):
                                    offset = True
                if not offset or p in cons:
                    continue
                if not self._taps_inside(p, part):
                    continue
                # always cast: literals print as C doubles, so the C type of
                # the expression can be wider than the DSL type says
                body = Cast(p.func.typ, e)
                for c in set(cons):
                    new_parts = []
                    for cp in c.parts:
                        m = {}
                        for r in self._part_refs(cp):
                            if r.objectRef is p.func:
                                m[r] = substitute_vars(body, dict(zip(p.func.variables, r.arguments)))
                        if not m:
                            new_parts.append(cp)
                            continue
                        new_parts.append(Part(cp.box, [substitute_refs(q, m) if False else q
                                                       for q in cp.pred],
                                              substitute_refs(cp.expr, m), cp.case_index))
                        if any(r.objectRef is p.func for q in cp.pred for r in q.collect(Reference)):
                            raise PlanError("offset inlining into a predicate")
                    c.parts = new_parts
                self.stages.remove(p)
                self.offset_inlined = getattr(self, 'offset_inlined', []) + [p.name]
                changed = True
                break
        self.by_comp = dict((s.comp, s) for s in self.stages)

    @staticmethod
    def _own_point(r, p):
        """True when reference r reads its function at p's own point: its
        arguments are p's variables in order, with integer constants
        allowed in between (a plane of a multi-channel producer, like a
        camera pipeline's ``r_r(x, y) = deinterleaved(1, x, y)``)."""
        from .expr_ast import Value
        vs = []
        for a in r.arguments:
            if isinstance(a, Value) and isinstance(a.value, int):
                continue
            vs.append(a)
        return len(vs) == len(p.func.variables) and \
            all(a is v for a, v in zip(vs, p.func.variables))

    def _taps_inside(self, p, part):
        """True when every point at which an in-group consumer reads p
        (over the consumer part's box) lies in p's Case box, for all
        parameter values: then substituting p's expression reads nothing
        the original program leaves undefined."""
        from .expression import isAffine, get_affine_var_and_param_coeff, get_constant_from_expr
        from .expr_ast import Value

        def const_diff(a, b):
            e = Value.numericToValue(a) - Value.numericToValue(b)
            if not isAffine(e, include_div=False):
                return None
            if any(v != 0 for v in get_affine_var_and_param_coeff(e).values()):
                return None
            return get_constant_from_expr(e, affine=True)

        for c in self.stages:
            for cp in c.parts:
                for r in self._part_refs(cp):
                    if r.objectRef is not p.func:
                        continue
                    for i, arg in enumerate(r.arguments):
                        sa = shift_access(arg, c.func.variables)
                        if sa is None:
                            return False
                        plo, phi = part.box[i]
                        plo = p.func.domain[i].lowerBound if plo is None else plo
                        phi = p.func.domain[i].upperBound if phi is None else phi
                        if sa[0] == 'const':
                            lo = hi = sa[1]
                        else:
                            k, off = sa
                            clo, chi = cp.box[k]
                            clo = c.func.domain[k].lowerBound if clo is None else clo
                            chi = c.func.domain[k].upperBound if chi is None else chi
                            lo, hi = clo + off, chi + off
                        d1 = const_diff(lo, plo)
                        d2 = const_diff(phi, hi)
                        if d1 is None or d2 is None or d1 < 0 or d2 < 0:
                            return False
        return True

    def _consumers(self):
        for s in self.stages:
            for part in s.parts:
                for r in self._part_refs(part):
                    if self.in_group(r.objectRef):
                        p = self.stage_of(r.objectRef)
                        p.consumers.append((s, part, r))

    def _anchor(self):
        lo = [s for s in self.stages if s.liveout]
        self.anchor = max(lo, key=lambda s: (s.comp.level, self.pipe.comp_size(s.comp), -s.index))

    def _grid(self):
        best = self.anchor
        dom = self.pipe.domain_box(best.func)
        self.grid_dims = []      # [(anchor dim, tile size)]
        k = 0
        for d in range(best.nd):
            ext = dom[d][1] - dom[d][0] + 1
            if k < len(self.tile_sizes) and ext > self.tile_sizes[k]:
                self.grid_dims.append((d, int(self.tile_sizes[k])))
                k += 1
            elif d < best.nd - 1 and self._identity_dim(d):
                if self.opts.get('kblock') and d == 0 and best.nd >= 3 and ext <= KBLOCK_MAX:
                    # (gpu option kblock) a short leading dimension (the
                    # local Laplacian's K levels) stays whole in the tile;
                    # threads then block over it (_blocks)
                    continue
                # (never the innermost dimension: splitting it would leave
                # each workgroup a strided, uncoalesced slice of the rows)
                self.grid_dims.append((d, 1))
        # split dimensions whose workgroups read the same input tile (the
        # local Laplacian's K levels all read the gray image) vary fastest
        # in the workgroup order, so those workgroups run together on one
        # XCD and the shared tile is read from HBM once, not once per level
        self.split_fast = [d for (d, T) in self.grid_dims
                           if T == 1 and self._split_shares_inputs(d)]

    def _split_shares_inputs(self, d):
        """True when some stage reads a buffer from outside the group at an
        index that does not involve its dimension aligned with anchor dim
        d: every workgroup of a split along d then reads the same points."""
        for s in self.stages:
            own = [s.func.variables[i] for i in range(s.nd)
                   if s.align[i] is not None and s.align[i][0] == d]
            if not own:
                continue
            for part in s.parts:
                for r in self._part_refs(part):
                    if self.in_group(r.objectRef):
                        continue
                    used = set()
                    for arg in r.arguments:
                        used.update(arg.collect(Variable) if hasattr(arg, 'collect') else [])
                    if not any(v in used for v in own):
                        return True
        return False

    def _identity_dim(self, d):
        """True when every in-group dependence along anchor dim d is the
        identity and every stage has a dimension aligned with d, so
        splitting it one point per tile costs no halo and no recompute (a
        stage without such a dimension, like the demosaic stages feeding a
        camera pipeline's colour-channel output, would be computed again by
        every workgroup of the split)."""
        for s in self.stages:
            if not any(al is not None and al[0] == d for al in s.align):
                return False
        for s in self.stages:
            for (c, part, r) in s.consumers:
                for i, arg in enumerate(r.arguments):
                    if s.align[i] is None or s.align[i][0] != d:
                        continue
                    acc = classify_access(arg, c.func.variables)
                    if acc.kind != 'var' or not acc.identity:
                        return False
                    if c.align[acc.dim] is None or c.align[acc.dim][0] != d:
                        return False
        return True

    def _align(self):
        a = self.anchor
        a.align = [(d, Fraction(1)) for d in range(a.nd)]
        for s in reversed(self.stages):
            if s is a:
                continue
            cand = [None] * s.nd
            for (c, part, r) in s.consumers:
                for i, arg in enumerate(r.arguments):
                    acc = classify_access(arg, c.func.variables)
                    if acc.kind != 'var' or c.align[acc.dim] is None:
                        continue
                    ad, sc = c.align[acc.dim]
                    if cand[i] is None:
                        cand[i] = (ad, sc * acc.slope)
            s.align = cand

    def _inline(self):
        if self.opts.get('no_inline'):
            return
        for s in self.stages:
            if s.liveout or not s.consumers:
                continue
            ok = True
            for (c, part, r) in s.consumers:
                if len(r.arguments) != c.nd or s.nd != c.nd:
                    ok = False
                    break
                for d, arg in enumerate(r.arguments):
                    if arg is not c.func.variables[d]:
                        ok = False
                        break
            s.inline = ok

    # -- regions -----------------------------------------------------------------
    def _regions(self):
        ir = RegionIR()
        self.ir = ir
        pipe = self.pipe
        a = self.anchor
        big = ir.const(BIG)
        nbig = ir.const(-BIG)
        # anchor tile box and tile counts
        self.ntiles = {}
        tile_box = []
        for d in range(a.nd):
            lb = ir.expr(a.func.domain[d].lowerBound)
            ub = ir.expr(a.func.domain[d].upperBound)
            tile_box.append((lb, ub, None))
        for (d, T) in self.grid_dims:
            lb, ub, _ = tile_box[d]
            t = ir.tile(d)
            lo = ir.add(lb, ir.mul(t, T))
            hi = ir.add(lo, ir.const(T - 1))
            tile_box[d] = (lo, hi, t)
        self.tile_box = tile_box

        def dom_ids(s, i):
            return (ir.expr(s.func.domain[i].lowerBound),
                    ir.expr(s.func.domain[i].upperBound))

        def case_hull(s, i):
            lo_h = hi_h = None
            dlo, dhi = dom_ids(s, i)
            for part in s.parts:
                blo, bhi = part.box[i]
                plo = ir.expr(blo) if blo is not None else dlo
                phi = ir.expr(bhi) if bhi is not None else dhi
                lo_h = plo if lo_h is None else ir.min(lo_h, plo)
                hi_h = phi if hi_h is None else ir.max(hi_h, phi)
            return lo_h, hi_h

        # ownership of live-outs (scaled tile ranges, open at the edges)
        ntile_ids = {}
        for (d, T) in self.grid_dims:
            lb = ir.expr(a.func.domain[d].lowerBound)
            ub = ir.expr(a.func.domain[d].upperBound)
            ext = ir.add(ir.sub(ub, lb), ir.const(1))
            nt = ir.fdiv(ir.add(ext, ir.const(T - 1)), T)
            ntile_ids[d] = nt
        self.ntile_ids = ntile_ids

        def own_range(s, i):
            al = s.align[i]
            dlo, dhi = dom_ids(s, i)
            if al is None:
                return dlo, dhi
            ad, sc = al
            lo_t, hi_t, t = tile_box[ad]
            if t is None:
                return dlo, dhi
            last = ir.sub(ntile_ids[ad], ir.const(1))
            lo_open = ir.sel_eq(t, ir.const(0), nbig, lo_t)
            hi_open = ir.sel_eq(t, last, big, hi_t)
            if sc == 1:
                return lo_open, hi_open
            if sc <= 0:
                raise PlanError("live-out %s has a non-positive scale" % s.name)
            P, Q = sc.numerator, sc.denominator
            # lo = ceil(lo_t * P / Q); hi = ceil((hi_t + 1) * P / Q) - 1
            lo_s = ir.sub(ir.const(0), ir.fdiv(ir.sub(ir.const(0), ir.mul(lo_t, P)), Q))
            hi_s = ir.sub(ir.sub(ir.const(0), ir.fdiv(ir.sub(ir.const(0),
                                                            ir.mul(ir.add(hi_t, ir.const(1)), P)), Q)),
                          ir.const(1))
            lo_s = ir.sel_eq(t, ir.const(0), nbig, lo_s)
            hi_s = ir.sel_eq(t, last, big, hi_s)
            return lo_s, hi_s

        self._requests = []     # (ref, dim, producer, valid boxes, lo id, hi id)
        for s in reversed(self.stages):
            lo = [big] * s.nd
            hi = [nbig] * s.nd
            if s.liveout:
                own = []
                for i in range(s.nd):
                    if s is a:
                        lo_t, hi_t, t = tile_box[i]
                        own.append((lo_t, hi_t))
                    else:
                        own.append(own_range(s, i))
                s.own = own
                for i in range(s.nd):
                    lo[i] = ir.min(lo[i], own[i][0])
                    hi[i] = ir.max(hi[i], own[i][1])
            # the references of one consumer part share its sub-region, so
            # their requests are united first and guarded once; requests
            # that shift the same sub-region bound by constants fold into
            # that bound plus the extreme shift (a 3x3 stencil's nine taps
            # give one min and one max per dimension, not nine)
            groups = []
            gidx = {}
            for (c, part, r) in s.consumers:
                k = (id(c), id(part))
                if k not in gidx:
                    gidx[k] = len(groups)
                    groups.append((c, part, []))
                groups[gidx[k]][2].append(r)
            for (c, part, refs) in groups:
                # sub-region of the consumer that this part covers
                sub = []
                for j in range(c.nd):
                    clo, chi = c.region[j]
                    blo, bhi = part.box[j]
                    if blo is not None:
                        clo = ir.max(clo, ir.expr(blo))
                    if bhi is not None:
                        chi = ir.min(chi, ir.expr(bhi))
                    sub.append((clo, chi))
                valid = [(l, h) for (l, h) in sub]
                want_lo = [None] * s.nd
                want_hi = [None] * s.nd
                shifts = {}
                for r in refs:
                    for i, arg in enumerate(r.arguments):
                        acc = classify_access(arg, c.func.variables)
                        sa = shift_access(arg, c.func.variables)
                        if acc.kind == 'var':
                            l, h = sub[acc.dim]
                            al = ir.access(arg, acc.var, l)
                            ah = ir.access(arg, acc.var, h)
                            if not acc.increasing:
                                al, ah = ah, al
                        elif acc.kind == 'const':
                            al = ah = ir.expr(arg)
                        else:
                            al, ah = dom_ids(s, i)
                        self._requests.append((r, i, s, valid, al, ah))
                        if acc.kind == 'var' and sa is not None and sa[0] != 'const' \
                                and isinstance(sa[1], int):
                            key = (i, sa[0])
                            mn, mx = shifts.get(key, (sa[1], sa[1]))
                            shifts[key] = (min(mn, sa[1]), max(mx, sa[1]))
                            continue
                        want_lo[i] = al if want_lo[i] is None else ir.min(want_lo[i], al)
                        want_hi[i] = ah if want_hi[i] is None else ir.max(want_hi[i], ah)
                for (i, j), (mn, mx) in sorted(shifts.items()):
                    l, h = sub[j]
                    al = ir.add(l, ir.const(mn)) if mn else l
                    ah = ir.add(h, ir.const(mx)) if mx else h
                    want_lo[i] = al if want_lo[i] is None else ir.min(want_lo[i], al)
                    want_hi[i] = ah if want_hi[i] is None else ir.max(want_hi[i], ah)
                for i in range(s.nd):
                    if want_lo[i] is None:
                        continue
                    lo[i] = ir.sel_valid(valid, ir.min(lo[i], want_lo[i]), lo[i])
                    hi[i] = ir.sel_valid(valid, ir.max(hi[i], want_hi[i]), hi[i])
            reg = []
            for i in range(s.nd):
                dlo, dhi = dom_ids(s, i)
                hlo, hhi = case_hull(s, i)
                reg.append((ir.max(ir.max(lo[i], dlo), hlo),
                            ir.min(ir.min(hi[i], dhi), hhi)))
            s.region = reg

    def tile_space(self, env=None):
        env = self.env if env is None else env
        vals = self.ir.evaluate(env, {})  if False else None
        counts = []
        a = self.anchor
        for (d, T) in self.grid_dims:
            lb = ceval_int(a.func.domain[d].lowerBound, env)
            ub = ceval_int(a.func.domain[d].upperBound, env)
            counts.append(max(0, (ub - lb + 1 + T - 1) // T))
        return counts

    def _sizes(self):
        counts = self.tile_space()
        self.tile_counts = counts
        total = 1
        for n in counts:
            total *= n
        self.total_tiles = total
        # sample tiles: all if few, else a window around the middle and the
        # edges of every grid dimension (region extents are periodic in the
        # tile index with a period dividing the scale denominators)
        per_dim = []
        for n in counts:
            if n <= 96:
                per_dim.append(list(range(n)))
            else:
                mid = n // 2
                s = set(range(0, 8)) | set(range(n - 8, n)) | set(range(mid - 40, mid + 40))
                per_dim.append(sorted(x for x in s if 0 <= x < n))
        dims = [d for (d, T) in self.grid_dims]
        import numpy as np
        if dims:
            mesh = np.meshgrid(*[np.asarray(p, dtype=np.int64) for p in per_dim], indexing='ij')
            tiles = dict((d, m.ravel()) for d, m in zip(dims, mesh))
        else:
            tiles = {}
        vals = self.ir.evaluate_many(self.env, tiles)
        # references that can ask for producer points outside the producer's
        # computed region (its domain or Case hull clipped the request: the
        # program reads points it never defines) get clamped reads
        self.clamp_refs = set()
        for (r, i, p, valid, al, ah) in self._requests:
            ok = None
            for (l, h) in valid:
                c = np.asarray(vals[l]) <= np.asarray(vals[h])
                ok = c if ok is None else (ok & c)
            rl, rh = p.region[i]
            bad = (np.asarray(vals[al]) < np.asarray(vals[rl])) | \
                  (np.asarray(vals[ah]) > np.asarray(vals[rh]))
            if ok is not None:
                bad = bad & ok
            if np.any(bad):
                self.clamp_refs.add((id(r), i))
        ext = {}
        for s in self.stages:
            ext[s] = []
            for (l, h) in s.region:
                e = np.asarray(vals[h]) - np.asarray(vals[l]) + 1
                ext[s].append(int(np.max(e)) if np.size(e) else 0)
        for s in self.stages:
            s.static_ext = [max(1, e) for e in ext[s]]
        # sampled region bounds, to recognise stages swept over equal boxes
        self._region_samples = {}
        for s in self.stages:
            self._region_samples[s] = [(np.asarray(vals[l]), np.asarray(vals[h])) for (l, h) in s.region]

    # -- phases and storage ---------------------------------------------------------
    def _phases(self):
        memo = {}

        def dep_phase(s):
            # phase in which a value of s can first be computed
            if s in memo:
                return memo[s]
            ph = 0
            for part in s.parts:
                for r in self._part_refs(part):
                    if not self.in_group(r.objectRef):
                        continue
                    p = self.stage_of(r.objectRef)
                    if p is s:
                        raise PlanError("self reference in %s" % s.name)
                    if p.inline:
                        ph = max(ph, dep_phase(p))
                    else:
                        ph = max(ph, dep_phase(p) + 1)
            memo[s] = ph
            return ph
        for s in self.stages:
            if not s.inline:
                s.phase = dep_phase(s)
        self.nphases = 1 + max(s.phase for s in self.stages if not s.inline)

        # phases in which an inline stage gets evaluated
        def eval_phases(s):
            out = set()
            for (c, part, r) in s.consumers:
                if c.inline:
                    out |= eval_phases(c)
                else:
                    out.add(c.phase)
            return out
        self.eval_phases = {}
        for s in self.stages:
            if s.inline:
                self.eval_phases[s] = eval_phases(s)
        for s in self.stages:
            if s.inline:
                continue
            reads = set()
            for (c, part, r) in s.consumers:
                reads |= (eval_phases(c) if c.inline else {c.phase})
            s.needs_lds = bool(reads)
            s.last_read = max(reads) if reads else s.phase

    # -- per-thread register blocks ------------------------------------------------
    def closure_refs(self, s):
        """[(stage whose variables the reference's arguments use, Reference)]
        over the parts of s and of every stage inlined into it."""
        out = []
        seen = set()

        def visit(t):
            for part in t.parts:
                for r in self._part_refs(part):
                    out.append((t, r))
                    if self.in_group(r.objectRef):
                        p = self.stage_of(r.objectRef)
                        if p.inline and p not in seen:
                            seen.add(p)
                            visit(p)
        visit(s)
        return out

    def _blocks(self):
        """Choose each materialised stage's per-thread block: U consecutive
        points along its second-innermost dimension times V along the
        innermost one.  A thread computes the U*V points of an item in
        straight-line code, so loads (and their conversions) shared by
        neighbouring points are issued once and consecutive loads merge
        into 128-bit accesses.  V = 4 only for stages that read no LDS
        (lane-strided LDS accesses would conflict on banks), or whose LDS
        reads are all row runs (_vector_row_read; gpu option ``no_vlds``
        turns this off): the block then loads each row it touches with
        aligned 128-bit LDS reads, conflict-free; U and V
        minimise an estimate of thread time: sweep iterations times points
        per item times per-point cost, where the per-point load count
        shrinks with the reuse a block gives to the stage's stencil reach."""
        threads = int(self.opts.get('threads', 256))
        forced = self.opts.get('block')
        # fixed cost of one sweep iteration, in the same units as a point's
        # cost (8 + 2 per load)
        it_cost = float(self.opts.get('iter_cost', 6))
        for s in self.stages:
            s.block = (1, 1)
            s.bdims = (s.nd - 2 if s.nd >= 2 else None, s.nd - 1 if s.nd >= 1 else None)
            if s.inline or s.nd == 0:
                continue
            ext = s.static_ext
            rd, cd = s.bdims
            if ext[-1] < 16 and s.nd >= 3:
                # short innermost dimension (colour channels): lanes run over
                # the two innermost dimensions, blocks along the third
                rd = s.nd - 3
                s.bdims = (rd, cd)
            elif self._kblock_dim(s):
                # a whole short leading dimension: each thread computes all
                # of its points for one (row, column), so what they share
                # (loads, index arithmetic, a gray value's table index) is
                # done once
                s.bdims = (0, cd)
                s.block = (ext[0], 1)
                continue
            ext_c = ext[cd]
            ext_r = ext[rd] if rd is not None else 1
            outer = 1
            for i, e in enumerate(ext):
                if i not in (rd, cd):
                    outer *= e
            reads_lds = False
            vec_ok = not (self.opts.get('no_vlds') or self.opts.get('lanes'))
            lds_prods = []
            rv = rh = 0
            for (t, r) in self.closure_refs(s):
                if self.in_group(r.objectRef) and not self.stage_of(r.objectRef).inline:
                    reads_lds = True
                    p = self.stage_of(r.objectRef)
                    lds_prods.append(p)
                    vec_ok = vec_ok and self._vector_row_read(t, r, p, cd)
                for arg in r.arguments:
                    sa = shift_access(arg, t.func.variables)
                    if sa is None or sa[0] == 'const' or not isinstance(sa[1], int):
                        continue
                    if sa[0] == cd:
                        rh = max(rh, abs(sa[1]))
                    elif sa[0] == rd:
                        rv = max(rv, abs(sa[1]))
            lds_v = not reads_lds or vec_ok
            own = forced
            if reads_lds and self.opts.get('block_lds') is not None:
                # (gpu option block_lds: the block of the stages that read LDS)
                own = self.opts.get('block_lds')
            if own is not None:
                U, V = own
                U = U if rd is not None else 1
                V = V if (V in (1, 4) and lds_v and ext_c >= 4 * V) else 1
                s.block = (max(1, min(U, ext_r)), V)
                if V == 4:
                    for p in lds_prods:
                        p.vread = True
                continue
            Vs = [4, 1] if (lds_v and ext_c >= 16) else [1]
            best = None
            for V in Vs:
                for U in (1, 2, 4, 8):
                    if U > 1 and (rd is None or ext_r < 2 * U) or U * V > 16:
                        continue
                    if V == 4 and reads_lds and U > 2:
                        # the rows a 4-wide block loads from LDS stay in
                        # registers: taller blocks measured slower (unsharp's
                        # mask, Harris' sums: 2x4 best)
                        continue
                    items = outer * (-(-ext_r // U)) * (-(-ext_c // V))
                    iters = -(-items // threads)
                    taps = (U + 2.0 * rv) / U * (V + 2.0 * rh) / V
                    cost = iters * (U * V * (8.0 + 2.0 * taps) + it_cost)
                    key = (cost, -U * V)
                    if best is None or key < best[0]:
                        best = (key, (U, V))
            s.block = best[1]
            if s.block[1] == 4:
                for p in lds_prods:
                    p.vread = True
        self._sweeps()

    def _kblock_dim(self, s):
        """True when (gpu option kblock) stage s's leading dimension is a
        whole, unsplit, short dimension of the tile (2 to KBLOCK_MAX
        points) and s has at least three dimensions."""
        if not self.opts.get('kblock') or s.nd < 3:
            return False
        # (a leading dimension its consumers read at constants, like the
        # bilateral grid's value/weight pair, is aligned with no anchor
        # dimension and is whole in every tile)
        if s.align[0] is not None and any(d == s.align[0][0] for (d, T) in self.grid_dims):
            return False
        return 2 <= s.static_ext[0] <= KBLOCK_MAX

    def _vector_row_read(self, t, r, p, cd):
        """True when the LDS read ``r`` (in stage t's variables) of producer
        p can be served from 128-bit row loads by a 4-wide block: p holds
        4-byte elements, r's innermost argument is the block's innermost
        variable plus an integer constant, and its other arguments are
        shifts of other variables (or constants), so each row the block
        touches is a run of consecutive elements at a fixed offset."""
        if cd is None or p.func.typ.nbytes() != 4 or p.lane:
            return False
        tv = t.func.variables
        if len(r.arguments) != p.nd or p.nd == 0:
            return False
        for j, arg in enumerate(r.arguments):
            sa = shift_access(arg, tv)
            if sa is None:
                # a downsampling read (2 * y + c) of the row is a run too
                sc = scaled_access(arg, tv)
                if sc is None:
                    return False
                sa = (sc[0], sc[2])
            if j == p.nd - 1:
                if sa[0] != cd or not isinstance(sa[1], int):
                    return False
            elif sa[0] == cd:
                return False
        return True

    def _sweeps(self):
        """Stages of one phase with the same static box, block and (on every
        sampled tile) the same region are swept by one loop: each point of
        an item computes all of them, so loads they share (Ix and Iy read
        the same 3x3 neighbourhood) are issued once."""
        import numpy as np
        self.sweeps = []
        for ph in range(self.nphases):
            groups = []
            for s in self.stages:
                if s.inline or s.phase != ph:
                    continue
                placed = False
                if not self.opts.get('no_fuse') and s.nd > 0:
                    for g in groups:
                        t = g[0]
                        if (t.nd, t.static_ext, t.block, t.bdims) != (s.nd, s.static_ext, s.block, s.bdims):
                            continue
                        if all(np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
                               for a, b in zip(self._region_samples[t], self._region_samples[s])):
                            g.append(s)
                            placed = True
                            break
                if not placed:
                    groups.append([s])
            for g in groups:
                self.sweeps.append((ph, g))

    # widest column reach a lane-fused sweep reads across lanes
    LANE_REACH = 3

    def _lane_fuse(self):
        """Lane fusion (gpu option ``lanes``): a sweep whose stages are read
        only by the next phase's single stored stage, at small offsets
        (rows: any; innermost dimension: at most LANE_REACH), is not stored
        in LDS.  Each thread of the consumer's sweep computes those stages
        in registers for its column, on the rows its block needs; values of
        neighbouring columns come from neighbouring lanes of the wavefront
        (a lane shuffle).  A wavefront covers 64 consecutive columns, of
        which the 2*reach at its edges only feed their neighbours.  The
        barrier between the two phases and the LDS traffic go away.
        ``self.lane_pair`` = (producer sweep, consumer sweep, row reach below,
        row reach above, column reach) or None."""
        self.lane_pair = None
        if not self.opts.get('lanes'):
            return
        for li in range(len(self.sweeps) - 1):
            ph, prod = self.sweeps[li]
            ph2, cons = self.sweeps[li + 1]
            if ph2 != ph + 1 or len(cons) != 1:
                continue
            if any(q[0] == ph for q in self.sweeps if q is not self.sweeps[li]) or \
                    any(q[0] == ph2 for q in self.sweeps if q is not self.sweeps[li + 1]):
                continue
            c = cons[0]
            U, V = c.block
            if V != 1 or c.bdims[1] is None:
                continue
            rd, cd = c.bdims
            if any(m.liveout or m.nd != c.nd or m.bdims != c.bdims for m in prod):
                continue
            mset = set(prod)
            lo = hi = reach = 0
            ok = True
            closure = set([c])
            for (t, r) in self.closure_refs(c):
                if self.in_group(r.objectRef) and self.stage_of(r.objectRef).inline:
                    closure.add(self.stage_of(r.objectRef))
            for m in prod:
                for (cs, part, r) in m.consumers:
                    if cs not in closure:
                        ok = False
            for (t, r) in self.closure_refs(c):
                if not (self.in_group(r.objectRef) and self.stage_of(r.objectRef) in mset):
                    continue
                for j, arg in enumerate(r.arguments):
                    sa = shift_access(arg, t.func.variables)
                    if sa is None or sa[0] != j or not isinstance(sa[1], int):
                        ok = False
                        break
                    if j == cd:
                        reach = max(reach, abs(sa[1]))
                    elif j == rd:
                        lo = min(lo, sa[1])
                        hi = max(hi, sa[1])
                    elif sa[1] != 0:
                        ok = False
            if not ok or reach > self.LANE_REACH:
                continue
            for m in prod:
                m.needs_lds = False
                m.lane = True
            self.lane_pair = (li, li + 1, -lo, hi, reach)
            return

    def lds_ext(self, s):
        """LDS layout extents of a stage: the static box, with the innermost
        extent padded to a multiple of 4 for 4-wide blocks (aligned
        128-bit LDS stores)."""
        ext = list(s.static_ext)
        if s.block[1] == 4 or s.vread:
            ext[-1] = (ext[-1] + 3) // 4 * 4
        return ext

    def _storage(self):
        """LDS layout with reuse.  Storage class = (element type, rank), as
        the reference's optimize_storage keys classes on type and
        dimensionality and sizes them by the largest member
        (storage_mapping.py:152-299); a buffer whose last reader ran in an
        earlier phase is reused by a later stage of its class when large
        enough (best fit), otherwise new space is appended."""
        free = {}
        offset = 0
        order = sorted([s for s in self.stages if s.needs_lds],
                       key=lambda s: (s.phase, s.index))
        live = []
        for s in order:
            still = []
            for (lr, t) in live:
                if lr < s.phase:
                    free.setdefault(t.lds_class, []).append((t.lds_offset, t.lds_size))
                else:
                    still.append((lr, t))
            live = still
            nbytes = s.func.typ.nbytes()
            for e in self.lds_ext(s):
                nbytes *= e
            nbytes = (nbytes + 15) // 16 * 16
            if s.vread:
                # a row load of a 4-wide block may run up to three elements
                # past the last row (values no point uses)
                nbytes += 16
            cls = (s.func.typ.c_name(), s.nd)
            s.lds_class = cls
            cands = sorted([fs for fs in free.get(cls, []) if fs[1] >= nbytes], key=lambda fs: fs[1])
            if cands:
                slot = cands[0]
                free[cls].remove(slot)
                s.lds_offset, s.lds_size = slot
            else:
                s.lds_offset = offset
                s.lds_size = nbytes
                offset += nbytes
            live.append((s.last_read, s))
        self.lds_bytes = offset
        if self.lds_bytes > LDS_LIMIT_BYTES:
            raise PlanError("group %d needs %d bytes of LDS (> %d); use smaller tiles"
                            % (self.gindex, self.lds_bytes, LDS_LIMIT_BYTES))

    # -- reporting ------------------------------------------------------------------
    def describe(self):
        lines = ['group %d: anchor %s, grid dims %s, tiles %s, LDS %d B' % (
            self.gindex, self.anchor.name, self.grid_dims, self.tile_counts, self.lds_bytes)]
        for s in self.stages:
            kind = 'inline' if s.inline else 'phase %d%s' % (s.phase, ' lds@%d' % s.lds_offset
                                                           if s.needs_lds else '')
            lines.append('  %-12s %-18s ext=%s%s' % (s.name, kind, s.static_ext,
                                                   ' liveout' if s.liveout else ''))
        return '\n'.join(lines)
