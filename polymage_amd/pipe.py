"""The Pipeline: DAG of compute objects, passes, grouping and code emission.

Restates the planning half of sandbox/pipe.py:596-759 without isl:

  1. collect every Function/Image/Reduction reachable from the outputs
     (pipe.py:101-120) and clone the non-input ones (630-645);
  2. build compute objects with parent/child maps (828-858) and a level
     order (utils.py:120-138);
  3. a static bounds check of affine references (bounds.py:36-118);
  4. the inline pass for explicit ``inline_directives`` (inline.py:29-199);
  5. greedy grouping, ``auto_group`` (grouping.py:56-239), with two extra
     fusion rules the GPU tile needs (see ``Pipeline._fusible``);
  6. level-ordered group schedule (schedule.py:85-89) and liveness
     (liveness.py:35-54);
  7. ``generate_code`` emits HIP for every group (one kernel per group, one
     workgroup per overlapped tile) plus a C++ host stub that exports the
     reference's ``pipeline_<name>`` entry point (codegen.py:920-1067).

Unlike the reference (pipe.py:220-486 iterate ``set``s of id-hashed objects)
every ordering here is deterministic.
"""

import logging

from .constructs import (Function, Image, Reduction, Case, Reduce,
                         Parameter, Variable, Reference, Condition, Value)
from .expr_types import Int
from .expression import substitute_vars, substitute_refs
from .analysis import (PlanError, case_parts, classify_access, ceval_int,
                       eval_bexpr, interval_image, unconditional_refs)
from .liveness import compute_liveness, level_order

LOG = logging.getLogger("polymage_amd.pipe")


# ---------------------------------------------------------------------------
# DAG collection
# ---------------------------------------------------------------------------

def _def_refs(func):
    """References appearing anywhere in a function's definition."""
    refs = []
    for case in func.defn:
        if isinstance(case, Case):
            refs += case.condition.collect(Reference)
            e = case.expression
        else:
            e = case
        if isinstance(e, Reduce):
            refs += [r for r in e.accumulate_ref.collect(Reference)
                     if r is not e.accumulate_ref]
            refs += e.expression.collect(Reference)
        else:
            refs += e.collect(Reference)
    if isinstance(func, Reduction):
        refs += func.default.collect(Reference)
    return refs


def get_funcs(outputs):
    """BFS from the outputs over references; returns (funcs, parents,
    children) with deterministic orders."""
    funcs = []
    parents = {}
    children = {}
    queue = list(outputs)
    seen = set()
    while queue:
        f = queue.pop(0)
        if id(f) in seen:
            continue
        seen.add(id(f))
        funcs.append(f)
        parents.setdefault(f, [])
        children.setdefault(f, [])
        if isinstance(f, Image):
            continue
        for r in _def_refs(f):
            p = r.objectRef
            if p is f:
                continue       # self reference (reductions)
            if p not in parents[f]:
                parents[f].append(p)
            children.setdefault(p, [])
            if f not in children[p]:
                children[p].append(f)
            if id(p) not in seen:
                queue.append(p)
    return funcs, parents, children


class ComputeObject(object):
    def __init__(self, func, is_output=False):
        self.func = func
        self.is_output = is_output
        self.parents = []
        self.children = []
        self.level = 0
        self.group = None
        self.parts = None        # list of analysis.Part for Functions

    @property
    def name(self):
        return self.func.name

    @property
    def is_image(self):
        return isinstance(self.func, Image)

    @property
    def is_reduction(self):
        return isinstance(self.func, Reduction)

    @property
    def is_const(self):
        return (not self.is_image) and self.func.is_const_func

    def __repr__(self):
        return 'Comp(%s)' % self.func.name


class Group(object):
    def __init__(self, comps):
        self.comps = list(comps)
        self.parents = []
        self.children = []

    @property
    def name(self):
        return '[' + ', '.join(c.name for c in self.comps) + ']'

    def __repr__(self):
        return 'Group' + self.name


class DotGraph(object):
    """Minimal stand-in for the pygraphviz graph the reference returns
    (pipe.py:901-932): DOT text with ``.write(path)`` and ``str()``."""

    def __init__(self, text):
        self.text = text

    def write(self, path):
        with open(path, 'w') as f:
            f.write(self.text)

    def __str__(self):
        return self.text

    def to_string(self):
        return self.text


class Pipeline(object):
    def __init__(self, _outputs, _param_estimates=None, _param_constraints=None,
                 _grouping=None, _group_size=None, _inline_directives=None,
                 _tile_sizes=None, _size_threshold=None, _options=None,
                 _name=None, _literal_mode='double', _gpu_options=None):
        if _name is None:
            _name = ''.join(o.name for o in _outputs)
        self._name = _name
        self._orig_outputs = list(_outputs)
        self._param_estimates_list = list(_param_estimates or [])
        self._param_constraints = list(_param_constraints or [])
        self._grouping = list(_grouping or [])
        self._group_size = _group_size if _group_size is not None else 5
        self._inline_directives = list(_inline_directives or [])
        self._tile_sizes = list(_tile_sizes or [])
        self._size_threshold = _size_threshold if _size_threshold is not None else 200 * 200
        self._options = list(_options or [])
        self._literal_mode = _literal_mode
        self._gpu_options = dict(_gpu_options or {})

        # -- DAG of clones ------------------------------------------------
        orig_funcs, _, _ = get_funcs(self._orig_outputs)
        self._clone_map = {}
        self._inputs = []
        for f in orig_funcs:
            if isinstance(f, Image):
                self._clone_map[f] = f
                self._inputs.append(f)
            else:
                self._clone_map[f] = f.clone()
        self._outputs = [self._clone_map[o] for o in self._orig_outputs]
        for f in orig_funcs:
            cln = self._clone_map[f]
            if isinstance(cln, Image):
                continue
            for r in cln.getObjects(Reference):
                if not isinstance(r.objectRef, Image):
                    r._replace_ref_object(self._clone_map[r.objectRef])

        self._env = self._estimate_env()
        self._build_comps()

        # -- passes ---------------------------------------------------------
        self.bounds_check()
        self.inline_pass()
        self._build_comps()
        if not (self._gpu_options or {}).get('no_gather_inline'):
            self.gather_inline_pass()
        self._groups = [Group([c]) for c in self._comps]
        for g in self._groups:
            g.comps[0].group = g
        self._link_groups()
        if self._grouping:
            self.manual_group()
        else:
            self.auto_group()
        self._order_groups()

        # liveness of live-outs over the group schedule (liveness.py:57-80)
        sched = dict((g, i) for i, g in enumerate(self._groups))
        ch = dict((g, g.children) for g in self._groups)
        self._group_liveness = compute_liveness(ch, sched)

    # -- properties --------------------------------------------------------
    @property
    def name(self):
        return self._name

    @property
    def inputs(self):
        return self._inputs

    @property
    def outputs(self):
        return self._outputs

    @property
    def orig_outputs(self):
        return self._orig_outputs

    @property
    def comps(self):
        return self._comps

    @property
    def func_map(self):
        return self._func_map

    @property
    def groups(self):
        return self._groups

    @property
    def group_schedule(self):
        return dict((g, i) for i, g in enumerate(self._groups))

    @property
    def options(self):
        return self._options

    @property
    def tile_sizes(self):
        return self._tile_sizes

    @property
    def group_size(self):
        return self._group_size

    @property
    def literal_mode(self):
        return self._literal_mode

    @property
    def gpu_options(self):
        return self._gpu_options

    @property
    def param_env(self):
        """{Parameter: int} from estimates and equality constraints."""
        return dict(self._env)

    def get_parameters(self):
        params = []
        seen = set()
        for c in self._comps:
            objs = c.func.getObjects(Parameter) if not c.is_image else []
            if isinstance(c.func, Reduction):
                pass
            for p in objs:
                if id(p) not in seen:
                    seen.add(id(p))
                    params.append(p)
        for img in self._inputs:
            for d in img.dimensions:
                for p in d.collect(Parameter):
                    if id(p) not in seen:
                        seen.add(id(p))
                        params.append(p)
        return params

    def sorted_parameters(self):
        return sorted(self.get_parameters(), key=lambda p: p.name)

    def sorted_inputs(self):
        return sorted(self._inputs, key=lambda f: f.name)

    def sorted_outputs(self):
        return sorted(self._outputs, key=lambda f: f.name)

    # -- construction helpers ------------------------------------------------
    def _estimate_env(self):
        env = {}
        for p, v in self._param_estimates_list:
            env[p] = int(v)
        for c in self._param_constraints:
            if isinstance(c, Condition) and c.conditional == '==' and \
                    isinstance(c.lhs, Parameter) and c.lhs not in env:
                try:
                    env[c.lhs] = ceval_int(c.rhs, env)
                except PlanError:
                    pass
        return env

    def _build_comps(self):
        funcs, parents, children = get_funcs(self._outputs)
        self._func_map = {}
        self._comps = []
        for f in funcs:
            if isinstance(f, Image):
                comp = ComputeObject(f)
                self._func_map[f] = comp
                continue
            comp = ComputeObject(f, f in self._outputs)
            self._func_map[f] = comp
            self._comps.append(comp)
        for f in funcs:
            comp = self._func_map[f]
            comp.parents = [self._func_map[p] for p in parents[f]
                            if not isinstance(p, Image)]
            comp.image_parents = [p for p in parents[f] if isinstance(p, Image)]
            comp.children = [self._func_map[c] for c in children[f]]
            if not comp.is_image and not comp.is_reduction:
                comp.parts = case_parts(f)
        order = level_order(self._comps, dict((c, c.parents) for c in self._comps))
        for c in self._comps:
            c.level = order[c]
        # stable: by level, then by discovery order reversed (producers first)
        idx = dict((c, i) for i, c in enumerate(self._comps))
        self._comps.sort(key=lambda c: (c.level, -idx[c]))

    def domain_box(self, func, env=None):
        """Concrete [(lo, hi)] of a function's domain."""
        env = self._env if env is None else env
        return [(ceval_int(d.lowerBound, env), ceval_int(d.upperBound, env))
                for d in func.domain]

    def comp_size(self, comp):
        n = 1
        for lo, hi in self.domain_box(comp.func):
            n *= max(0, hi - lo + 1)
        return n

    # -- bounds check (bounds.py:36-118) ---------------------------------------
    def bounds_check(self):
        env = self._env
        if not env:
            return
        for comp in self._comps:
            if comp.is_reduction:
                continue
            f = comp.func
            dom = self.domain_box(f)
            for part in comp.parts:
                box = []
                for d, (lo, hi) in enumerate(dom):
                    blo, bhi = part.box[d]
                    if blo is not None:
                        lo = max(lo, eval_bexpr(blo, env))
                    if bhi is not None:
                        hi = min(hi, eval_bexpr(bhi, env))
                    box.append((lo, hi))
                if any(lo > hi for lo, hi in box):
                    continue
                # references inside Select branches only run when their
                # branch is taken; the GPU evaluates Selects lazily, so only
                # unconditional references must stay inside the domain
                refs = unconditional_refs(part.expr)
                for c in part.pred:
                    refs += c.collect(Reference)
                for r in refs:
                    p = r.objectRef
                    pdom = self.domain_box(p)
                    for i, arg in enumerate(r.arguments):
                        acc = classify_access(arg, f.variables)
                        if acc.kind == 'const':
                            v = ceval_int(arg, env)
                            lo_i = hi_i = v
                        elif acc.kind == 'var':
                            lo_i, hi_i = interval_image(acc, box[acc.dim][0],
                                                        box[acc.dim][1], env)
                        else:
                            continue
                        if lo_i < pdom[i][0] or hi_i > pdom[i][1]:
                            raise TypeError("Reference out of domain",
                                            f.name, str(r), (lo_i, hi_i), pdom[i])

    # -- gather inlining ---------------------------------------------------------
    GATHER_OPS = 48
    GATHER_REFS = 128

    def gather_inline_pass(self):
        """Substitute a function into the consumers that read it at a
        data-dependent index.

        The local Laplacian blends, per pixel, two of K intensity levels of
        each Laplacian pyramid level (``lPyramid(k, x, y)`` with ``k`` computed
        from the pixel's gray value).  Materialising a level for all K values
        of k writes and reads K full planes, of which each pixel uses two.
        A function that is not a live-out, has a single unconditional
        definition of at most GATHER_OPS operations and is read by some
        consumer with a data-dependent ('*') argument is substituted into
        those references (cast to its own type, so values are unchanged);
        its other consumers keep reading
        it.  Repeats until nothing changes, so chains (the upsampled level
        read by the blend, the coarser level read by the upsampling)
        collapse into reads of the coarser stored level at the two needed k.
        A reader is not grown past GATHER_REFS references (gpu option
        ``gather_refs``): at 128 the local Laplacian's finest blend reads
        the level-1 Gaussian planes directly instead of materialising the
        upsampled level for all K values of k in the tile."""
        from .expr_ast import (AbstractBinaryOpNode, AbstractUnaryOpNode, Select,
                               InbuiltFunction, Cast)
        max_refs = int((self._gpu_options or {}).get('gather_refs', self.GATHER_REFS))
        for _ in range(64):
            done = True
            for comp in list(self._comps):
                P = comp.func
                if P in self._outputs or comp.is_reduction or comp.is_const:
                    continue
                body = P.defn
                if len(body) != 1 or isinstance(body[0], Case):
                    continue
                e = body[0]
                nops = 0
                stack = [e]
                while stack:
                    x = stack.pop()
                    if isinstance(x, (AbstractBinaryOpNode, AbstractUnaryOpNode, Cast, Select,
                                      InbuiltFunction)):
                        nops += 1
                    if hasattr(x, 'children'):
                        stack.extend(x.children())
                if nops > self.GATHER_OPS:
                    continue
                prefs = len(e.collect(Reference))
                # only functions that read their producers at their own
                # resolution or coarser (pointwise, shifted, upsampled, table
                # look-ups): recomputing a downsampling stage per use would
                # redo the finer level's work for every read
                fine = False
                for r in e.collect(Reference):
                    for a in r.arguments:
                        acc = classify_access(a, P.variables)
                        if acc.kind == 'var' and abs(acc.slope) > 1:
                            fine = True
                if fine:
                    continue
                # always cast: literals print as C doubles, so the C type of
                # the expression can be wider than the DSL type says
                sub = Cast(P.typ, e)
                hit = False
                for ch in comp.children:
                    cf = ch.func
                    if ch.is_reduction:
                        continue
                    cond_refs = set()
                    for case in cf.defn:
                        if isinstance(case, Case):
                            cond_refs |= set(id(r) for r in case.condition.collect(Reference))
                    m = {}
                    for r in cf.getObjects(Reference):
                        if r.objectRef is not P or id(r) in cond_refs:
                            continue
                        if any(classify_access(a, cf.variables).kind == '*' for a in r.arguments):
                            m[r] = substitute_vars(sub, dict(zip(P.variables, r.arguments)))
                    # bound the growth: each substituted reference brings
                    # P's references along (a pyramid would otherwise
                    # unfold level by level into its finest level)
                    total = sum(len(case.collect(Reference)) for case in cf.defn)
                    if m and total + len(m) * (prefs - 1) > max_refs:
                        m = {}
                    if m:
                        cf.replace_refs(m)
                        hit = True
                if hit:
                    self.gather_inlined = getattr(self, 'gather_inlined', []) + [P.name]
                    self._build_comps()
                    done = False
                    break
            if done:
                return

    # -- inline pass (inline.py:29-199) ----------------------------------------
    def inline_pass(self):
        for directive in self._inline_directives:
            assert isinstance(directive, Function)
            clone = self._clone_map[directive]
            assert clone not in self._outputs, "cannot inline a live-out"
            comp = self._func_map[clone]
            assert not comp.is_reduction
            parts = comp.parts
            if len(parts) != 1 or parts[0].pred:
                continue    # only single-expression functions (no split)
            expr = parts[0].expr
            ok_all = True
            for child in list(comp.children):
                cf = child.func
                if child.is_reduction:
                    ok_all = False
                    continue
                mapping = {}
                for cpart in child.parts:
                    for r in cpart.expr.collect(Reference):
                        if r.objectRef is clone:
                            vmap = dict(zip(clone.variables, r.arguments))
                            mapping[r] = substitute_vars(expr, vmap)
                if mapping:
                    cf.replace_refs(mapping)
            if ok_all:
                self._outputs = [o for o in self._outputs]

    # -- grouping ----------------------------------------------------------------
    def _link_groups(self):
        for g in self._groups:
            g.parents = []
            g.children = []
        for g in self._groups:
            for c in g.comps:
                for p in c.parents:
                    pg = p.group
                    if pg is not g and pg not in g.parents:
                        g.parents.append(pg)
                for ch in c.children:
                    cg = ch.group
                    if cg is not g and cg not in g.children:
                        g.children.append(cg)

    def merge_groups(self, g1, g2):
        comps = g1.comps + [c for c in g2.comps if c not in g1.comps]
        merged = Group(comps)
        for c in comps:
            c.group = merged
        pos = min(self._groups.index(g1), self._groups.index(g2))
        self._groups = [g for g in self._groups if g is not g1 and g is not g2]
        self._groups.append(merged)
        self._link_groups()
        return merged

    def _fusible(self, comps):
        return self._unfusible_reason(comps) is None

    def _unfusible_reason(self, comps):
        """GPU-side fusion rule on top of the reference's (grouping.py): an
        in-group reference whose argument is data dependent (or not an
        affine map of the reader's variables) must hit a producer dimension
        small enough to keep whole in the tile (<= 64 points).  Returns
        None when the functions can form one group, else the reason."""
        cset = set(comps)
        for c in comps:
            f = c.func
            if c.is_reduction:
                return '%s is a Reduction' % f.name
            for part in c.parts:
                refs = part.expr.collect(Reference)
                for cond in part.pred:
                    refs += cond.collect(Reference)
                for r in refs:
                    pc = self._func_map.get(r.objectRef)
                    if pc is None or pc not in cset:
                        continue
                    pdom = self.domain_box(r.objectRef)
                    for i, arg in enumerate(r.arguments):
                        acc = classify_access(arg, f.variables)
                        if acc.kind == '*':
                            lo, hi = pdom[i]
                            if hi - lo + 1 > 64:
                                return ('%s reads %s at a data-dependent index along a dimension '
                                        'of %d points (at most 64 fit a tile)'
                                        % (f.name, r.objectRef.name, hi - lo + 1))
        return None

    def auto_group(self):
        """Greedy level-by-level fusion (grouping.py:56-239)."""
        grp_size = self._group_size
        sizes = dict((c, self.comp_size(c)) for c in self._comps)

        small = []
        for c in self._comps:
            is_small = sizes[c] <= self._size_threshold
            for p in c.parents:
                is_small = is_small and sizes[p] > self._size_threshold
            if is_small and c.parents:
                small.append(c)

        changed = True
        while changed:
            changed = False
            for g in [g for g in self._groups if g.children]:
                is_small_grp = all(c in small for c in g.comps)
                red = any(c.is_reduction for c in g.comps) or \
                    any(c.is_reduction for ch in g.children for c in ch.comps)
                const = any(c.is_const for c in g.comps) or \
                    any(c.is_const for ch in g.children for c in ch.comps)
                if is_small_grp or red or const or len(g.comps) >= grp_size:
                    continue
                n = len(g.comps) + sum(len(ch.comps) for ch in g.children)
                if n > grp_size:
                    continue
                if len(g.children) > 1:
                    all_parents = []
                    for ch in g.children:
                        all_parents += [p for p in ch.parents if p is not g]
                    if not all(p in g.children for p in all_parents):
                        continue
                comps = list(g.comps)
                for ch in g.children:
                    comps += ch.comps
                if not self._fusible(comps):
                    continue
                merged = g
                for ch in list(g.children):
                    merged = self.merge_groups(merged, ch)
                changed = True
                break

    def manual_group(self):
        """The caller's grouping (``buildPipeline(grouping=...)``, lists of
        Functions): each list becomes one group, every other function a
        group of its own.  The reference accepts the argument but does not
        apply it (pipe.py:680-692, ``if self._grouping and False``) and
        always runs auto_group."""
        for funcs in self._grouping:
            comps = []
            for f in funcs:
                cl = self._clone_map.get(f)
                comp = self._func_map.get(cl) if cl is not None else None
                if comp is None or comp not in self._comps:
                    raise ValueError("grouping: %s is not a computed function of the pipeline"
                                     % getattr(f, 'name', f))
                comps.append(comp)
            if len(comps) > 1 and any(c.is_reduction or c.is_const for c in comps):
                raise ValueError("grouping: a Reduction or const function forms a group "
                                 "of its own (%s)" % [c.func.name for c in comps])
            if len(comps) > 1:
                why = self._unfusible_reason(comps)
                if why:
                    raise ValueError("grouping: %s cannot form one group: %s"
                                     % ([c.func.name for c in comps], why))
            g = comps[0].group
            for c in comps[1:]:
                if c.group is not g:
                    g = self.merge_groups(g, c.group)
        # the groups must run in some order: no group may feed another
        # that feeds it back
        indeg = dict((g, len(g.parents)) for g in self._groups)
        ready = [g for g in self._groups if indeg[g] == 0]
        seen = 0
        while ready:
            g = ready.pop()
            seen += 1
            for ch in g.children:
                indeg[ch] -= 1
                if indeg[ch] == 0:
                    ready.append(ch)
        if seen != len(self._groups):
            raise ValueError("grouping: the groups depend on each other in a cycle")

    def _order_groups(self):
        order = level_order(self._groups, dict((g, g.parents) for g in self._groups))
        idx = dict((g, i) for i, g in enumerate(self._groups))
        self._groups.sort(key=lambda g: (order[g], idx[g]))
        for g in self._groups:
            g.comps.sort(key=lambda c: c.level)

    # -- live-outs --------------------------------------------------------------
    def liveouts(self, group):
        """Comps of ``group`` read outside it or pipeline outputs."""
        out = []
        for c in group.comps:
            if c.func in self._outputs or any(ch.group is not group for ch in c.children):
                out.append(c)
        return out

    # -- code generation -----------------------------------------------------------
    def generate_code(self, is_extern_c_func=True, are_io_void_ptrs=True):
        from .hip_emit import emit_pipeline
        return emit_pipeline(self, is_extern_c_func, are_io_void_ptrs)

    @property
    def pipeline_graph(self):
        return self.draw_pipeline_graph()

    def draw_pipeline_graph(self):
        lines = ['digraph "%s" {' % self._name]
        for gi, g in enumerate(self._groups):
            lines.append('  subgraph cluster_%d { label="group %d";' % (gi, gi))
            for c in g.comps:
                style = ' style=bold' if c in self.liveouts(g) else ''
                lines.append('    "%s" [shape=box%s];' % (c.name, style))
            lines.append('  }')
        for img in self._inputs:
            lines.append('  "%s" [shape=ellipse];' % img.name)
        for c in self._comps:
            for p in c.parents:
                lines.append('  "%s" -> "%s";' % (p.name, c.name))
            for img in getattr(c, 'image_parents', []):
                lines.append('  "%s" -> "%s";' % (img.name, c.name))
        lines.append('}')
        return DotGraph('\n'.join(lines) + '\n')
