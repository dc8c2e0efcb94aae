"""The reference's overlapped-tile geometry, for comparison and reporting.

The reference tiles a fused group with a trapezoid per tiled dimension:

* dependence vectors ``[level(consumer) - level(producer), d_1..d_n]`` with
  ``d_i`` the negated constant offset of a uniform access along aligned
  dimension i, '*' when the access is not uniform (poly.py:291-429);
* ``compute_tile_slope`` (poly_schedule.py:128-219): from the top level
  down, accumulate the widest positive and negative reach per level; the
  slope of a dimension is reach/(hmax - h) maximised over levels;
* ``overlap_tile`` (poly_schedule.py:429-536): ``left = ceil(max slope)``,
  ``right = floor(min slope)``, ``overlap = |left*h| + |right*h|`` with h
  the group height; every stage of the group is computed on
  ``ts + overlap`` points per tile and dimension, and
  ``enable_tile_scratchpad`` (305-346) sizes each scratchpad
  ``ceil((ts + overlap)/scale)``.

The GPU backend does not use this geometry (``group_plan`` propagates the
exact region of each stage, which is smaller and never reads uncomputed
producer points); it is kept to report how much redundant work and
scratch the two schemes imply, and it is pinned in the tests by the
scratchpad sizes of the reference's checked-in generated code.
"""

import math
from fractions import Fraction

from .analysis import classify_access
from .constructs import Reference


def dependence_vectors(pipe, group, anchor_dims=None):
    """[(vec, parent_level)] for every in-group reference; vec[0] is the
    level difference, vec[1:] per anchor dimension (d = -offset)."""
    comps = group.comps
    inset = set(comps)
    anchor = max(comps, key=lambda c: (c.level, pipe.comp_size(c)))
    nd = len(anchor.func.variables)
    out = []
    for c in comps:
        if c.parts is None:
            continue
        for part in c.parts:
            refs = part.expr.collect(Reference)
            for cond in part.pred:
                refs += cond.collect(Reference)
            for r in refs:
                pc = pipe.func_map.get(r.objectRef)
                if pc is None or pc not in inset:
                    continue
                vec = [c.level - pc.level] + [0] * nd
                for i, arg in enumerate(r.arguments):
                    acc = classify_access(arg, c.func.variables)
                    # align producer dim i with the consumer dim it reads,
                    # and consumer dims with anchor dims by variable name
                    if acc.kind == 'var':
                        name = acc.var.name
                        j = [k for k, v in enumerate(anchor.func.variables) if v.name == name]
                        if not j:
                            continue
                        j = j[0]
                        if acc.slope != 1:
                            vec[1 + j] = '*'
                            continue
                        from .analysis import ceval_int
                        off = ceval_int(arg, {acc.var: 0})
                        if vec[1 + j] != '*':
                            vec[1 + j] = -off
                    elif acc.kind == '*':
                        pass
                out.append((vec, pc.level))
    return out


def tile_slopes(dep_vecs, hmax):
    if not dep_vecs:
        return [], []
    n = len(dep_vecs[0][0]) - 1
    slope_min = [Fraction(0)] * n
    slope_max = [Fraction(0)] * n
    unknown = [False] * n
    hmin = min(h for _, h in dep_vecs)
    min_w = [0] * n
    max_w = [0] * n
    widths = []
    for h in range(hmax - 1, hmin - 1, -1):
        lmax = [0] * n
        lmin = [0] * n
        for vec, ph in dep_vecs:
            if ph != h:
                continue
            for i in range(n):
                d = vec[i + 1]
                if d == '*':
                    unknown[i] = True
                    continue
                if d > 0:
                    lmax[i] = max(lmax[i], d)
                if d < 0:
                    lmin[i] = min(lmin[i], d)
        for i in range(n):
            max_w[i] += lmax[i]
            min_w[i] += lmin[i]
        widths.append((list(min_w), h))
        widths.append((list(max_w), h))
    for w, h in widths:
        sc = hmax - h
        for i in range(n):
            f = Fraction(w[i], sc)
            if w[i] < 0 and f < slope_min[i]:
                slope_min[i] = f
            if w[i] > 0 and f > slope_max[i]:
                slope_max[i] = f
    return (['*' if unknown[i] else slope_min[i] for i in range(n)],
            ['*' if unknown[i] else slope_max[i] for i in range(n)])


def reference_geometry(pipe, group, tile_sizes=None):
    """Per tiled anchor dimension: tile size, left/right slopes, overlap
    and the scratchpad extent every stage of the group gets."""
    tile_sizes = list(tile_sizes if tile_sizes is not None else pipe.tile_sizes)
    comps = group.comps
    anchor = max(comps, key=lambda c: (c.level, pipe.comp_size(c)))
    levels = [c.level for c in comps]
    hmax, hmin = max(levels), min(levels)
    h = hmax - hmin
    deps = dependence_vectors(pipe, group)
    smin, smax = tile_slopes(deps, hmax)
    dom = pipe.domain_box(anchor.func)
    dims = []
    k = 0
    for d in range(len(dom)):
        ext = dom[d][1] - dom[d][0] + 1
        if k < len(tile_sizes) and ext > tile_sizes[k] and (not smin or smin[d] != '*'):
            ts = tile_sizes[k]
            k += 1
            lo = smin[d] if smin else Fraction(0)
            hi = smax[d] if smax else Fraction(0)
            right = int(math.floor(lo))
            left = int(math.ceil(hi))
            overlap = abs(left * h) + abs(right * h)
            dims.append({'dim': d, 'tile': ts, 'left': left, 'right': right,
                         'overlap': overlap, 'scratch': ts + overlap})
        else:
            dims.append({'dim': d, 'tile': None, 'scratch': ext})
    return {'h': h, 'dims': dims, 'dep_vecs': deps}
