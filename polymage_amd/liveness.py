"""Liveness of compute objects over a schedule.

Same contract as sandbox/liveness.py:35-54: given the children map of a
DAG and a schedule (object -> time stamp), return a map from time stamp to
the objects whose last use happens at that time.  Objects with no children
in the map are not listed; objects in the map with no children get -1.
"""


def compute_liveness(children_map, schedule):
    liveness_map = {}
    for comp in schedule:
        if comp not in children_map:
            continue
        last = -1
        for child in children_map[comp]:
            last = max(last, schedule[child])
        liveness_map.setdefault(last, []).append(comp)
    return liveness_map


def level_order(objs, parent_map):
    """Longest-path level numbering of a DAG (utils.py:120-138):
    level(o) = 1 + max(level(parent)), roots at 0."""
    order = dict((o, 0) for o in objs)
    changed = True
    while changed:
        changed = False
        for o in objs:
            for p in (parent_map.get(o) or []):
                if p in order and order[p] >= order[o]:
                    order[o] = order[p] + 1
                    changed = True
    return order
