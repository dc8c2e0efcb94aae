"""Liveness known answers (the reference's DAGs and schedules,
sandbox/tests/test_liveness_computation.py:52-200), checked exactly."""
from polymage_amd.liveness import compute_liveness, level_order


class Node(object):
    def __init__(self, name, parents=None):
        self.name = name
        self.parents = parents


def children_map(nodes):
    cm = {}
    for n in nodes:
        for p in (n.parents or []):
            cm.setdefault(p, []).append(n)
    return cm


def names(m):
    return dict((k, sorted(n.name for n in v)) for k, v in m.items())


def test_chain():
    a = Node('a'); b = Node('b', [a]); c = Node('c', [b]); d = Node('d', [c])
    e = Node('e', [d]); f = Node('f', [e])
    nodes = [a, b, c, d, e, f]
    sched = dict((n, i) for i, n in enumerate(nodes))
    assert names(compute_liveness(children_map(nodes), sched)) == \
        {1: ['a'], 2: ['b'], 3: ['c'], 4: ['d'], 5: ['e']}


def test_chain_multi_children():
    a = Node('a'); b = Node('b', [a]); c = Node('c', [a, b]); d = Node('d', [b, c])
    e = Node('e', [c, d]); f = Node('f', [d, e])
    nodes = [a, b, c, d, e, f]
    sched = dict((n, i) for i, n in enumerate(nodes))
    assert names(compute_liveness(children_map(nodes), sched)) == \
        {2: ['a'], 3: ['b'], 4: ['c'], 5: ['d', 'e']}


def test_tree():
    a = Node('a'); b = Node('b', [a]); c = Node('c', [a]); d = Node('d', [b])
    e = Node('e', [b]); f = Node('f', [c]); g = Node('g', [c])
    nodes = [a, b, c, d, e, f, g]
    sched = dict((n, i) for i, n in enumerate(nodes))
    assert names(compute_liveness(children_map(nodes), sched)) == \
        {2: ['a'], 4: ['b'], 6: ['c']}


def _harris_nodes():
    img = Node('img'); Ix = Node('Ix', [img]); Iy = Node('Iy', [img])
    Ixx = Node('Ixx', [Ix]); Ixy = Node('Ixy', [Ix, Iy]); Iyy = Node('Iyy', [Iy])
    Sxx = Node('Sxx', [Ixx]); Sxy = Node('Sxy', [Ixy]); Syy = Node('Syy', [Iyy])
    det = Node('det', [Sxx, Sxy, Syy]); trace = Node('trace', [Sxx, Syy])
    harris = Node('harris', [det, trace])
    return dict((n.name, n) for n in [img, Ix, Iy, Ixx, Ixy, Iyy, Sxx, Sxy, Syy, det, trace, harris])


def test_harris_schedules():
    N = _harris_nodes()
    cm = children_map(N.values())
    s1 = ['img', 'Ix', 'Iy', 'Ixx', 'Ixy', 'Iyy', 'Sxx', 'Sxy', 'Syy', 'det', 'trace', 'harris']
    sched = dict((N[k], i) for i, k in enumerate(s1))
    assert names(compute_liveness(cm, sched)) == {
        2: ['img'], 4: ['Ix'], 5: ['Iy'], 6: ['Ixx'], 7: ['Ixy'], 8: ['Iyy'],
        9: ['Sxy'], 10: ['Sxx', 'Syy'], 11: ['det', 'trace']}
    s2 = ['img', 'Iy', 'Iyy', 'Syy', 'Ix', 'Ixy', 'Sxy', 'Ixx', 'Sxx', 'trace', 'det', 'harris']
    sched = dict((N[k], i) for i, k in enumerate(s2))
    assert names(compute_liveness(cm, sched)) == {
        3: ['Iyy'], 4: ['img'], 5: ['Iy'], 6: ['Ixy'], 7: ['Ix'], 8: ['Ixx'],
        10: ['Sxx', 'Sxy', 'Syy'], 11: ['det', 'trace']}


def test_level_order():
    N = _harris_nodes()
    pm = dict((n, n.parents or []) for n in N.values())
    lv = level_order(list(N.values()), pm)
    assert lv[N['img']] == 0 and lv[N['Ixy']] == 2 and lv[N['det']] == 4 and lv[N['harris']] == 5
