"""Planner: grouping, the reference's trapezoid geometry (pinned by the
scratchpad sizes of the reference's generated C++) and the exact regions
the GPU plan uses instead."""
import pytest

from polymage_amd.apps.harris import harris_pipe
from polymage_amd.apps.unsharp import unsharp_mask
from polymage_amd.ref_geometry import reference_geometry
from polymage_amd.hip_emit import plan_pipeline

import vd_pipelines
from helpers import build


def _harris(tiles, gs=11):
    pd = {}
    out = harris_pipe(pd)
    return build(out, pd, {'R': 2046, 'C': 2046}, 'harris', tiles, group_size=gs)


def test_grouping_harris():
    assert len(_harris([32, 256]).groups) == 1            # g_size 11 -> one group
    p5 = _harris([32, 256], gs=5)
    assert len(p5.groups) > 1 and max(len(g.comps) for g in p5.groups) <= 5


def test_reference_trapezoid_harris():
    # 11 stages, height 4, slopes +-1/3 -> every stage on 40x264 (poly_schedule.py:429-536)
    pipe = _harris([32, 256])
    geo = reference_geometry(pipe, pipe.groups[0])
    assert geo['h'] == 4
    assert [d['scratch'] for d in geo['dims']] == [40, 264]


def test_reference_trapezoid_video_demo_harris():
    # video_demo/harris_polymage.cpp:19-29 declares [26][138] at 16x128
    pd = {}
    out = vd_pipelines.vd_harris(pd)
    pipe = build(out, pd, {'R': 1000, 'C': 1000}, 'vdh', [16, 128], group_size=12,
                 literal_mode='float12')
    big = max(pipe.groups, key=lambda g: len(g.comps))
    geo = reference_geometry(pipe, big)
    assert geo['h'] == 5
    assert [d['scratch'] for d in geo['dims']] == [26, 138]


def test_reference_trapezoid_unsharp():
    # video_demo/unsharp_polymage.cpp:27-29 declares [3][32][262] at 32x256
    pd = {}
    out = unsharp_mask(pd)
    pipe = build(out, pd, {'R': 2844, 'C': 4252}, 'u', [32, 256], group_size=5)
    geo = reference_geometry(pipe, pipe.groups[0])
    assert [d['scratch'] for d in geo['dims']] == [3, 32, 262]


def test_exact_regions_are_tighter():
    pipe = _harris([32, 256])
    plan = plan_pipeline(pipe)[0]
    ext = dict((s.name, s.static_ext) for s in plan.stages)
    # the products are recomputed at each tap of the 3x3 sums, so Ix and Iy
    # are the stored stages, on tile + 1 each side: 34 x 258, the scratch
    # size of the reference's generated Harris at 32x256 tiles
    # (ipp_bench/harris/harris_function.cpp:17-21, 8772 = 34 * 258 floats)
    assert plan.offset_inlined == ['Ixx', 'Iyy', 'Ixy']
    assert ext['Ix'] == [34, 258] and ext['Iy'] == [34, 258] and ext['harris'] == [32, 256]
    # they read global memory only, so they are swept in 4-wide blocks and
    # their LDS rows are padded to a multiple of 4 floats; the sums read
    # them in 4-wide blocks of 128-bit row loads, which may run up to three
    # floats past the last row (16 bytes of slack per array)
    assert all(s.block[1] == 4 for s in plan.stages if s.name in ('Ix', 'Iy'))
    assert all(s.vread for s in plan.stages if s.name in ('Ix', 'Iy'))
    assert plan.lds_bytes == 2 * (34 * 260 * 4 + 16)
    # the 3x3 sums, det and trace are computed in registers
    inl = sorted(s.name for s in plan.stages if s.inline)
    assert inl == ['Sxx', 'Sxy', 'Syy', 'det', 'trace']
    # the reference trapezoid computes every stage on 40 x 264 points
    geo = reference_geometry(pipe, pipe.groups[0])
    assert [d['scratch'] for d in geo['dims']] == [40, 264]


def test_offset_inlining_needs_taps_inside_the_case():
    # a pointwise stage whose consumer reads it outside its Case box keeps
    # its storage (the program reads undefined points there)
    from polymage_amd.constructs import Variable, Parameter, Interval, Image, Function, Case, Condition
    from polymage_amd.expr_types import Int, Float

    def plan_for(lo, hi_off):
        R = Parameter(Int, 'R')
        x = Variable(Int, 'x')
        img = Image(Float, 'img', [R + 2])
        dom = Interval(Int, 0, R + 1)
        a = Function(([x], [dom]), Float, 'a')
        a.defn = [img(x) * 2.0]
        b = Function(([x], [dom]), Float, 'b')
        b.defn = [Case(Condition(x, '>=', 1) & Condition(x, '<=', R), a(x) * a(x))]
        c = Function(([x], [dom]), Float, 'c')
        c.defn = [Case(Condition(x, '>=', lo) & Condition(x, '<=', R - hi_off), b(x - 1) + b(x + 1))]
        pipe = build(c, {'R': R}, {'R': 1000}, 'oi', [64], group_size=5)
        return plan_pipeline(pipe)[0]

    plan = plan_for(1, 0)                      # c reads b(0) and b(R+1)
    assert 'b' in [s.name for s in plan.stages]
    plan = plan_for(2, 1)                      # every tap inside b's Case
    # b goes, then a (now read with offsets by c, unconditionally defined)
    assert getattr(plan, 'offset_inlined', []) == ['b', 'a']
    assert [s.name for s in plan.stages] == ['c']


def test_magic24_division_is_exact_over_the_sweep_range():
    """The sweep's item decomposition replaces x / d by a 24-bit multiply
    and a shift only when that is exact for every dividend the sweep
    produces; brute force over the whole range, including its last value."""
    import numpy as np
    from polymage_amd.hip_emit import _magic24
    for n, d in [(1122, 66), (520, 260), (4096, 162), (65536, 1188), (33 * 19, 19), (7, 3)]:
        m, s = _magic24(n, d)
        x = np.arange(n, dtype=np.uint64)
        assert m < (1 << 24) and (n - 1) * m < (1 << 32)
        assert np.array_equal((x * np.uint64(m)) >> np.uint64(s), x // np.uint64(d))
    assert _magic24(100000, 7) is None          # no 24-bit multiplier reaches that range
    assert _magic24(10, 1) is None


def test_halo_aligned_tiles():
    """unsharp's vertical blur adds 2 columns on each side of a tile, Harris'
    Sobel 1: the aligned tiles make the widest stage a power of two wide."""
    from polymage_amd import tuner
    assert tuner.row_halo('unsharp') == 4 and tuner.row_halo('harris') == 2
    assert (16, 252) in tuner.aligned_tiles('unsharp')
    assert (16, 510) in tuner.aligned_tiles('harris')
    assert tuner.aligned_tiles('bilateral') == []


def test_manual_grouping():
    """buildPipeline(grouping=...) forms the listed groups (every other
    function alone); a Reduction cannot join a group, and groups that feed
    each other in a cycle are an error."""
    from polymage_amd.apps.bilateral import bilateral_grid_full
    from polymage_amd.compiler import buildPipeline
    from polymage_amd.constructs import Condition
    from polymage_amd.pipe import get_funcs

    def pipe(names):
        pd = {}
        out = bilateral_grid_full(pd)
        f = dict((x.name, x) for x in get_funcs([out])[0])
        return buildPipeline([out], param_estimates=[(pd['rows'], 64), (pd['cols'], 48)],
                             param_constraints=[Condition(pd['rows'], '==', 64),
                                                Condition(pd['cols'], '==', 48)],
                             grouping=[[f[n] for n in g] for g in names], tile_sizes=[16, 16])

    p = pipe([['blurz', 'blurx'], ['interpolated', 'filtered']])
    assert [[c.func.name for c in g.comps] for g in p._groups] == \
        [['grid'], ['blurz', 'blurx'], ['blury'], ['interpolated', 'filtered']]
    with pytest.raises(ValueError, match="Reduction"):
        pipe([['grid', 'blurz']])
    with pytest.raises(ValueError, match="cycle"):
        pipe([['blurz', 'blury']])


def test_manual_grouping_rejects_unfusible_groups_with_the_reason():
    """A caller's group in which one function reads another at a
    data-dependent index along more than 64 points cannot be one kernel:
    buildPipeline says so, instead of failing later at code generation.
    Functions of one group asking for different tiles are refused too."""
    from polymage_amd.compiler import buildPipeline
    from polymage_amd.constructs import (Parameter, Variable, Interval, Image, Function, Cast,
                                         Case, Condition, Min, Max, Int, Float)
    N = Parameter(Int, "N")
    x = Variable(Int, "x")
    img = Image(Float, "img", [N])
    inner = Interval(Int, 1, N - 2)
    f = Function(([x], [inner]), Float, "f")
    # (a Case keeps f from being substituted into its data-dependent reader)
    f.defn = [Case(Condition(x, '>=', 1), img(x - 1) + img(x) + img(x + 1))]
    g = Function(([x], [inner]), Float, "g")
    g.defn = [f(Min(Max(Cast(Int, img(x) * 100.0), 1), N - 2)) + 1.0]
    with pytest.raises(ValueError, match="data-dependent"):
        buildPipeline([g], param_estimates=[(N, 1000)], grouping=[[f, g]], tile_sizes=[64])
    h = Function(([x], [Interval(Int, 1, N - 3)]), Float, "h")
    h.defn = [f(x) + f(x + 1)]
    p = buildPipeline([h], param_estimates=[(N, 1000)], tile_sizes=[64], group_size=5)
    assert [[c.func.name for c in grp.comps] for grp in p._groups] == [['f', 'h']]
    q = buildPipeline([h], param_estimates=[(N, 1000)], tile_sizes=[64], group_size=5,
                      gpu_options={'group_tiles': {'f': [32], 'h': [64]}})
    with pytest.raises(ValueError, match="group_tiles"):
        q.generate_code()
