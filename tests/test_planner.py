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
    # 11 stages, height 4, slopes +-1/3 -> every stage on 40x264
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
    # exact halo: Ixx/Iyy/Ixy need tile + 1 on each side, the response the tile
    assert ext['Ixx'] == [34, 258] and ext['harris'] == [32, 256]
    # the products read global memory only, so they are swept in 4-wide
    # blocks and their LDS rows are padded to a multiple of 4 floats
    assert all(s.block[1] == 4 for s in plan.stages if s.name in ('Ixx', 'Iyy', 'Ixy'))
    assert plan.lds_bytes == 3 * 34 * 260 * 4
    # Ix/Iy and the 3x3 sums are computed in registers
    inl = sorted(s.name for s in plan.stages if s.inline)
    assert inl == ['Ix', 'Iy', 'Sxx', 'Sxy', 'Syy', 'det', 'trace']


def test_lds_reuse_across_phases():
    # without inlining every stage is materialised; buffers of one storage
    # class are reused once their last reader's phase has passed
    pd = {}
    out = harris_pipe(pd)
    pipe = build(out, pd, {'R': 510, 'C': 510}, 'harris', [16, 64], group_size=11,
                 gpu_options={'no_inline': True})
    plan = plan_pipeline(pipe)[0]
    offs = [s.lds_offset for s in plan.stages if s.needs_lds]
    assert len(offs) > len(set(offs))          # some offsets shared
    total = sum(4 * s.static_ext[0] * s.static_ext[1] for s in plan.stages if s.needs_lds)
    assert plan.lds_bytes < total
