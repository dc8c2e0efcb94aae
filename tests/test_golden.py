"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py
from the reference's own generated C++): the oracle and the HIP path
against the reference's outputs, bit for bit, with no reference tree
needed at test time.

  refharris_RxC  ipp_bench/harris/harris_function.cpp refHarris (32x256 tiles)
  vd_harris      video_demo/harris_naive.cpp (= harris_polymage.cpp)
  vd_unsharp     video_demo/unsharp_naive.cpp (= unsharp_polymage.cpp)
  vd_laplacian   video_demo/laplacian_naive.cpp, rows its bottom padding
                 cannot reach"""
import os

import numpy as np
import pytest

from oracle import cemit
from polymage_amd.apps.harris import harris_pipe
from polymage_amd.apps.local_laplacian import local_laplacian

import vd_pipelines
from helpers import bits_equal, build, gpu_build

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden')
REFHARRIS = ['refharris_64x96', 'refharris_37x133', 'refharris_2x7']


def _load(name):
    with np.load(os.path.join(GOLDEN, name + '.npz'), allow_pickle=False) as z:
        return dict((k, z[k]) for k in z.files)


def _cases():
    """(fixture, DSL output builder, values, output name, compare) per fixture."""
    cases = []
    for name in REFHARRIS:
        f = _load(name)
        R, C = int(f['R']), int(f['C'])
        cases.append((name, harris_pipe, {'R': R, 'C': C, 'img': f['img']}, 'harris',
                      f['harris'], (slice(2, R), slice(2, C))))
    f = _load('vd_harris')
    R, C = int(f['R']), int(f['C'])
    cases.append(('vd_harris', vd_pipelines.vd_harris, {'R': R, 'C': C, 'img': f['img']}, 'harris',
                  f['harris'], (slice(2, R), slice(2, C))))
    f = _load('vd_unsharp')
    cases.append(('vd_unsharp', vd_pipelines.vd_unsharp,
                  {'R': int(f['R']), 'C': int(f['C']), 'threshold': float(f['threshold']),
                   'weight': float(f['weight']), 'img': f['img']}, 'mask', f['mask'], (slice(None),)))
    f = _load('vd_laplacian')
    cases.append(('vd_laplacian', lambda pd: local_laplacian(pd, K=4, J=5, io='u8'),
                  {'R': int(f['R']), 'C': int(f['C']), 'alpha': float(f['alpha']),
                   'beta': float(f['beta']), 'img_colour': f['img']}, 'laplacian', f['laplacian'],
                  (slice(0, int(f['valid_rows'])),)))
    return cases


CASES = _cases()
IDS = [c[0] for c in CASES]


def test_fixtures_present():
    assert len(CASES) == 6
    assert all(c[4].size > 0 for c in CASES)


@pytest.mark.parametrize("case", CASES, ids=IDS)
def test_oracle_vs_golden(case):
    name, fn, vals, oname, expect, sl = case
    out = fn({})
    prog = cemit.build([out], 'golden_oracle_' + name, literal_mode='float12')
    o = np.zeros_like(expect)
    prog(dict(vals, **{oname: o}))
    assert bits_equal(o[sl], expect[sl])


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES, ids=IDS)
def test_hip_vs_golden(gpu, case):
    """The HIP kernels through the C ABI (pipeline_<name>, host pointers)
    against the reference's outputs."""
    name, fn, vals, oname, expect, sl = case
    print("case golden", name, flush=True)
    pd = {}
    out = fn(pd)
    sizes = dict((k, vals[k]) for k in ('R', 'C'))
    gs = 5 if name == 'vd_laplacian' else 11
    pipe = build(out, pd, sizes, 'g_' + name, [16, 32], group_size=gs, literal_mode='float12')
    g = np.zeros_like(expect)
    gpu_build(pipe).run_host(dict(vals, **{oname: g}))
    assert bits_equal(g[sl], expect[sl])
