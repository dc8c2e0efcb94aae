"""GPU parity: HIP path (through the generated C ABI) vs the oracle.

Float pipelines are compared bit for bit: both sides evaluate the same C
expression semantics with FP contraction off, so any difference is a bug.
"""
import numpy as np
import pytest

from oracle import cemit
from polymage_amd import runtime
from polymage_amd.apps.harris import harris_pipe
from polymage_amd.apps.unsharp import unsharp_mask, unsharp_mask_full
from polymage_amd.apps import inputs

from helpers import build, bits_equal, gpu_build

pytestmark = pytest.mark.gpu


def _harris(R, C, tiles, mode='double', seed=7, gpu_options=None):
    pd = {}
    out = harris_pipe(pd)
    pipe = build(out, pd, {'R': R, 'C': C}, 'harris', tiles, literal_mode=mode,
                 gpu_options=gpu_options)
    img = np.random.default_rng(seed).random((R + 2, C + 2)).astype(np.float32)
    g = np.full((R + 2, C + 2), -7.0, np.float32)
    o = np.full((R + 2, C + 2), -7.0, np.float32)
    orc = cemit.build([out], 'harris_oracle_' + mode, literal_mode=mode)
    gpu_build(pipe).run_host({'R': R, 'C': C, 'img': img, 'harris': g})
    orc({'R': R, 'C': C, 'img': img, 'harris': o})
    return g, o


@pytest.mark.parametrize("R,C,tiles", [(62, 70, [16, 32]), (253, 301, [32, 64]),
                                       (130, 517, [32, 256]), (5, 9, [4, 4])])
def test_harris_small_bitexact(gpu, R, C, tiles):
    print("case harris", R, C, tiles, flush=True)
    g, o = _harris(R, C, tiles)
    assert bits_equal(g, o)          # defined region and untouched border


def test_harris_float12_bitexact(gpu):
    g, o = _harris(200, 180, [32, 64], mode='float12')
    assert bits_equal(g, o)


def test_harris_full_size(gpu):
    print("case harris 2046", flush=True)
    g, o = _harris(2046, 2046, [32, 256], seed=1001)
    assert bits_equal(g, o)


def _unsharp(fn, R, C, tiles, seed=3):
    pd = {}
    out = fn(pd)
    pipe = build(out, pd, {'R': R, 'C': C}, 'unsharp', tiles, group_size=5)
    img = inputs.smooth_field(np.random.default_rng(seed), R + 4, C + 4, 3)
    g = np.full((3, R, C), -7.0, np.float32)
    o = np.full((3, R, C), -7.0, np.float32)
    vals = {'R': R, 'C': C, 'threshold': 0.001, 'weight': 3.0, 'img': img}
    orc = cemit.build([out], 'unsharp_oracle_%s' % fn.__name__)
    gpu_build(pipe).run_host(dict(vals, mask=g))
    orc(dict(vals, mask=o))
    return g, o


@pytest.mark.parametrize("R,C,tiles", [(40, 300, [8, 64]), (97, 130, [32, 256])])
def test_unsharp_full_small_bitexact(gpu, R, C, tiles):
    print("case unsharp", R, C, tiles, flush=True)
    g, o = _unsharp(unsharp_mask_full, R, C, tiles)
    assert bits_equal(g, o)


def test_unsharp_app_defined_region(gpu):
    # the reference app leaves y in {2, C-1, C} undefined (its Case quirk);
    # y index 0 is y = 2, so the defined columns are indices 1 .. C-4
    g, o = _unsharp(unsharp_mask, 64, 200, [16, 64])
    assert bits_equal(g[:, :, 1:-3], o[:, :, 1:-3])


def test_unsharp_full_size(gpu):
    print("case unsharp 2844x4252", flush=True)
    g, o = _unsharp(unsharp_mask_full, 2844, 4252, [32, 256], seed=1002)
    assert bits_equal(g, o)


def _bilateral(rows, cols, tiles, seed=5, app=False):
    from polymage_amd.apps.bilateral import bilateral_grid, bilateral_grid_full
    pd = {}
    out = (bilateral_grid if app else bilateral_grid_full)(pd)
    pipe = build(out, pd, {'rows': rows, 'cols': cols}, 'bilateral', tiles, group_size=5)
    img = inputs.smooth_field(np.random.default_rng(seed), rows, cols, 1)[0]
    g = np.full((1, rows, cols), -7.0, np.float32)
    o = np.full((1, rows, cols), -7.0, np.float32)
    orc = cemit.build([out], 'bilateral_oracle_%s' % ('app' if app else 'full'))
    gpu_build(pipe).run_host({'rows': rows, 'cols': cols, 'input': img, 'filtered': g})
    orc({'rows': rows, 'cols': cols, 'input': img, 'filtered': o})
    return g, o


@pytest.mark.parametrize("rows,cols,tiles", [(64, 48, [16, 16]), (256, 192, [32, 64]),
                                             (100, 77, [16, 32])])
def test_bilateral_bitexact(gpu, rows, cols, tiles):
    print("case bilateral", rows, cols, tiles, flush=True)
    g, o = _bilateral(rows, cols, tiles)
    assert np.all(np.isfinite(o))
    assert bits_equal(g, o)


def test_bilateral_full_size(gpu):
    print("case bilateral 2560x1536", flush=True)
    g, o = _bilateral(2560, 1536, [32, 64], seed=1003)
    assert bits_equal(g, o)
