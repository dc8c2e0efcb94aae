"""The C++/OpenMP backend (polymage_amd/cpu_emit.py): every pipeline of the
GPU parity suite, built for the host from the same group plans, equals
the oracle bit for bit (untouched output points included), and so do the
reductions.  These run in the CPU suite."""
import numpy as np
import pytest

from polymage_amd import runtime
from polymage_amd.apps.unsharp import unsharp_mask_full

import test_gpu_parity as P
import test_reductions as RED
from helpers import bits_equal


@pytest.fixture
def cpu(monkeypatch):
    monkeypatch.setattr(P, 'gpu_build', runtime.build_cpu)
    monkeypatch.setenv('OMP_NUM_THREADS', '4')


@pytest.mark.parametrize("R,C,tiles", [(62, 70, [16, 32]), (130, 517, [32, 256]), (5, 9, [4, 4])])
def test_cpu_harris(cpu, R, C, tiles):
    g, o = P._harris(R, C, tiles)
    assert bits_equal(g, o)


def test_cpu_harris_float12(cpu):
    g, o = P._harris(200, 180, [32, 64], mode='float12')
    assert bits_equal(g, o)


@pytest.mark.parametrize("R,C,tiles", [(40, 300, [8, 64]), (97, 130, [32, 256])])
def test_cpu_unsharp(cpu, R, C, tiles):
    g, o = P._unsharp(unsharp_mask_full, R, C, tiles)
    assert bits_equal(g, o)


@pytest.mark.parametrize("rows,cols,tiles", [(64, 48, [16, 16]), (100, 77, [16, 32])])
def test_cpu_bilateral(cpu, rows, cols, tiles):
    g, o = P._bilateral(rows, cols, tiles)
    assert bits_equal(g, o)


def test_cpu_laplacian_u16(cpu):
    g, o = P._laplacian(96, 80, [16, 32])
    assert np.array_equal(g, o)


def test_cpu_laplacian_video_demo_form(cpu):
    g, o = P._laplacian(160, 128, [32, 64], K=4, J=5, io='u8', mode='float12')
    assert np.array_equal(g, o)


def test_cpu_interpolate(cpu):
    g, o = P._interpolate(70, 90, [16, 32], 5)
    assert bits_equal(g, o)


def test_cpu_blend(cpu):
    g, o = P._blend(96, 130, [16, 32], 5)
    assert bits_equal(g, o)


@pytest.mark.parametrize("variant,gs", [('reference', 5), ('batch', 40)])
def test_cpu_campipe(cpu, variant, gs):
    from oracle import cemit
    from polymage_amd.apps.campipe import camera_pipe
    pd = {}
    out = camera_pipe(pd, variant=variant)
    img = P._campipe_frames(1, 21, variant, sizes=[(288, 416)])[0]
    R, C = img.shape
    pipe = P.build(out, pd, {'R': R, 'C': C}, 'campipe', [16, 64], group_size=gs)
    vals = P._campipe_vals(img)
    g = np.zeros((3, R - 48, C - 64), np.uint8)
    o = np.zeros_like(g)
    runtime.build_cpu(pipe).run_host(dict(vals, process=g))
    cemit.build([out], 'cam_oracle_' + variant)(dict(vals, process=o))
    assert np.array_equal(g, o)


@pytest.mark.parametrize("case", sorted(RED.CASES))
def test_cpu_reductions(cpu, case):
    g, o = RED._run(case, 23, 37, 3, runtime.build_cpu)
    assert bits_equal(g, o)
