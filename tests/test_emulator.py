"""The generated kernels and host stub, run on the CPU emulator
(tests/kernel_emu.py): per-thread blocks, partial blocks at region edges,
LDS layouts and the persistent work queue give the oracle's results bit for
bit, and an AddressSanitizer build finds no out-of-bounds access.  These
run without a GPU; the same cases run on the device in test_gpu_parity."""
import os
import subprocess
import sys

import numpy as np
import pytest

from oracle import cemit
from polymage_amd.apps.harris import harris_pipe
from polymage_amd.apps.unsharp import unsharp_mask_full
from polymage_amd.apps import inputs

from helpers import build, bits_equal
import kernel_emu

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.parametrize("R,C,tiles,block", [(37, 101, [16, 32], None), (62, 70, [32, 64], (2, 4)),
                                             (9, 13, [4, 8], (4, 4))])
def test_emulated_harris(R, C, tiles, block):
    pd = {}
    out = harris_pipe(pd)
    opts = {'block': block} if block else None
    pipe = build(out, pd, {'R': R, 'C': C}, 'harris', tiles, gpu_options=opts)
    img = np.random.default_rng(R).random((R + 2, C + 2)).astype(np.float32)
    g = np.full((R + 2, C + 2), -7.0, np.float32)
    o = g.copy()
    kernel_emu.build(pipe).run_host({'R': R, 'C': C, 'img': img, 'harris': g})
    cemit.build([out], 'harris_oracle_double')({'R': R, 'C': C, 'img': img, 'harris': o})
    assert bits_equal(g, o)


@pytest.mark.parametrize("block", [None, (8, 1), (1, 4), (4, 4)])
def test_emulated_unsharp_blocks(block):
    R, C = 45, 77
    pd = {}
    out = unsharp_mask_full(pd)
    pipe = build(out, pd, {'R': R, 'C': C}, 'unsharp', [16, 32], group_size=5,
                 gpu_options={'block': block} if block else None)
    img = inputs.smooth_field(np.random.default_rng(5), R + 4, C + 4, 3)
    vals = {'R': R, 'C': C, 'threshold': 0.001, 'weight': 3.0, 'img': img}
    g = np.full((3, R, C), -7.0, np.float32)
    o = g.copy()
    kernel_emu.build(pipe).run_host(dict(vals, mask=g))
    cemit.build([out], 'unsharp_oracle_unsharp_mask_full')(dict(vals, mask=o))
    assert bits_equal(g, o)


def test_emulated_kernels_under_asan():
    """One harris and one unsharp case of the device parity suite, run on
    the emulator built with AddressSanitizer (in a child process that
    preloads the ASan runtime)."""
    env = dict(os.environ, PMX_EMULATE='asan', ASAN_OPTIONS='detect_leaks=0',
               LD_PRELOAD=kernel_emu.asan_runtime())
    cmd = [sys.executable, '-m', 'pytest', '-x', '-q', '-p', 'no:cacheprovider',
           os.path.join(HERE, 'test_gpu_parity.py'),
           '-k', 'harris_small_bitexact and 62 or unsharp_full_small_bitexact and 97']
    r = subprocess.run(cmd, env=env, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       timeout=900)
    out = r.stdout.decode(errors='replace')
    assert r.returncode == 0, out[-3000:]
    assert '2 passed' in out, out[-3000:]
