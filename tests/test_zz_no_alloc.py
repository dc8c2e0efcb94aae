"""The allocation-free entry point pipeline_<name>_dev_noalloc.

It takes the arguments of pipeline_<name>_dev and keeps every buffer the
library provides (status word, constant tables, intermediates between
kernels) in the runtime's static device scratch instead of the device
arena.  Its output must equal the plain entry point's, bit for bit, on two
calls in a row (the second with a smaller input), the device arena must not
grow across its calls, and sizes beyond the static scratch must be refused
with an error before anything is launched.
"""
import ctypes

import numpy as np
import pytest

from oracle import cemit
from polymage_amd.apps import inputs
from polymage_amd.utils import convert_to_ctype

from helpers import build, bits_equal, gpu_build, emulating

pytestmark = pytest.mark.gpu


def _bilateral_case():
    from polymage_amd.apps.bilateral import bilateral_grid_full
    pd = {}
    out = bilateral_grid_full(pd)
    pipe = build(out, pd, {'rows': 256, 'cols': 192}, 'bilateral', [32, 64], group_size=5)

    def values(rows, cols, seed):
        img = inputs.smooth_field(np.random.default_rng(seed), rows, cols, 1)[0]
        return {'rows': rows, 'cols': cols, 'input': img}, (1, rows, cols)
    return out, pipe, values, [(256, 192), (100, 77)]


def _laplacian_case():
    from polymage_amd.apps.local_laplacian import local_laplacian
    pd = {}
    out = local_laplacian(pd, K=8, J=4, io='u16', lut_hi_inclusive=True)
    lo, hi = pd['pad']
    pipe = build(out, pd, {'R': 192 + lo + hi, 'C': 160 + lo + hi}, 'laplacian', [32, 64], group_size=5)

    def values(rows, cols, seed):
        f = inputs.smooth_field(np.random.default_rng(seed), rows, cols, 3)
        img = np.ascontiguousarray(np.round(f * 65535.0).astype(np.uint16).transpose(1, 2, 0))
        return {'R': rows + lo + hi, 'C': cols + lo + hi, 'alpha': 1.0 / 7, 'beta': 1.0,
                'img_colour': img}, (rows, cols, 3)
    return out, pipe, values, [(192, 160), (96, 72)]


def _campipe_case():
    from polymage_amd.apps.campipe import camera_pipe, MATRIX_3200, MATRIX_7000
    pd = {}
    out = camera_pipe(pd, variant='batch')
    pipe = build(out, pd, {'R': 416, 'C': 608}, 'campipe', [16, 64], group_size=40)

    def values(R, C, seed):
        img = np.random.default_rng(seed).integers(0, 1024, (R, C)).astype(np.uint16)
        return {'R': R, 'C': C, 'colour_temp': 3700.0, 'gamma': 2.2, 'contrast': 50.0, 'img': img,
                'matrix_3200': np.array(MATRIX_3200, np.float32),
                'matrix_7000': np.array(MATRIX_7000, np.float32)}, (3, R - 48, C - 64)
    return out, pipe, values, [(416, 608), (256, 320)]


CASES = {'bilateral': _bilateral_case, 'laplacian': _laplacian_case, 'campipe': _campipe_case}


class _Dev(object):
    """Device buffers through torch on a GPU; host arrays on the kernel
    emulator (its device pointers are host pointers)."""

    def __init__(self):
        self.emu = emulating()
        if not self.emu:
            import torch
            self.torch = torch
            self.dev = torch.device('cuda:0')

    def put(self, a):
        if self.emu:
            return np.ascontiguousarray(a)
        raw = a.view(np.int16) if a.dtype == np.uint16 else a
        return self.torch.from_numpy(np.ascontiguousarray(raw)).to(self.dev)

    def ptr(self, t):
        return t.ctypes.data if self.emu else t.data_ptr()

    def get(self, t, dtype, shape):
        if self.emu:
            return t.view(dtype).reshape(shape)
        return t.cpu().numpy().view(dtype).reshape(shape)

    def sync(self):
        if not self.emu:
            self.torch.cuda.synchronize()

    def stream(self):
        return 0 if self.emu else self.torch.cuda.current_stream().cuda_stream


def _call(bp, fname, vals, bufs, D):
    args = [convert_to_ctype(p.typ.c_type_name(), vals[p.name]) for p in bp.params]
    for f in bp.inputs + bp.outputs:
        args.append(ctypes.c_void_p(D.ptr(bufs[f.name])))
    args.append(ctypes.c_void_p(D.stream()))
    fn = bp.lib['pipeline_%s_%s' % (bp.name, fname)]
    fn.restype = ctypes.c_int
    return fn(*args)


def _arena_bytes():
    if emulating():
        return None
    from polymage_amd import runtime
    return int(runtime.rt().pmx_arena_bytes())


@pytest.mark.parametrize("app", sorted(CASES))
def test_noalloc_equals_dev_on_two_calls(gpu, app):
    out, pipe, values, sizes = CASES[app]()
    print("case noalloc", app, sizes, flush=True)
    bp = gpu_build(pipe)
    orc = cemit.build([out], 'noalloc_%s_oracle' % app)
    D = _Dev()
    err = bp.lib['pmx_last_error']
    err.restype = ctypes.c_char_p
    arena = None
    for k, (a, b) in enumerate(sizes):
        vals, oshape = values(a, b, 31 + k)
        dt = out.typ.np_dtype()
        fill = np.full(oshape, -7.0 if np.dtype(dt).kind == 'f' else 7, dt)
        ref = fill.copy()
        orc(dict(vals, **{out.name: ref}))
        bufs_a = dict((f.name, D.put(vals[f.name])) for f in bp.inputs)
        bufs_b = dict(bufs_a)
        bufs_a[out.name] = D.put(fill.copy())
        bufs_b[out.name] = D.put(fill.copy())
        assert _call(bp, 'dev', vals, bufs_a, D) == 0, err()
        D.sync()
        if k == 0:
            arena = _arena_bytes()
        assert _call(bp, 'dev_noalloc', vals, bufs_b, D) == 0, err()
        D.sync()
        assert bp.lib['pipeline_%s_status' % bp.name]() == 0
        assert _arena_bytes() == arena
        ga = D.get(bufs_a[out.name], dt, oshape)
        gb = D.get(bufs_b[out.name], dt, oshape)
        assert bits_equal(ga, ref), (app, a, b)
        assert bits_equal(gb, ga), (app, a, b)


def test_noalloc_refuses_sizes_beyond_the_static_scratch(gpu):
    """A local Laplacian of 2560x1536 needs more intermediate storage than
    the static scratch holds: the call returns -12 with a message and
    launches nothing (the output keeps its bytes)."""
    from polymage_amd.apps import workloads
    w = workloads.laplacian(tiles=(32, 64), with_data=False)
    print("case noalloc oversize", flush=True)
    bp = gpu_build(w.pipe)
    D = _Dev()
    img = np.zeros((2560, 1536, 3), np.uint16)
    fill = np.full((2560, 1536, 3), 7, np.uint16)
    bufs = {'img_colour': D.put(img), w.out_func.name: D.put(fill)}
    rc = _call(bp, 'dev_noalloc', w.values, bufs, D)
    D.sync()
    err = bp.lib['pmx_last_error']
    err.restype = ctypes.c_char_p
    assert rc == -12, err()
    assert b'static scratch' in err()
    assert np.array_equal(D.get(bufs[w.out_func.name], np.uint16, fill.shape), fill)
