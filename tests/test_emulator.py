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


@pytest.mark.parametrize("R,C,tiles,opts", [(37, 101, [16, 32], None), (62, 70, [32, 64], {'block': (2, 4)}),
                                            (9, 13, [4, 8], {'block': (4, 4)}),
                                            (61, 75, [8, 32], {'roll': 3}),
                                            (53, 97, [16, 64], {'no_offset_inline': 1, 'threads': 512})])
def test_emulated_harris(R, C, tiles, opts):
    pd = {}
    out = harris_pipe(pd)
    pipe = build(out, pd, {'R': R, 'C': C}, 'harris', tiles, gpu_options=opts)
    img = np.random.default_rng(R).random((R + 2, C + 2)).astype(np.float32)
    g = np.full((R + 2, C + 2), -7.0, np.float32)
    o = g.copy()
    kernel_emu.build(pipe).run_host({'R': R, 'C': C, 'img': img, 'harris': g})
    cemit.build([out], 'harris_oracle_double')({'R': R, 'C': C, 'img': img, 'harris': o})
    assert bits_equal(g, o)


@pytest.mark.parametrize("opts", [None, {'block': (8, 1)}, {'block': (1, 4)}, {'block': (4, 4)},
                                  {'stage': 1}, {'stage': 1, 'nt_stores': 1}, {'no_fuse': 1},
                                  {'roll': 3}, {'roll': 2, 'block': (2, 1)}])
def test_emulated_unsharp_options(opts):
    R, C = 45, 77
    pd = {}
    out = unsharp_mask_full(pd)
    pipe = build(out, pd, {'R': R, 'C': C}, 'unsharp', [16, 32], group_size=5, gpu_options=opts)
    if opts and opts.get('stage'):
        from polymage_amd.hip_emit import plan_pipeline
        assert plan_pipeline(pipe)[0].staged == ['img']
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


def test_emulated_campipe_batch_queues():
    """Frames of three sizes through the persistent per-XCD work queues
    (emulated grid smaller than the item count, so queues are stolen
    from) give the oracle's output for every frame."""
    from polymage_amd.apps.campipe import camera_pipe, MATRIX_3200, MATRIX_7000
    pd = {}
    out = camera_pipe(pd, variant='batch')
    g = np.random.default_rng(3)
    sizes = [(128, 192), (160, 256), (96, 160)]
    pipe = build(out, pd, {'R': 160, 'C': 256}, 'campipe', [16, 64], group_size=40)
    orc = cemit.build([out], 'cam_oracle_batch')
    recs, refs = [], []
    for (R, C) in sizes:
        img = g.integers(0, 1024, (R, C)).astype(np.uint16)
        vals = {'R': R, 'C': C, 'colour_temp': 3700.0, 'gamma': 2.2, 'contrast': 50.0, 'img': img,
                'matrix_3200': np.array(MATRIX_3200, np.float32),
                'matrix_7000': np.array(MATRIX_7000, np.float32)}
        o = np.zeros((3, R - 48, C - 64), np.uint8)
        orc(dict(vals, process=o))
        refs.append(o)
        recs.append(dict(vals, process=np.full_like(o, 7)))
    kernel_emu.build(pipe).run_batch(recs, 'persistent')
    for rec, r in zip(recs, refs):
        assert np.array_equal(rec['process'], r)


def test_recomputed_stage_keeps_its_float_rounding():
    """A pointwise stage recomputed at its consumer's taps (offset
    inlining) computes in double (C literals) and is stored as float; the
    substituted expression must be rounded to float at each tap exactly as
    the stored value is, or sums differ in the last bit."""
    from polymage_amd.constructs import Variable, Parameter, Interval, Image, Function, Case, Condition
    from polymage_amd.expr_types import Int, Float
    from polymage_amd.hip_emit import plan_pipeline
    R = Parameter(Int, 'R')
    x, y = Variable(Int, 'x'), Variable(Int, 'y')
    img = Image(Float, 'img', [R + 2, R + 2])
    dom = Interval(Int, 0, R + 1)
    a = Function(([x, y], [dom, dom]), Float, 'a')
    a.defn = [img(x, y) * 0.1]
    cond = Condition(x, '>=', 1) & Condition(x, '<=', R) & Condition(y, '>=', 1) & Condition(y, '<=', R)
    b = Function(([x, y], [dom, dom]), Float, 'b')
    b.defn = [Case(cond, a(x - 1, y) + a(x + 1, y) + a(x, y - 1) + a(x, y + 1))]
    pipe = build(b, {'R': R}, {'R': 61}, 'ri', [16, 32], group_size=5)
    assert plan_pipeline(pipe)[0].offset_inlined == ['a']
    v = np.random.default_rng(9).random((63, 63)).astype(np.float32)
    g = np.zeros((63, 63), np.float32)
    o = g.copy()
    kernel_emu.build(pipe).run_host({'R': 61, 'img': v, 'b': g})
    cemit.build([b], 'ri_oracle')({'R': 61, 'img': v, 'b': o})
    assert bits_equal(g, o)


def test_size_limits_and_bad_batches_are_errors():
    """Sizes whose arrays overflow 32-bit indexing are refused before any
    memory is touched (status -75 and a message naming the array), and an
    empty batch is refused (-22): the C ABI reports instead of faulting."""
    import ctypes
    pd = {}
    out = unsharp_mask_full(pd)
    pipe = build(out, pd, {'R': 64, 'C': 64}, 'unsharp', [16, 32], group_size=5)
    ep = kernel_emu.build(pipe)
    dev = ep.lib['pipeline_unsharp_dev']
    dev.restype = ctypes.c_int
    big = 40000
    rc = dev(ctypes.c_int(big), ctypes.c_int(big), ctypes.c_float(0.001), ctypes.c_float(3.0),
             ctypes.c_void_p(0), ctypes.c_void_p(0), ctypes.c_void_p(0))
    assert rc == -75
    assert b'too large for 32-bit indexing' in ep.lib.pmx_last_error()
    create = ep.lib['pipeline_unsharp_batch_create']
    create.restype = ctypes.c_int
    h = ctypes.c_void_p()
    rc = create(ctypes.c_int(0), None, None, None, None, None, None, ctypes.byref(h))
    assert rc == -22 and b'no frames' in ep.lib.pmx_last_error()


@pytest.mark.parametrize("app,R,C,tiles,gs", [('interpolate', 70, 90, [16, 32], 5),
                                              ('blend', 96, 130, [16, 32], 5)])
def test_emulated_pyramid_apps(app, R, C, tiles, gs):
    """The reference's interpolation and pyramid-blending apps: many small
    groups over scaled accesses and Selects on parity, bit for bit."""
    g_ = np.random.default_rng(R + C)
    pd = {}
    if app == 'interpolate':
        from polymage_amd.apps.interpolate import interpolate
        out = interpolate(pd)
        vals = {'R': R, 'C': C, 'img': g_.random((4, R + 2, C + 2)).astype(np.float32)}
        oname = 'interpolate'
    else:
        from polymage_amd.apps.pyramid_blend import pyramid_blend
        out = pyramid_blend(pd)
        vals = {'R': R, 'C': C, 'img1': g_.random((3, R, C)).astype(np.float32),
                'img2': g_.random((3, R, C)).astype(np.float32),
                'mask': g_.random((R, C)).astype(np.float32)}
        oname = 'blend'
    pipe = build(out, pd, {'R': R, 'C': C}, app, tiles, group_size=gs)
    g = np.full((3, R, C), -7.0, np.float32)
    o = g.copy()
    kernel_emu.build(pipe).run_host(dict(vals, **{oname: g}))
    cemit.build([out], app + '_oracle')(dict(vals, **{oname: o}))
    assert bits_equal(g, o)


def _two_handles_case(asan):
    """A small batch handle, then a larger one of the same pipeline; the
    first handle still gives the oracle's output afterwards (each handle
    owns its intermediates, so the second cannot free or share them)."""
    from polymage_amd.apps.bilateral import bilateral_grid_full
    pd = {}
    out = bilateral_grid_full(pd)
    pipe = build(out, pd, {'rows': 128, 'cols': 128}, 'bilateral', [16, 32], group_size=5)
    ep = kernel_emu.build(pipe, asan=asan)
    status = ep.lib['pipeline_bilateral_status']
    orc = cemit.build([out], 'bilateral_oracle_full')

    def frames(sizes, seed):
        recs, refs = [], []
        for k, (rows, cols) in enumerate(sizes):
            img = inputs.smooth_field(np.random.default_rng(seed + k), rows, cols, 1)[0]
            o = np.full((1, rows, cols), -7.0, np.float32)
            orc({'rows': rows, 'cols': cols, 'input': img, 'filtered': o})
            refs.append(o)
            recs.append({'rows': rows, 'cols': cols, 'input': img,
                         'filtered': np.full_like(o, -7.0)})
        return recs, refs

    small, small_ref = frames([(48, 40), (40, 56)], 11)
    big, big_ref = frames([(112, 96), (96, 120)], 21)
    ha = ep.batch_create(small)
    hb = ep.batch_create(big)
    ep.batch_run(hb)
    for rec in small:
        rec['filtered'][...] = -7.0
    ep.batch_run(ha)
    assert status() == 0
    for rec, r in zip(small + big, small_ref + big_ref):
        assert bits_equal(rec['filtered'], r)
    ep.batch_destroy(hb)
    ep.batch_run(ha)
    assert status() == 0
    for rec, r in zip(small, small_ref):
        assert bits_equal(rec['filtered'], r)
    ep.batch_destroy(ha)


def test_emulated_batch_handles_own_intermediates():
    _two_handles_case(asan=False)


def test_emulated_batch_handles_under_asan():
    """The same two-handle case with AddressSanitizer: a handle whose
    intermediates a later handle freed would be a use-after-free."""
    if os.environ.get('PMX_EMU_CHILD'):
        pytest.skip("runs only as the parent's child case")
    env = dict(os.environ, PMX_EMU_CHILD='1', ASAN_OPTIONS='detect_leaks=0',
               LD_PRELOAD=kernel_emu.asan_runtime())
    code = ('import sys; sys.path[:0] = [%r, %r]; import test_emulator as t; '
            't._two_handles_case(asan=True); print("two handles ok")' % (HERE, os.path.dirname(HERE)))
    r = subprocess.run([sys.executable, '-c', code], env=env, stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, timeout=900, cwd=os.path.dirname(HERE))
    out = r.stdout.decode(errors='replace')
    assert r.returncode == 0 and 'two handles ok' in out, out[-3000:]


@pytest.mark.parametrize("R,C,tiles,threads", [(45, 77, [16, 60], 128), (37, 130, [16, 124], 256),
                                               (16, 252, [16, 252], 512)])
def test_emulated_unsharp_forwarded_loads(R, C, tiles, threads):
    """The benchmark's unsharp shape (one item per thread, 8x1 blocks): the
    mask takes the image values the blur loaded from registers instead of
    reading them again; partial tiles and the columns the mask does not
    cover fall back to loads.  Bit-exact against the oracle, also under
    AddressSanitizer."""
    pd = {}
    out = unsharp_mask_full(pd)
    pipe = build(out, pd, {'R': R, 'C': C}, 'unsharp', tiles, group_size=5,
                 gpu_options={'threads': threads, 'block': (8, 1)})
    assert 'fw_ok' in pipe.generate_code().hip_src
    img = inputs.smooth_field(np.random.default_rng(R + C), R + 4, C + 4, 3)
    vals = {'R': R, 'C': C, 'threshold': 0.001, 'weight': 3.0, 'img': img}
    o = np.full((3, R, C), -7.0, np.float32)
    cemit.build([out], 'unsharp_oracle_unsharp_mask_full')(dict(vals, mask=o))
    for asan in (False, True):
        if asan and not os.environ.get('PMX_EMU_CHILD'):
            env = dict(os.environ, PMX_EMU_CHILD='1', ASAN_OPTIONS='detect_leaks=0',
                       LD_PRELOAD=kernel_emu.asan_runtime())
            code = ('import sys; sys.path[:0] = [%r, %r]; import test_emulator as t; '
                    't._forwarded_asan(%d, %d, %r, %d); print("forwarded ok")'
                    % (HERE, os.path.dirname(HERE), R, C, tiles, threads))
            r = subprocess.run([sys.executable, '-c', code], env=env, stdout=subprocess.PIPE,
                               stderr=subprocess.STDOUT, timeout=900, cwd=os.path.dirname(HERE))
            txt = r.stdout.decode(errors='replace')
            assert r.returncode == 0 and 'forwarded ok' in txt, txt[-3000:]
            continue
        if asan:
            continue
        g = np.full((3, R, C), -7.0, np.float32)
        kernel_emu.build(pipe).run_host(dict(vals, mask=g))
        assert bits_equal(g, o)


def _forwarded_asan(R, C, tiles, threads):
    pd = {}
    out = unsharp_mask_full(pd)
    pipe = build(out, pd, {'R': R, 'C': C}, 'unsharp', tiles, group_size=5,
                 gpu_options={'threads': threads, 'block': (8, 1)})
    img = inputs.smooth_field(np.random.default_rng(R + C), R + 4, C + 4, 3)
    vals = {'R': R, 'C': C, 'threshold': 0.001, 'weight': 3.0, 'img': img}
    o = np.full((3, R, C), -7.0, np.float32)
    cemit.build([out], 'unsharp_oracle_unsharp_mask_full')(dict(vals, mask=o))
    g = np.full((3, R, C), -7.0, np.float32)
    kernel_emu.build(pipe, asan=True).run_host(dict(vals, mask=g))
    assert bits_equal(g, o)


@pytest.mark.parametrize("app,R,C,tiles,threads,block", [
    ('unsharp', 45, 77, [16, 60], 128, (8, 1)), ('unsharp', 37, 130, [16, 124], 256, (4, 1)),
    ('unsharp', 16, 252, [16, 252], 512, (8, 1)), ('harris', 62, 70, [16, 32], 256, (8, 1)),
    ('harris', 130, 517, [16, 128], 512, (8, 1)), ('harris', 5, 9, [4, 4], 64, (8, 1))])
def test_emulated_lane_fusion(app, R, C, tiles, threads, block):
    """gpu_options lanes=1: the producer sweep (unsharp's x-blur, Harris'
    Sobel Ix/Iy) is computed in registers inside its consumer's sweep and
    read across lanes with wavefront shuffles instead of through LDS;
    partial tiles, the wavefront edge lanes and tiny images included.
    Bit-exact against the oracle (the emulator runs __shfl on 64-thread
    wavefronts)."""
    pd = {}
    out = unsharp_mask_full(pd) if app == 'unsharp' else harris_pipe(pd)
    opts = {'threads': threads, 'block': block, 'lanes': 1}
    if app == 'unsharp':
        pipe = build(out, pd, {'R': R, 'C': C}, 'unsharp', tiles, group_size=5, gpu_options=opts)
        img = inputs.smooth_field(np.random.default_rng(R + C), R + 4, C + 4, 3)
        vals = {'R': R, 'C': C, 'threshold': 0.001, 'weight': 3.0, 'img': img}
        shape, oname, orc = (3, R, C), 'mask', 'unsharp_oracle_unsharp_mask_full'
    else:
        pipe = build(out, pd, {'R': R, 'C': C}, 'harris', tiles, gpu_options=opts)
        vals = {'R': R, 'C': C, 'img': np.random.default_rng(R).random((R + 2, C + 2)).astype(np.float32)}
        shape, oname, orc = (R + 2, C + 2), 'harris', 'harris_oracle_double'
    assert 'pmx_shfl' in pipe.generate_code().hip_src
    o = np.full(shape, -7.0, np.float32)
    g = o.copy()
    cemit.build([out], orc)(dict(vals, **{oname: o}))
    kernel_emu.build(pipe).run_host(dict(vals, **{oname: g}))
    assert bits_equal(g, o)


@pytest.mark.parametrize("app,R,C,tiles,threads,block,fw", [
    ('unsharp', 45, 77, [16, 60], 128, (2, 4), 0), ('unsharp', 37, 130, [16, 124], 256, (4, 4), 0),
    ('unsharp', 16, 252, [16, 252], 512, (2, 4), 1), ('harris', 62, 70, [16, 32], 256, (4, 4), 0),
    ('harris', 130, 517, [16, 128], 512, (2, 4), 0), ('harris', 5, 9, [4, 4], 64, (4, 4), 0)])
def test_emulated_vector_lds_rows(app, R, C, tiles, threads, block, fw):
    """gpu_options vlds=1 with a 4-wide block: a stage reading LDS loads
    each producer row its block touches with aligned 128-bit reads
    (``pmx_ld4``; the emulator fails the run on a misaligned one) and
    takes its taps from those registers; blocks whose rows are not
    aligned (with load forwarding the item grid starts at the producer's
    region) and partial blocks take the per-point path.  Bit-exact
    against the oracle."""
    pd = {}
    out = unsharp_mask_full(pd) if app == 'unsharp' else harris_pipe(pd)
    opts = {'threads': threads, 'block': block, 'vlds': 1}
    if not fw:
        opts['no_forward'] = 1
    if app == 'unsharp':
        pipe = build(out, pd, {'R': R, 'C': C}, 'unsharp', tiles, group_size=5, gpu_options=opts)
        img = inputs.smooth_field(np.random.default_rng(R + C), R + 4, C + 4, 3)
        vals = {'R': R, 'C': C, 'threshold': 0.001, 'weight': 3.0, 'img': img}
        shape, oname, orc = (3, R, C), 'mask', 'unsharp_oracle_unsharp_mask_full'
    else:
        pipe = build(out, pd, {'R': R, 'C': C}, 'harris', tiles, gpu_options=opts)
        vals = {'R': R, 'C': C, 'img': np.random.default_rng(R).random((R + 2, C + 2)).astype(np.float32)}
        shape, oname, orc = (R + 2, C + 2), 'harris', 'harris_oracle_double'
    assert 'pmx_ld4' in pipe.generate_code().hip_src
    o = np.full(shape, -7.0, np.float32)
    g = o.copy()
    cemit.build([out], orc)(dict(vals, **{oname: o}))
    kernel_emu.build(pipe).run_host(dict(vals, **{oname: g}))
    assert bits_equal(g, o)


@pytest.mark.parametrize("tiles,opts", [([16, 32], {'block_lds': (1, 4)}),
                                        ([16, 32], {'block_lds': (2, 4)}),
                                        ([32, 64], {'block': (1, 4), 'block_lds': (2, 4)}),
                                        ([8, 16], {'block_lds': (2, 4), 'threads': 64})])
def test_emulated_laplacian_downsampling_rows(monkeypatch, tiles, opts):
    """The local Laplacian's downsampling stages read their LDS producer
    at 2 * x + c (rows) and 2 * y + c (columns): in 4-wide blocks those
    reads come from 128-bit row loads (per-item base at twice the item's
    point); the padded input inlined, gather-inlined blends.  Bit-exact
    against the oracle, every 128-bit load aligned."""
    import test_gpu_parity as P
    monkeypatch.setattr(P, 'gpu_build', kernel_emu.build)
    g, o = P._laplacian(96, 80, tiles, gpu_options=opts, inline=['colour'], group_size=8)
    assert np.array_equal(g, o)


@pytest.mark.parametrize("tiles,opts", [([8, 16], {'kblock': 1}), ([16, 32], {'kblock': 1}),
                                        ([8, 32], {'kblock': 1, 'threads': 128})])
def test_emulated_laplacian_level_blocks(monkeypatch, tiles, opts):
    """gpu option kblock: the K intensity levels stay whole in each tile
    and every thread computes all K levels of its (row, column), sharing
    the gray value, its table index and the index arithmetic; partial
    tiles included.  Bit-exact against the oracle."""
    import test_gpu_parity as P
    from polymage_amd.hip_emit import plan_pipeline
    monkeypatch.setattr(P, 'gpu_build', kernel_emu.build)
    g, o = P._laplacian(96, 80, tiles, gpu_options=opts, inline=['colour'], group_size=8)
    assert np.array_equal(g, o)


@pytest.mark.parametrize("rows,cols,tiles", [(100, 77, [16, 32]), (256, 192, [32, 64])])
def test_emulated_bilateral_pair_blocks(monkeypatch, rows, cols, tiles):
    """gpu option kblock on the bilateral grid: the (value, weight) pair of
    the slice is one block, so both read the pixel, its grid cell and
    weights once.  Bit-exact against the oracle."""
    import test_gpu_parity as P
    monkeypatch.setattr(P, 'gpu_build', kernel_emu.build)
    g, o = P._bilateral(rows, cols, tiles, gpu_options={'kblock': 1})
    assert bits_equal(g, o)
