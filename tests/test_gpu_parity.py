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

from helpers import build, bits_equal, gpu_build, emulating

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


@pytest.mark.parametrize("R,C,tiles,threads", [(253, 301, [32, 64], 256), (130, 517, [16, 512], 512),
                                               (2046, 2046, [16, 512], 512)])
def test_harris_stored_products_bitexact(gpu, R, C, tiles, threads):
    """Products stored in LDS instead of recomputed at every box-filter tap."""
    print("case harris stored products", R, C, tiles, threads, flush=True)
    g, o = _harris(R, C, tiles, gpu_options={'no_offset_inline': 1, 'threads': threads})
    assert bits_equal(g, o)


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


def _bilateral(rows, cols, tiles, seed=5, app=False, grouping=None, gpu_options=None):
    from polymage_amd.apps.bilateral import bilateral_grid, bilateral_grid_full
    pd = {}
    out = (bilateral_grid if app else bilateral_grid_full)(pd)
    pipe = build(out, pd, {'rows': rows, 'cols': cols}, 'bilateral', tiles, group_size=5,
                 grouping=grouping, gpu_options=gpu_options)
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


@pytest.mark.parametrize("rows,cols,tiles,grid_tiles", [(100, 77, [16, 32], None), (256, 192, [32, 64], None),
                                                        (2560, 1536, [32, 64], None),
                                                        (100, 77, [16, 32], [4, 8]),
                                                        (2560, 1536, [32, 64], [8, 8])])
def test_bilateral_sliced_apart_bitexact(gpu, rows, cols, tiles, grid_tiles):
    """The caller's grouping: the grid blurs one kernel each, the slice
    (interpolation and division) one kernel; with ``grid_tiles`` the blur
    kernels get tile sizes of their own (gpu_options['group_tiles'])."""
    print("case bilateral grouped", rows, cols, tiles, grid_tiles, flush=True)
    opts = {'group_tiles': dict((n, grid_tiles) for n in ('blurz', 'blurx', 'blury'))} \
        if grid_tiles else None
    g, o = _bilateral(rows, cols, tiles, seed=1003, grouping=[['interpolated', 'filtered']],
                      gpu_options=opts)
    assert bits_equal(g, o)


def test_bilateral_full_size(gpu):
    print("case bilateral 2560x1536", flush=True)
    g, o = _bilateral(2560, 1536, [32, 64], seed=1003)
    assert bits_equal(g, o)


def _laplacian(Rin, Cin, tiles, K=8, J=4, io='u16', mode='double', seed=9, group_size=5,
               gpu_options=None, inline=None):
    from polymage_amd.apps.local_laplacian import local_laplacian
    pd = {}
    out = local_laplacian(pd, K=K, J=J, io=io, lut_hi_inclusive=(io == 'u16'))
    lo, hi = pd['pad']
    R, C = Rin + lo + hi, Cin + lo + hi
    pipe = build(out, pd, {'R': R, 'C': C}, 'laplacian', tiles, group_size=group_size,
                 literal_mode=mode, gpu_options=gpu_options, inline=inline)
    g_ = np.random.default_rng(seed)
    if io == 'u16':
        img = np.ascontiguousarray((inputs.smooth_field(g_, Rin, Cin, 3) * 65535).round()
                                   .astype(np.uint16).transpose(1, 2, 0))
        dt = np.uint16
    else:
        img = g_.integers(0, 256, (Rin, Cin, 3)).astype(np.uint8)
        dt = np.uint8
    alpha = 1.0 / (K - 1)
    vals = {'R': R, 'C': C, 'alpha': alpha, 'beta': 1.0, 'img_colour': img}
    g = np.zeros((Rin, Cin, 3), dt)
    o = np.zeros((Rin, Cin, 3), dt)
    orc = cemit.build([out], 'lap_oracle_%d_%d_%s_%s' % (K, J, io, mode), literal_mode=mode)
    gpu_build(pipe).run_host(dict(vals, laplacian=g))
    orc(dict(vals, laplacian=o))
    return g, o


@pytest.mark.parametrize("Rin,Cin,tiles", [(96, 80, [16, 32]), (200, 136, [32, 64])])
def test_laplacian_u16_bitexact(gpu, Rin, Cin, tiles):
    print("case laplacian u16", Rin, Cin, tiles, flush=True)
    g, o = _laplacian(Rin, Cin, tiles)
    assert np.array_equal(g, o)


def test_laplacian_video_demo_form(gpu):
    print("case laplacian u8 K4 J5", flush=True)
    g, o = _laplacian(160, 128, [32, 64], K=4, J=5, io='u8', mode='float12')
    assert np.array_equal(g, o)


def test_laplacian_full_size(gpu):
    print("case laplacian 2560x1536", flush=True)
    g, o = _laplacian(2560, 1536, [32, 64], seed=1004)
    assert np.array_equal(g, o)


def _campipe_frames(n, seed, variant='batch', sizes=None):
    g = np.random.default_rng(seed)
    frames = []
    for i in range(n):
        if sizes:
            R, C = sizes[i]
        else:
            R, C = 32 * int(g.integers(12, 20)), 32 * int(g.integers(16, 24))
        dt = np.uint16 if variant == 'batch' else np.int16
        frames.append(g.integers(0, 1024, (R, C)).astype(dt))
    return frames


def _campipe_vals(img):
    from polymage_amd.apps.campipe import MATRIX_3200, MATRIX_7000
    R, C = img.shape
    return {'R': R, 'C': C, 'colour_temp': 3700.0, 'gamma': 2.2, 'contrast': 50.0, 'img': img,
            'matrix_3200': np.array(MATRIX_3200, np.float32),
            'matrix_7000': np.array(MATRIX_7000, np.float32)}


def _interpolate(R, C, tiles, gs=5, seed=11, inline=None):
    from polymage_amd.apps.interpolate import interpolate
    pd = {}
    out = interpolate(pd)
    pipe = build(out, pd, {'R': R, 'C': C}, 'interpolate', tiles, group_size=gs, inline=inline)
    img = inputs.smooth_field(np.random.default_rng(seed), R + 2, C + 2, 4)
    img[3] = 0.05 + 0.95 * img[3]
    img = np.ascontiguousarray(img, np.float32)
    g = np.full((3, R, C), -7.0, np.float32)
    o = g.copy()
    gpu_build(pipe).run_host({'R': R, 'C': C, 'img': img, 'interpolate': g})
    cemit.build([out], 'interp_oracle')({'R': R, 'C': C, 'img': img, 'interpolate': o})
    return g, o


@pytest.mark.parametrize("R,C,tiles,gs", [(70, 90, [16, 32], 5), (301, 517, [16, 64], 5),
                                          (130, 94, [32, 64], 8)])
def test_interpolate_bitexact(gpu, R, C, tiles, gs):
    print("case interpolate", R, C, tiles, gs, flush=True)
    g, o = _interpolate(R, C, tiles, gs)
    assert bits_equal(g, o)


@pytest.mark.parametrize("R,C,tiles,gs", [(70, 90, [16, 32], 5), (301, 517, [8, 64], 5),
                                          (2560, 1536, [8, 64], 5)])
def test_interpolate_inlined_level0_bitexact(gpu, R, C, tiles, gs):
    """The premultiplied level 0 inlined into its two readers (the
    benchmark schedule) against the oracle of the pipeline as written."""
    print("case interpolate inlined level 0", R, C, tiles, gs, flush=True)
    g, o = _interpolate(R, C, tiles, gs, seed=1006, inline=['downsampled_L0'])
    assert bits_equal(g, o)


def test_interpolate_full_size(gpu):
    print("case interpolate 2560x1536", flush=True)
    g, o = _interpolate(2560, 1536, [32, 64], seed=1006)
    assert bits_equal(g, o)


def _blend(R, C, tiles, gs=5, seed=12):
    from polymage_amd.apps.pyramid_blend import pyramid_blend
    pd = {}
    out = pyramid_blend(pd)
    pipe = build(out, pd, {'R': R, 'C': C}, 'blend', tiles, group_size=gs)
    g_ = np.random.default_rng(seed)
    vals = {'R': R, 'C': C,
            'img1': np.ascontiguousarray(inputs.smooth_field(g_, R, C, 3), np.float32),
            'img2': np.ascontiguousarray(inputs.smooth_field(g_, R, C, 3), np.float32),
            'mask': np.ascontiguousarray(inputs.smooth_field(g_, R, C, 1)[0], np.float32)}
    g = np.full((3, R, C), -7.0, np.float32)
    o = g.copy()
    gpu_build(pipe).run_host(dict(vals, blend=g))
    cemit.build([out], 'blend_oracle')(dict(vals, blend=o))
    return g, o


@pytest.mark.parametrize("R,C,tiles,gs", [(96, 130, [16, 32], 5), (257, 391, [16, 64], 3),
                                          (150, 203, [16, 32], 6)])
def test_blend_bitexact(gpu, R, C, tiles, gs):
    print("case blend", R, C, tiles, gs, flush=True)
    g, o = _blend(R, C, tiles, gs)
    assert bits_equal(g, o)        # the written window and the untouched frame


def test_blend_full_size(gpu):
    print("case blend 2560x1536", flush=True)
    g, o = _blend(2560, 1536, [16, 64], seed=1007)
    assert bits_equal(g, o)


@pytest.mark.parametrize("variant,gs", [('reference', 5), ('batch', 5), ('batch', 40)])
def test_campipe_bitexact(gpu, variant, gs):
    from polymage_amd.apps.campipe import camera_pipe
    print("case campipe", variant, gs, flush=True)
    pd = {}
    out = camera_pipe(pd, variant=variant)
    img = _campipe_frames(1, 21, variant, sizes=[(416, 608)])[0]
    R, C = img.shape
    pipe = build(out, pd, {'R': R, 'C': C}, 'campipe', [16, 64], group_size=gs)
    vals = _campipe_vals(img)
    g = np.zeros((3, R - 48, C - 64), np.uint8)
    o = np.zeros_like(g)
    orc = cemit.build([out], 'cam_oracle_' + variant)
    gpu_build(pipe).run_host(dict(vals, process=g))
    orc(dict(vals, process=o))
    assert np.array_equal(g, o)


def test_campipe_batch_persistent_and_graph(gpu):
    """Frames of different sizes in one persistent launch per kernel (and
    the same captured in a hipGraph, replayed twice) give the oracle's
    output for every frame."""
    from polymage_amd.apps.campipe import camera_pipe
    print("case campipe batch", flush=True)
    pd = {}
    out = camera_pipe(pd, variant='batch')
    frames = _campipe_frames(6, 22)
    Rm = max(f.shape[0] for f in frames)
    Cm = max(f.shape[1] for f in frames)
    pipe = build(out, pd, {'R': Rm, 'C': Cm}, 'campipe', [16, 64], group_size=40)
    bp = gpu_build(pipe)
    orc = cemit.build([out], 'cam_oracle_batch')
    if emulating():
        # the kernel emulator takes host arrays; graph capture is not emulated
        for mode in ('persistent', 'frames'):
            recs, refs = [], []
            for img in frames:
                R, C = img.shape
                vals = _campipe_vals(img)
                o = np.zeros((3, R - 48, C - 64), np.uint8)
                orc(dict(vals, process=o))
                refs.append(o)
                recs.append(dict(vals, process=np.zeros_like(o)))
            bp.run_batch(recs, mode)
            for rec, r in zip(recs, refs):
                assert np.array_equal(rec['process'], r), mode
        return
    import torch
    dev = torch.device('cuda:0')
    recs, outs, refs = [], [], []
    m32 = torch.from_numpy(_campipe_vals(frames[0])['matrix_3200']).to(dev)
    m70 = torch.from_numpy(_campipe_vals(frames[0])['matrix_7000']).to(dev)
    for img in frames:
        R, C = img.shape
        vals = _campipe_vals(img)
        o = np.zeros((3, R - 48, C - 64), np.uint8)
        orc(dict(vals, process=o))
        refs.append(o)
        d_img = torch.from_numpy(img.view(np.int16)).to(dev)
        d_out = torch.zeros((3, R - 48, C - 64), dtype=torch.uint8, device=dev)
        outs.append(d_out)
        recs.append(dict(vals, img=d_img, process=d_out, matrix_3200=m32, matrix_7000=m70))
    h = bp.batch_create(recs)
    stream = torch.cuda.current_stream().cuda_stream
    for mode in ('persistent', 'graph', 'graph', 'frames'):
        for d in outs:
            d.zero_()
        bp.batch_run(h, stream, mode)
        torch.cuda.synchronize()
        for d, r in zip(outs, refs):
            assert np.array_equal(d.cpu().numpy(), r), mode
    bp.batch_destroy(h)
