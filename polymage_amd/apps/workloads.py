"""The benchmark workloads at their configured sizes.

Each entry builds the DSL pipeline with the schedule named in the config
and describes the data: arrays, MPix (output spatial pixels) and the
ideal HBM traffic (all input bytes + bytes of the output region written).
"""

import numpy as np

from .. import buildPipeline
from ..constructs import Condition
from . import inputs


class Workload(object):
    """``pipe`` may be given as a zero-argument callable; the pipeline is
    then built (buildPipeline) on first access, so the data, shapes and
    output function of a workload are available without planning it."""

    def __init__(self, name, pipe, out_func, values, out_shape, mpix, ideal_bytes,
                 defined=None, dtype='f32'):
        self.name = name
        self._pipe = pipe
        self.out_func = out_func
        self.values = values            # host arrays + scalar params
        self.out_shape = out_shape
        self.mpix = mpix
        self.ideal_bytes = ideal_bytes
        self.defined = defined          # slice tuple of the written region
        self.dtype = dtype

    @property
    def pipe(self):
        if callable(self._pipe):
            self._pipe = self._pipe()
        return self._pipe


def _pipe(out, pd, sizes, name, tiles, group_size, literal_mode='double', gpu_options=None,
          extra_est=(), inline=()):
    est = [(pd[k], v) for k, v in sizes.items()] + list(extra_est)
    cons = [Condition(pd[k], '==', v) for k, v in sizes.items()]
    return lambda: buildPipeline([out], param_estimates=est, param_constraints=cons,
                                 tile_sizes=tiles, group_size=group_size, pipe_name=name,
                                 inline_directives=list(inline),
                                 literal_mode=literal_mode, gpu_options=gpu_options)


def harris(rows=2048, cols=2048, tiles=(32, 256), gpu_options=None, with_data=True,
           group_size=11, literal_mode='double'):
    from .harris import harris_pipe
    pd = {}
    out = harris_pipe(pd)
    R, C = rows - 2, cols - 2
    pipe = _pipe(out, pd, {'R': R, 'C': C}, 'harris', list(tiles), group_size,
                 gpu_options=gpu_options, literal_mode=literal_mode)
    vals = {'R': R, 'C': C}
    if with_data:
        vals['img'] = inputs.harris_input(rows, cols)
    out_px = (R - 2) * (C - 2)
    ideal = rows * cols * 4 + out_px * 4
    return Workload('harris', pipe, out, vals, (rows, cols), out_px / 1e6, ideal,
                    defined=(slice(2, R), slice(2, C)))


def unsharp(rows=2848, cols=4256, tiles=(32, 256), gpu_options=None, with_data=True,
            group_size=5, literal_mode='double'):
    from .unsharp import unsharp_mask_full
    pd = {}
    out = unsharp_mask_full(pd)
    R, C = rows - 4, cols - 4
    pipe = _pipe(out, pd, {'R': R, 'C': C}, 'unsharp', list(tiles), group_size,
                 gpu_options=gpu_options, literal_mode=literal_mode)
    vals = {'R': R, 'C': C, 'threshold': 0.001, 'weight': 3.0}
    if with_data:
        vals['img'] = inputs.unsharp_input(rows, cols)
    out_px = R * C
    ideal = 3 * rows * cols * 4 + 3 * R * C * 4
    return Workload('unsharp', pipe, out, vals, (3, R, C), out_px / 1e6, ideal)


def bilateral(rows=2560, cols=1536, tiles=(32, 64), gpu_options=None, with_data=True,
              group_size=5, literal_mode='double'):
    from .bilateral import bilateral_grid_full
    pd = {}
    out = bilateral_grid_full(pd)
    # the slice's blocks without 128-bit LDS row loads (48.1 -> 47.6 us)
    gpu_options = dict({'no_vlds': 1}, **(gpu_options or {}))
    pipe = _pipe(out, pd, {'rows': rows, 'cols': cols}, 'bilateral', list(tiles), group_size,
                 gpu_options=gpu_options, literal_mode=literal_mode)
    vals = {'rows': rows, 'cols': cols}
    if with_data:
        vals['input'] = inputs.bilateral_input(rows, cols)
    ideal = rows * cols * 4 * 2
    return Workload('bilateral', pipe, out, vals, (1, rows, cols), rows * cols / 1e6, ideal)


def laplacian(rows=2560, cols=1536, tiles=(32, 64), gpu_options=None, with_data=True,
              group_size=5, K=8, J=4, literal_mode='double', inline=('colour',)):
    """Local Laplacian, uint16 RGB, K intensity levels, J pyramid levels,
    alpha = 1/(K-1), beta = 1; the input is padded by edge replication
    inside the pipeline so the output has the input's size.  ``inline``
    names functions to inline into their readers (inline_directives): by
    default the padded, widened input ``colour``, whose definition is a
    cast of a clamped read of the input, so its readers read the input
    directly (same values) instead of a padded copy written and read
    back."""
    from .local_laplacian import local_laplacian
    from ..pipe import get_funcs
    pd = {}
    out = local_laplacian(pd, K=K, J=J, io='u16', lut_hi_inclusive=True)
    lo, hi = pd['pad']
    R, C = rows + lo + hi, cols + lo + hi
    funcs = [f for f in get_funcs([out])[0] if f.name in set(inline)]
    # the K-level groups (the remap with the first downsampling, and the
    # level-2 and level-3 downsamplings) keep the K levels whole in a tile,
    # each thread computing all K for its point (the gray value, its table
    # index and the index arithmetic shared), on 8x16 tiles of their own so
    # K times the LDS still leaves several workgroups per CU
    # (148.8 -> 141.2 us per call at the learned schedule); with that, the
    # larger per-thread blocks of a higher fixed cost per sweep iteration
    # in the block choice (141.2 -> 139.6 us)
    kgroups = dict(('D_gPyramid_L%d' % j if j > 1 else 'gPyramid_L0', [8, 16])
                   for j in range(1, min(J, 4)))
    gpu_options = dict({'kblock': 1, 'group_tiles': kgroups, 'iter_cost': 10},
                       **(gpu_options or {}))
    pipe = _pipe(out, pd, {'R': R, 'C': C}, 'laplacian', list(tiles), group_size,
                 gpu_options=gpu_options, literal_mode=literal_mode, inline=funcs)
    vals = {'R': R, 'C': C, 'alpha': 1.0 / (K - 1), 'beta': 1.0}
    if with_data:
        vals['img_colour'] = inputs.laplacian_input(rows, cols)
    ideal = rows * cols * 3 * 2 * 2
    return Workload('laplacian', pipe, out, vals, (rows, cols, 3), rows * cols / 1e6, ideal,
                    dtype='u16')


def campipe(rows=1216, cols=1664, tiles=(16, 64), gpu_options=None, with_data=True,
            group_size=40, literal_mode='double'):
    """One frame of the batched camera pipeline (uint16 10-bit GRBG Bayer,
    3700 K, gamma 2.2, contrast 50, 3x3 sharpen, uint8 RGB out)."""
    from .campipe import camera_pipe, MATRIX_3200, MATRIX_7000
    pd = {}
    out = camera_pipe(pd, variant='batch')
    # the largest per-thread blocks the sweeps allow (a high fixed cost per
    # sweep iteration in the block choice): its 26 stages each pay a loop
    # guard and a fast/slow split per item (48-frame batch 1.48 -> 1.44 ms)
    gpu_options = dict({'iter_cost': 80}, **(gpu_options or {}))
    pipe = _pipe(out, pd, {'R': rows, 'C': cols}, 'campipe', list(tiles), group_size,
                 gpu_options=gpu_options, literal_mode=literal_mode)
    vals = {'R': rows, 'C': cols, 'colour_temp': 3700.0, 'gamma': 2.2, 'contrast': 50.0,
            'matrix_3200': np.array(MATRIX_3200, np.float32),
            'matrix_7000': np.array(MATRIX_7000, np.float32)}
    if with_data:
        vals['img'] = inputs.rng(5).integers(0, 1024, (rows, cols)).astype(np.uint16)
    opx = (rows - 48) * (cols - 64)
    ideal = rows * cols * 2 + 3 * opx
    return Workload('campipe', pipe, out, vals, (3, rows - 48, cols - 64), opx / 1e6, ideal,
                    dtype='u8')


def interpolate(rows=2560, cols=1536, tiles=(32, 64), gpu_options=None, with_data=True,
                group_size=5, literal_mode='double', inline_level0=True):
    """Multiscale interpolation (10 pyramid levels) of an RGBA float image
    with a one-pixel ghost border; RGB float out.  Colour and alpha are
    smooth random fields, alpha kept in [0.05, 1] so the normalisation
    never divides by zero.

    ``inline_level0``: the premultiplied level 0 (``img(c) * img(3)``, a
    float product) is inlined into its two readers (the first downsampling
    and the level-0 blend), the reference's ``inline_directives``
    (inline.py:29-199); otherwise auto_group leaves it a kernel of its own
    that writes 4 full-resolution planes for two later kernels to read."""
    from .interpolate import interpolate as interp
    from ..pipe import get_funcs
    pd = {}
    out = interp(pd)
    R, C = rows, cols
    inline = [f for f in get_funcs([out])[0] if f.name == 'downsampled_L0'] if inline_level0 else []
    pipe = _pipe(out, pd, {'R': R, 'C': C}, 'interpolate', list(tiles), group_size,
                 gpu_options=gpu_options, literal_mode=literal_mode, inline=inline)
    vals = {'R': R, 'C': C}
    if with_data:
        g = inputs.rng(6)
        img = inputs.smooth_field(g, R + 2, C + 2, 4)
        img[3] = 0.05 + 0.95 * img[3]
        vals['img'] = np.ascontiguousarray(img, np.float32)
    ideal = 4 * (R + 2) * (C + 2) * 4 + 3 * R * C * 4
    return Workload('interpolate', pipe, out, vals, (3, R, C), R * C / 1e6, ideal)


def blend(rows=2560, cols=1536, tiles=(16, 64), gpu_options=None, with_data=True,
          group_size=5, literal_mode='double'):
    """Laplacian-pyramid blending (4 levels) of two float RGB images under
    a float mask; float RGB out on the window the reference defines (a
    31-point frame is not written)."""
    from .pyramid_blend import pyramid_blend
    pd = {}
    out = pyramid_blend(pd)
    R, C = rows, cols
    # smaller per-thread blocks than the default choice (a low fixed cost
    # per sweep iteration in the block choice): 207 -> 201 us per call
    gpu_options = dict({'iter_cost': 3}, **(gpu_options or {}))
    pipe = _pipe(out, pd, {'R': R, 'C': C}, 'blend', list(tiles), group_size,
                 gpu_options=gpu_options, literal_mode=literal_mode)
    vals = {'R': R, 'C': C}
    if with_data:
        g = inputs.rng(7)
        vals['img1'] = np.ascontiguousarray(inputs.smooth_field(g, R, C, 3), np.float32)
        vals['img2'] = np.ascontiguousarray(inputs.smooth_field(g, R, C, 3), np.float32)
        vals['mask'] = np.ascontiguousarray(inputs.smooth_field(g, R, C, 1)[0], np.float32)
    wr, wc = R - 60, C - 60           # rows/cols 31 .. R-30 are written
    ideal = (2 * 3 + 1) * R * C * 4 + 3 * wr * wc * 4
    return Workload('blend', pipe, out, vals, (3, R, C), wr * wc / 1e6, ideal)


DEFAULT_TILES = {'unsharp': (32, 256), 'harris': (32, 256), 'bilateral': (32, 64),
                 'laplacian': (32, 64), 'campipe': (16, 64), 'interpolate': (32, 64),
                 'blend': (16, 64)}


def prebuild_default():
    """Compile (no GPU needed) the default schedules and the schedules the
    autotuner recorded, so benchmark runs find them in the kernel cache."""
    from .. import runtime
    for name in ('harris', 'unsharp', 'bilateral', 'laplacian', 'campipe', 'interpolate', 'blend'):
        for which in ('default', 'learned', 'exhaustive'):
            runtime.build(tuned(name, which, with_data=False).pipe)
    for which in ('learned', 'exhaustive'):
        for (_, tiles, gs, gopts, _) in class_schedules('campipe', which):
            runtime.build(campipe(tiles=tiles, group_size=gs, gpu_options=gopts,
                                  with_data=False).pipe)


BUILDERS = {'harris': harris, 'unsharp': unsharp, 'bilateral': bilateral,
            'laplacian': laplacian, 'campipe': campipe, 'interpolate': interpolate,
            'blend': blend}


def schedule(app, which='learned'):
    """(tiles, group_size, gpu_options, label) of ``app`` under the
    autotuner's recorded pick ('learned' or 'exhaustive', from
    polymage_amd/tuned.json) or the default tiles ('default', or when no
    pick is recorded)."""
    from .. import tuner
    rec = tuner.tuned_schedule(app, which) if which != 'default' else None
    if rec is None:
        return tuple(DEFAULT_TILES[app]), None, None, 'default'
    gopts = {'threads': rec['threads']}
    if rec.get('lanes'):
        gopts['lanes'] = 1
    return tuple(rec['tiles']), rec['group_size'], gopts, '%s: %s' % (which, rec['key'])


def class_schedules(app, which='learned'):
    """The frame-size classes recorded for ``app`` by the autotuner
    (``autotune --size-classes``): a list of (max_area, tiles, group_size,
    gpu_options, label), smallest class first, or [] when none."""
    import json
    import os
    from .. import tuner
    if not os.path.exists(tuner.TUNED_PATH):
        return []
    with open(tuner.TUNED_PATH) as f:
        rec = json.load(f).get(app) or {}
    out = []
    for c in rec.get('classes', []):
        key = c.get('learned_best' if which == 'learned' else 'exhaustive_best')
        if not key:
            return []
        k = tuner.parse_key(key)
        gopts = {'threads': k['threads']}
        if k.get('lanes'):
            gopts['lanes'] = 1
        out.append((c['max_area'], tuple(k['tiles']), k['group_size'], gopts,
                    '%s %dx%d: %s' % (which, c['rows'], c['cols'], key)))
    return out


def tuned(app, which='learned', **kw):
    """The workload ``app`` built with the schedule ``schedule(app, which)``:
    the pipeline bench.py times and the full-size parity tests check."""
    tiles, gs, gopts, label = schedule(app, which)
    args = dict(kw)
    args['tiles'] = tiles
    if gs is not None:
        args['group_size'] = gs
    if gopts:
        merged = dict(gopts)
        merged.update(args.get('gpu_options') or {})
        args['gpu_options'] = merged
    w = BUILDERS[app](**args)
    w.schedule = label
    return w
