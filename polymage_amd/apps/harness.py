"""The reference applications' harness, for every pipeline here.

Reference: ``sandbox/apps/python/img_proc/<app>/`` — ``main.py`` (mode
dispatch), ``builder.py:86-113`` (``create_lib``), ``arg_parser.py:10-81``
and ``apps/python/pipe_options.py`` (command line), ``exec_pipe.py``
(``call_pipe`` and the timed run loop) and ``app_tuner.py`` (``auto_tune``
through ``tuner.generate`` / ``tuner.execute``).

Modes, as in the reference:

* ``new``: build the pipeline (buildPipeline), optionally write its graph,
  write the generated source to ``<dir>/<app>.cpp`` (here the HIP kernels
  followed by the host stub), compile it to ``<dir>/<app>.so`` and load it;
* ``existing``: compile the source already in ``<dir>/<app>.cpp`` (no
  pipeline is built) and load it;
* ``ready``: load ``<dir>/<app>.so`` as it is;
* ``tune``: generate one library per tile-size x group-size configuration
  and time each (``tuner.generate`` + ``tuner.execute``).

The compiled source and library stay on disk between runs, which is what
``existing`` and ``ready`` reuse.  Inputs are the seeded synthetic images
of ``apps/inputs.py`` at the requested size (the reference reads an image
file through PIL, which this environment does not have), or a ``.npy``
file given with ``--img``.

Usage: ``python -m polymage_amd.apps.harness <app> [--mode new|existing|ready|tune]
[--rows N --cols N] [--runs N] [--timer] [--tiles 32x256] [--group-size G]
[--dir PATH] [--graph-gen]``.
"""

import ctypes
import optparse
import os
import sys
import time

import numpy as np

from . import workloads
from ..compat.cpp_compiler import c_compile
from ..compat.loader import load_lib

MODES = ('new', 'existing', 'ready', 'tune')

# default image sizes (rows, cols) of each application's input
DEFAULT_SIZE = {'harris': (2048, 2048), 'unsharp': (2848, 4256), 'bilateral': (2560, 1536),
                'laplacian': (2560, 1536), 'campipe': (1216, 1664), 'interpolate': (2560, 1536),
                'blend': (2560, 1536)}

# the reference apps' tuning grid (harris/app_tuner.py:33-54)
TUNE_TILES = [[64, 256], [64, 128], [32, 512], [32, 256], [32, 128], [32, 64], [16, 512],
              [16, 256], [16, 128], [16, 64], [8, 512], [8, 256], [8, 128], [8, 64], [8, 32]]
TUNE_GROUPS = [3, 5, 7, 9, 10]


def parse_args(argv=None):
    help_str = '"new" : from scratch | "existing" : compile and run |  "ready" : just run | "tune"'
    p = optparse.OptionParser(usage='%prog <app> [options]')
    p.add_option('-m', '--mode', type='choice', choices=list(MODES), default='new', help=help_str)
    p.add_option('-i', '--img', dest='img_file', default=None,
                 help='input image (.npy) for single-input apps; default: seeded synthetic input')
    p.add_option('-x', '--rows', type='int', default=0, help='number of rows of the image')
    p.add_option('-y', '--cols', type='int', default=0, help='number of cols of the image')
    p.add_option('-n', '--runs', type='int', default=1, help='number of runs')
    p.add_option('-t', '--timer', action='store_true', default=False, help='report execution time')
    p.add_option('--graph-gen', action='store_true', dest='graph_gen', default=False,
                 help='write the pipeline graph as <dir>/<app>.dot')
    p.add_option('--tiles', default=None, help='tile sizes, e.g. 32x256 (default: the app default)')
    p.add_option('--group-size', type='int', dest='group_size', default=None)
    p.add_option('--dir', default='.', help='where <app>.cpp and <app>.so are kept')
    p.add_option('--tune-dir', dest='tune_dir', default='/tmp', help="tuner's _tuner_dst_path")
    # the reference's storage options (pipe_options.py), accepted for the
    # same command lines: the GPU plan decides storage itself
    for opt in ('optimize_storage', 'early_free', 'pool_alloc', 'flatten_scratchpad', 'multipar'):
        p.add_option('--' + opt, action='store_true', dest=opt, default=False)
    opts, args = p.parse_args(argv)
    if len(args) != 1 or args[0] not in workloads.BUILDERS:
        p.error('give one application: %s' % ', '.join(sorted(workloads.BUILDERS)))
    opts.app = args[0]
    return opts


def init_all(app_args):
    """app_data of the reference harness: sizes, pipe data, host images."""
    app = app_args.app
    rows, cols = DEFAULT_SIZE[app]
    rows = app_args.rows or rows
    cols = app_args.cols or cols
    kw = {}
    if app_args.tiles:
        kw['tiles'] = tuple(int(t) for t in app_args.tiles.split('x'))
    if app_args.group_size:
        kw['group_size'] = app_args.group_size
    w = workloads.BUILDERS[app](rows=rows, cols=cols, **kw)
    if app_args.img_file:
        img = np.load(app_args.img_file, allow_pickle=False)
        names = [k for k, v in w.values.items() if hasattr(v, 'shape')]
        if len(names) != 1:
            raise ValueError('--img is for single-input applications')
        want = w.values[names[0]]
        if img.shape != want.shape:
            raise ValueError('--img: shape %s, this app at this size takes %s' % (img.shape, want.shape))
        w.values[names[0]] = np.ascontiguousarray(img.astype(want.dtype))
    opts = []
    for o in ('early_free', 'optimize_storage', 'pool_alloc'):
        if getattr(app_args, o):
            opts.append(o)
    out = np.zeros(w.out_shape, w.out_func.typ.np_dtype())
    return {'app': app, 'mode': app_args.mode, 'runs': app_args.runs, 'timer': app_args.timer,
            'graph_gen': app_args.graph_gen, 'app_args': app_args, 'rows': rows, 'cols': cols,
            'workload': w, 'opts': opts, 'dir': app_args.dir,
            'img_data': {'values': dict(w.values), 'OUT': out}}


def codegen(pipe, file_name, app_data):
    code = pipe.generate_code(is_extern_c_func=True, are_io_void_ptrs=True)
    with open(file_name, 'w') as f:
        f.write(str(code))


def generate_graph(pipe, file_name, app_data):
    pipe.pipeline_graph.write(file_name + '.dot')


def build_app(app_data):
    return app_data['workload'].pipe


def create_lib(build_func, pipe_name, app_data):
    """builder.py:86-113: new -> build + codegen; every mode but ready ->
    compile; then load pipeline_<pipe_name> into app_data."""
    mode = app_data['mode']
    base = os.path.join(app_data.get('dir', '.'), pipe_name)
    pipe_src = base + '.cpp'
    pipe_so = base + '.so'
    if build_func is not None and mode == 'new':
        pipe = build_func(app_data)
        if app_data.get('graph_gen'):
            generate_graph(pipe, base, app_data)
        codegen(pipe, pipe_src, app_data)
    if mode != 'ready':
        c_compile(pipe_src, pipe_so, app_data)
    load_lib(pipe_so, 'pipeline_' + pipe_name, app_data)


def ordered_args(values, out_name, out):
    """The ctypes arguments of pipeline_<name> (codegen.py:957-999):
    parameters sorted by name, input images sorted by name, then the
    output; scalars are int or float parameters."""
    scal = sorted((k, v) for k, v in values.items() if not hasattr(v, 'shape'))
    imgs = sorted((k, v) for k, v in values.items() if hasattr(v, 'shape'))
    args = []
    for _, v in scal:
        args.append(ctypes.c_int(v) if isinstance(v, (int, np.integer)) else ctypes.c_float(v))
    for _, v in imgs:
        args.append(ctypes.c_void_p(v.ctypes.data))
    args.append(ctypes.c_void_p(out.ctypes.data))
    return args


def call_pipe(app_data):
    w = app_data['workload']
    d = app_data['img_data']
    fn = app_data['pipeline_' + app_data['app']]
    fn(*ordered_args(d['values'], w.out_func.name, d['OUT']))


def exec_pipe(app_data):
    """exec_pipe.py: ``runs`` calls, the mean time printed with --timer."""
    t1 = time.time()
    for _ in range(app_data['runs']):
        call_pipe(app_data)
    dt = time.time() - t1
    app_data['time_ms'] = dt * 1000 / max(1, app_data['runs'])
    if app_data['timer']:
        print('[exec_pipe] : time taken to execute = ', app_data['time_ms'], ' ms')


def auto_tune(app_data):
    """app_tuner.py: every (tile, group size) configuration generated and
    timed on the application's input."""
    from .. import tuner
    from ..constructs import Condition
    w = app_data['workload']
    pipe = w.pipe
    pd_params = dict((p.name, p) for p in pipe.get_parameters())
    vals = app_data['img_data']['values']
    est = [(pd_params[k], vals[k]) for k in pd_params if isinstance(vals.get(k), int)]
    cons = [Condition(p, '==', v) for p, v in est]
    app_args = app_data['app_args']
    tiles = [[int(t) for t in app_args.tiles.split('x')]] if app_args.tiles else TUNE_TILES
    groups = [app_args.group_size] if app_args.group_size else TUNE_GROUPS
    gen = {'_tuner_app_name': app_data['app'], '_tuner_pipe_name': app_data['app'],
           '_tuner_live_outs': [w.out_func], '_tuner_param_constraints': cons,
           '_tuner_param_estimates': est, '_tuner_tile_size_configs': tiles,
           '_tuner_group_size_configs': groups, '_tuner_opts': app_data['opts'],
           '_tuner_dst_path': app_args.tune_dir, '_tuner_debug_flag': False}
    src, n, tpipe = tuner.generate(gen)
    arg_data = dict(vals)
    arg_data[w.out_func.name] = app_data['img_data']['OUT']
    ex = {'_tuner_app_name': app_data['app'], '_tuner_pipe': tpipe, '_tuner_src_path': src,
          '_tuner_configs_count': n, '_tuner_nruns': max(1, app_data['runs']),
          '_tuner_pipe_arg_data': arg_data, '_tuner_debug_flag': True}
    return tuner.execute(ex)


def main(argv=None):
    app_args = parse_args(argv)
    app_data = init_all(app_args)
    if app_data['mode'] == 'tune':
        return auto_tune(app_data)
    create_lib(build_app, app_data['app'], app_data)
    exec_pipe(app_data)
    return app_data


if __name__ == '__main__':
    main(sys.argv[1:])
