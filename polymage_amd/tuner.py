"""Autotuning of tile sizes, group sizes and workgroup shape on the GPU.

Two layers:

* The reference interface (sandbox/tuner.py:100-631), kept as is:
  ``generate(dict)`` builds, emits and compiles one library per
  (tile size, group size) configuration into ``<dst>/Poly<rand>Mage/``
  with ``configurations.txt``, and returns ``(dst_dir, n_configs, pipe)``;
  ``execute(dict)`` loads every ``<app>_polymage_<k>.so``, calls
  ``pipeline_<name>`` ``_tuner_nruns`` times on the host arrays of
  ``_tuner_pipe_arg_data`` (or through ``_tuner_custom_executor``) and
  writes ``tuning_report_<date>.txt`` / ``plot_<date>.dat``.  Build,
  compile, load and run errors are logged as ``[ERROR]`` and skipped, as
  in the reference.

* The GPU-native search used for measurement: ``Space`` enumerates
  candidate schedules, ``compile_all`` compiles them in parallel worker
  processes (hiprtc + g++; no GPU needed), ``time_candidates`` times each
  with device-resident inputs and HIP events (median of several runs on a
  cold buffer set), and ``TimingDB`` appends one JSON record per
  measurement with the schedule's static features (``costmodel``), which
  is what the learned cost model trains on.
"""

import ctypes
import itertools
import json
import os
import random
import string
import sys
import time

from .compiler import buildPipeline
from . import runtime


# ---------------------------------------------------------------------------
# reference interface
# ---------------------------------------------------------------------------

def _print_to(msg, files, end='\n'):
    for f in files:
        f.write(str(msg) + end)
        f.flush()


def _line(files):
    _print_to('-' * 72, files)


def generate(arg):
    app = arg['_tuner_app_name']
    pipe_name = arg.get('_tuner_pipe_name')
    live_outs = arg['_tuner_live_outs']
    constraints = arg.get('_tuner_param_constraints') or []
    estimates = arg.get('_tuner_param_estimates') or []
    tiles = arg.get('_tuner_tile_size_configs') or [[16, 16, 16]]
    groups = arg.get('_tuner_group_size_configs')
    if not groups:          # the reference's default (it crashes instead, tuner.py:262-267)
        groups = [1, 2, 3, 200]
    opts = arg.get('_tuner_opts') or []
    gpu_opts = arg.get('_tuner_gpu_options') or None
    dst = arg.get('_tuner_dst_path', '/tmp')
    debug = arg.get('_tuner_debug_flag', False)
    dump = [sys.stdout] if debug else []

    def rnd():
        return ''.join(random.SystemRandom().choice(string.ascii_letters + string.digits)
                       for _ in range(10))
    sub = os.path.join(str(dst), 'Poly' + rnd() + 'Mage') + '/'
    while os.path.exists(sub):
        sub = os.path.join(str(dst), 'Poly' + rnd() + 'Mage') + '/'
    os.makedirs(sub)
    prefix = sub + app + '_polymage_'
    cfg_path = sub + 'configurations.txt'
    pipe = None
    k = 0
    t_all = time.time()
    with open(cfg_path, 'a') as cfg:
        files = dump + [cfg]
        _line(files)
        _print_to('App name                : "%s"' % app, files)
        _print_to('Code Path               : "%s"' % sub, files)
        _print_to('Live-out Functions      : ' + ' '.join('"%s",' % f.name for f in live_outs), files)
        _print_to('Tile Sizes              : %s' % (tiles,), files)
        _print_to('Group Sizes             : %s' % (groups,), files)
        for ts in tiles:
            for gs in groups:
                k += 1
                _line(files)
                _print_to('Config     : #%d' % k, files)
                _print_to('Tile Sizes : %s' % (ts,), files)
                _print_to('Group Size : %s' % (gs,), files)
                t1 = time.time()
                try:
                    pipe = buildPipeline(live_outs, param_constraints=constraints,
                                         param_estimates=estimates, tile_sizes=ts,
                                         group_size=gs, options=opts, pipe_name=pipe_name,
                                         gpu_options=gpu_opts)
                    code = pipe.generate_code()
                except Exception as e:
                    _print_to('[ERROR] Build fail ... (%s)' % (str(e)[:200],), files)
                    continue
                with open(prefix + '%d.hip' % k, 'w') as f:
                    f.write(str(code))
                t2 = time.time()
                _print_to('Code Generation Time   : %sms' % ((t2 - t1) * 1000), files)
                try:
                    bp = runtime.build(pipe)
                    so = prefix + '%d.so' % k
                    if os.path.lexists(so):
                        os.remove(so)
                    os.symlink(bp.so_path, so)
                except Exception as e:
                    _print_to('[ERROR] Compilation aborted ... (%s)' % (str(e)[:200],), files)
                    continue
                t3 = time.time()
                _print_to('Compilation Time       : %sms' % ((t3 - t2) * 1000), files)
                _print_to('Total                  : %sms' % ((t3 - t1) * 1000), files)
        _line(files)
        _print_to('Time taken to generate all variants : ', files)
        _print_to('%sms' % ((time.time() - t_all) * 1000), files)
        _line(files)
    return sub, k, pipe


def execute(arg):
    from .utils import get_ordered_cfunc_params, map_cfunc_args
    custom = arg.get('_tuner_custom_executor')
    arg_data = arg.get('_tuner_pipe_arg_data')
    app_data = arg.get('_tuner_app_data')
    app = arg['_tuner_app_name']
    pipe = arg['_tuner_pipe']
    src = arg.get('_tuner_src_path', '/tmp')
    n = arg.get('_tuner_configs_count', 0)
    nruns = arg.get('_tuner_nruns', 1)
    debug = arg.get('_tuner_debug_flag', True)
    dump = [sys.stdout] if debug else []
    prefix = os.path.join(str(src), app + '_polymage_')
    stamp = time.strftime("%d-%m-%Y_%H.%M.%S")
    report = os.path.join(str(src), 'tuning_report_%s.txt' % stamp)
    plot = os.path.join(str(src), 'plot_%s.dat' % stamp)
    fname = 'pipeline_' + pipe.name
    params = get_ordered_cfunc_params(pipe)
    args = map_cfunc_args(params, arg_data) if arg_data else None
    best_k, best_t = 0, None
    times = {}
    t0 = time.time()
    with open(report, 'a') as rep:
        files = dump + [rep]
        _line(files)
        _print_to('App Name              : "%s"' % app, files)
        _print_to('Total Configurations  :%d' % n, files)
        _print_to('Number of Tuning Runs :%d' % nruns, files)
        for k in range(1, n + 1):
            _line(files)
            _print_to('Config #%d : ' % k, files, ' ')
            try:
                lib = ctypes.cdll.LoadLibrary(prefix + '%d.so' % k)
                fn = lib[fname]
            except Exception:
                _print_to('[ERROR] in loading shared library ...', files)
                continue
            local = None
            err = False
            for _ in range(nruns):
                t1 = time.time()
                try:
                    if custom is None:
                        fn(*args)
                    else:
                        custom(lib, fn, params, args, arg, app_data)
                except Exception:
                    err = True
                dt = time.time() - t1
                local = dt if local is None else min(local, dt)
            if err:
                _print_to('[ERROR] Execution Aborted ...', files)
                continue
            times[k] = local
            if best_t is None or local < best_t:
                best_t, best_k = local, k
            _print_to('%s ms (%s ms)' % (local * 1000, best_t * 1000), files)
            with open(plot, 'a') as p:
                p.write('%d %s\n' % (k, local))
        _line(files)
        _print_to('Best Config :', files)
        _print_to('#%d : %s ms' % (best_k, (best_t or 0) * 1000), files)
        _print_to('Tuning time : %s s' % (time.time() - t0), files)
    return {'best_config': best_k, 'best_time': best_t, 'times': times, 'report': report}


# ---------------------------------------------------------------------------
# GPU-native search
# ---------------------------------------------------------------------------

class Candidate(object):
    def __init__(self, tiles, group_size, threads=256, lanes=False):
        self.tiles = list(tiles)
        self.group_size = group_size
        self.threads = threads
        self.lanes = bool(lanes)

    @property
    def key(self):
        return 'T%s_g%d_w%d%s' % ('x'.join(map(str, self.tiles)), self.group_size, self.threads,
                                  '_L' if self.lanes else '')

    @property
    def gpu_options(self):
        o = {'threads': self.threads}
        if self.lanes:
            o['lanes'] = 1
        return o

    def __repr__(self):
        return 'Candidate(%s)' % self.key


class Space(object):
    """Cartesian product of tile sizes, group sizes, threads and (for
    applications with a lane-fusable group) lane fusion on or off."""

    def __init__(self, tiles, group_sizes, threads=(256,), lanes=(False,)):
        self.tiles = list(tiles)
        self.group_sizes = list(group_sizes)
        self.threads = tuple(threads)
        self.candidates = [Candidate(t, g, w, l) for t, g, w, l in
                           itertools.product(tiles, group_sizes, threads, lanes)]

    def __iter__(self):
        return iter(self.candidates)

    def __len__(self):
        return len(self.candidates)


POW2_TILES = [(a, b) for a in (8, 16, 32, 64, 128, 256, 512) for b in (8, 16, 32, 64, 128, 256, 512)]


_HALO = {}


def row_halo(app):
    """Points the widest materialised stage of the application's main
    group adds to the tile along the innermost tiled dimension (unsharp's
    vertical blur is computed on 256 + 4 columns of a 256-wide tile)."""
    if app not in _HALO:
        from .apps import workloads
        from .hip_emit import plan_pipeline
        from .group_plan import GroupPlan
        from .analysis import PlanError
        halo = 0
        plans = []
        for t in ((32, 256), (16, 64)):
            try:
                w = getattr(workloads, app)(tiles=t, with_data=False)
                plans = [p for p in plan_pipeline(w.pipe)
                         if isinstance(p, GroupPlan) and p.grid_dims]
                break
            except PlanError:
                continue
        if plans:
            p = max(plans, key=lambda q: len(q.stages))
            d, T = p.grid_dims[-1]
            for st in p.stages:
                if st.inline:
                    continue
                for i, al in enumerate(st.align):
                    if al is not None and al[0] == d and al[1] == 1:
                        halo = max(halo, st.static_ext[i] - T)
        _HALO[app] = halo
    return _HALO[app]


def aligned_tiles(app):
    """The power-of-two tiles with the innermost tile size reduced by the
    application's halo, so the widest stage region is a power of two wide
    and its sweep fills whole waves (unsharp: 16x252)."""
    h = row_halo(app)
    if h <= 0:
        return []
    return [(a, b - h) for (a, b) in POW2_TILES if b >= 32 and b - h > h]


_LANE_REACH = {}


def lane_tiles(app):
    """Tiles whose innermost size is a whole number of lane-fused wavefront
    steps (64 - 2 * column reach: unsharp 60, Harris 62), or [] when the
    application has no lane-fusable group."""
    if app not in _LANE_REACH:
        from .apps import workloads
        from .hip_emit import plan_pipeline
        from .group_plan import GroupPlan
        w = workloads.BUILDERS[app](with_data=False, gpu_options={'lanes': 1})
        reach = None
        for p in plan_pipeline(w.pipe):
            if isinstance(p, GroupPlan) and getattr(p, 'lane_pair', None) is not None:
                reach = p.lane_pair[4]
        _LANE_REACH[app] = reach
    r = _LANE_REACH[app]
    if r is None:
        return []
    step = 64 - 2 * r
    return [(a, k * step) for a in (8, 16, 32, 64, 128) for k in (1, 2, 4, 8)]


def default_space(app):
    """Tile sizes {8..512}^2 (powers of two, as the benchmark configuration
    states) and the same with the innermost size less the application's
    halo (``aligned_tiles``), times the application's group sizes times
    workgroup sizes.  Schedules that do not fit (LDS) fail at planning and
    are skipped.  Lane fusion (``lane_space``) is left out: on unsharp and
    Harris its best candidates measured slower than the LDS schedules."""
    tiles = POW2_TILES + aligned_tiles(app)
    if app == 'unsharp':
        return Space(tiles, [1, 5], (128, 256, 512))
    if app == 'harris':
        return Space(tiles, [1, 5, 11], (256, 512))
    if app == 'bilateral':
        return Space(tiles, [1, 5], (128, 256, 512))
    if app == 'laplacian':
        return Space(tiles, [1, 3, 5, 8], (256,))
    if app == 'campipe':
        return Space(tiles, [3, 5, 12, 40], (256,))
    if app == 'interpolate':
        return Space(tiles, [1, 5, 8], (256,))
    if app == 'blend':
        return Space(tiles, [1, 3, 5, 6], (256,))
    raise KeyError(app)


def lane_space(app):
    """Lane-fused candidates (gpu option ``lanes``) of an application whose
    main group has a lane-fusable pair of sweeps: the default tiles plus
    ``lane_tiles``, with lane fusion on."""
    base = default_space(app)
    return Space(list(base.tiles) + lane_tiles(app), list(base.group_sizes),
                 tuple(base.threads), lanes=(True,))


def build_candidate(workload_fn, cand, with_data=False):
    """Build the pipeline of ``workload_fn`` for one candidate."""
    return workload_fn(tiles=tuple(cand.tiles), group_size=cand.group_size,
                       gpu_options=cand.gpu_options, with_data=with_data)


def _worker_init(cache):
    runtime.CACHE = cache


def make_pool(procs):
    """A pool of fresh worker interpreters ('spawn': they never touch the
    GPU, and the parent may already have initialised it), reusable across
    compile_all / features_all calls."""
    import multiprocessing as mp
    return mp.get_context('spawn').Pool(procs, initializer=_worker_init, initargs=(runtime.CACHE,))


def _features_one(job):
    modname, fname, kw, cand_tuple, cache = job
    runtime.CACHE = cache
    import importlib
    mod = importlib.import_module(modname)
    fn = getattr(mod, fname)
    cand = Candidate(*cand_tuple)
    try:
        w = build_candidate(lambda **k: fn(**dict(kw, **k)), cand)
        from .costmodel import features
        return cand.key, features(w.pipe), None
    except Exception as e:
        return cand.key, None, str(e)[:300]


def features_all(workload, candidates, procs=8, pool=None, **kw):
    """Static features of every candidate (planning only, no compilation),
    in parallel worker processes; {key: features or None}."""
    jobs = [(workload.__module__, workload.__name__, kw,
             (c.tiles, c.group_size, c.threads), runtime.CACHE) for c in candidates]
    out = {}
    own = pool is None and procs > 1
    if own:
        pool = make_pool(procs)
    res = pool.imap_unordered(_features_one, jobs, chunksize=4) if pool else map(_features_one, jobs)
    for key, feats, err in res:
        out[key] = feats
    if own:
        pool.close()
        pool.join()
    return out


def _compile_one(job):
    modname, fname, kw, cand_tuple, cache = job
    runtime.CACHE = cache
    import importlib
    mod = importlib.import_module(modname)
    fn = getattr(mod, fname)
    cand = Candidate(*cand_tuple)
    t0 = time.time()
    try:
        w = build_candidate(lambda **k: fn(**dict(kw, **k)), cand)
        bp_path = runtime.build(w.pipe).so_path
        from .costmodel import features
        feats = features(w.pipe)
        return cand.key, bp_path, feats, time.time() - t0, None
    except Exception as e:
        return cand.key, None, None, time.time() - t0, str(e)[:300]


def compile_all(workload, candidates, procs=8, verbose=False, pool=None, **kw):
    """Compile candidates in parallel; returns {key: (so_path, features,
    seconds, error)}.  ``workload`` is a function of apps.workloads."""
    jobs = [(workload.__module__, workload.__name__, kw,
             (c.tiles, c.group_size, c.threads), runtime.CACHE) for c in candidates]
    out = {}
    own = pool is None and procs > 1
    if own:
        pool = make_pool(procs)
    res = pool.imap_unordered(_compile_one, jobs) if pool else map(_compile_one, jobs)
    for key, so, feats, dt, err in res:
        out[key] = (so, feats, dt, err)
        if verbose:
            print('  compiled %-22s %6.1fs%s' % (key, dt, ('  ERROR ' + err[:80]) if err else ''),
                  flush=True)
    if own:
        pool.close()
        pool.join()
    return out


class TimingDB(object):
    """Append-only JSON-lines store of measurements."""

    def __init__(self, path):
        self.path = path

    def add(self, rec):
        d = os.path.dirname(self.path)
        if d:
            os.makedirs(d, exist_ok=True)
        with open(self.path, 'a') as f:
            f.write(json.dumps(rec) + '\n')

    def load(self):
        if not os.path.exists(self.path):
            return []
        with open(self.path) as f:
            return [json.loads(l) for l in f if l.strip()]


def time_pipeline(bp, w, runs=9, warmup=2, nsets=None, reps=3):
    """Milliseconds per pipeline call with device-resident data: ``reps``
    batches of ``runs`` back-to-back calls, each batch between two HIP
    events on the launch stream (as bench.py times), the median batch
    average.  Calls rotate over enough buffer sets (about 1 GiB) that no
    call finds its data in the 256 MB Infinity Cache."""
    import numpy as np
    import torch
    dev = torch.device('cuda:0')
    if nsets is None:
        per = sum(w.values[f.name].nbytes for f in bp.inputs) + \
            sum(int(np.prod(w.out_shape)) * f.typ.nbytes() for f in bp.outputs)
        nsets = max(2, min(10, int((1 << 30) // max(1, per)) + 1))
    sets = []
    for _ in range(nsets):
        d = {}
        for f in bp.inputs:
            a = w.values[f.name]
            d[f.name] = torch.from_numpy(a.view(np.int16) if a.dtype == np.uint16 else a).to(dev)
        for f in bp.outputs:
            nb = int(np.prod(w.out_shape)) * f.typ.nbytes()
            d[f.name] = torch.empty(nb, dtype=torch.uint8, device=dev)
        sets.append(d)
    scal = dict((k, v) for k, v in w.values.items() if not hasattr(v, 'shape'))
    stream = torch.cuda.current_stream().cuda_stream
    launches = [bp.bind_dev(dict(scal, **d), stream) for d in sets]
    for i in range(warmup):
        launches[i % nsets]()
    torch.cuda.synchronize()
    ts = []
    k = 0
    for _ in range(reps):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for i in range(runs):
            launches[k % nsets]()
            k += 1
        e1.record()
        torch.cuda.synchronize()
        ts.append(e0.elapsed_time(e1) / runs)
    ts.sort()
    return ts[len(ts) // 2]


_DATA = {}


def workload_data(workload, **kw):
    """The workload with its synthetic input, generated once per process
    (input generation is not part of either search's cost)."""
    dk = (workload.__module__, workload.__name__, repr(sorted(kw.items())))
    if dk not in _DATA:
        _DATA[dk] = workload(with_data=True, **kw)
    return _DATA[dk]


def time_candidates(workload, compiled, runs=9, db=None, app=None, verbose=False, **kw):
    """Time compiled candidates on the GPU; returns {key: ms}."""
    w = workload_data(workload, **kw)
    out = {}
    for key, (so, feats, dt, err) in compiled.items():
        if so is None:
            continue
        bp = runtime.load(so, w.pipe)
        try:
            ms = time_pipeline(bp, w, runs=runs)
        except Exception as e:
            if db is not None:
                db.add({'app': app, 'key': key, 'error': str(e)[:200]})
            continue
        out[key] = ms
        if verbose:
            print('  timed    %-22s %8.4f ms' % (key, ms), flush=True)
        if db is not None:
            db.add({'app': app, 'key': key, 'ms': ms, 'mpix': w.mpix, 'features': feats,
                    'compile_s': dt})
    return out


def time_batch(bp, workload, frames, frame_input='img', runs=5, warmup=1, reps=3):
    """Milliseconds per persistent work-queue launch of ``bp`` over a batch
    of frames (``frames``: input arrays of the workload's ``frame_input``,
    of any sizes; the workload function gives every other value and the
    output shape at each frame's size), device-resident; median of ``reps``
    batches of ``runs`` launches between HIP events."""
    import numpy as np
    import torch
    dev = torch.device('cuda:0')
    recs = []
    for img in frames:
        R, C = img.shape[:2]
        w = workload(rows=R, cols=C, with_data=False)
        rec = dict((k, v) for k, v in w.values.items() if not hasattr(v, 'shape'))
        for f in bp.inputs:
            a = img if f.name == frame_input else w.values[f.name]
            rec[f.name] = torch.from_numpy(a.view(np.int16) if a.dtype == np.uint16 else a).to(dev)
        for f in bp.outputs:
            rec[f.name] = torch.empty(int(np.prod(w.out_shape)) * f.typ.nbytes(),
                                      dtype=torch.uint8, device=dev)
        recs.append(rec)
    h = bp.batch_create(recs)
    try:
        stream = torch.cuda.current_stream().cuda_stream
        for _ in range(warmup):
            bp.batch_run(h, stream, 'persistent')
        torch.cuda.synchronize()
        ts = []
        for _ in range(reps):
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(runs):
                bp.batch_run(h, stream, 'persistent')
            e1.record()
            torch.cuda.synchronize()
            ts.append(e0.elapsed_time(e1) / runs)
    finally:
        bp.batch_destroy(h)
    ts.sort()
    return ts[len(ts) // 2]


def time_candidates_batch(workload, compiled, frames, runs=5, db=None, app=None, verbose=False):
    """time_candidates with a batch of frames as the objective (the
    persistent launch over all of them, ``time_batch``); returns {key: ms}."""
    w = workload(rows=frames[0].shape[0], cols=frames[0].shape[1], with_data=False)
    mpix = 0.0
    out = {}
    for key, (so, feats, dt, err) in compiled.items():
        if so is None:
            continue
        bp = runtime.load(so, w.pipe)
        try:
            ms = time_batch(bp, workload, frames, runs=runs)
        except Exception as e:
            if db is not None:
                db.add({'app': app, 'key': key, 'error': str(e)[:200]})
            continue
        out[key] = ms
        if verbose:
            print('  timed    %-22s %8.4f ms (batch of %d)' % (key, ms, len(frames)), flush=True)
        if db is not None:
            db.add({'app': app, 'key': key, 'ms': ms, 'mpix': mpix, 'features': feats,
                    'compile_s': dt, 'batch': len(frames)})
    return out


# ---------------------------------------------------------------------------
# recorded schedules
# ---------------------------------------------------------------------------

TUNED_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'tuned.json')


def parse_key(key):
    """'T32x256_g5_w256' -> {'tiles': (32, 256), 'group_size': 5, 'threads': 256,
    'lanes': False}; a trailing '_L' marks lane fusion."""
    parts = key.split('_')
    t, g, w = parts[:3]
    return {'tiles': tuple(int(x) for x in t[1:].split('x')), 'group_size': int(g[1:]),
            'threads': int(w[1:]), 'lanes': parts[3:] == ['L'], 'key': key}


def tuned_schedule(app, which='learned', path=None):
    """The schedule the autotuner recorded for ``app`` ('learned': the cost
    model's best timed top-k candidate; 'exhaustive': the exhaustive best),
    or None."""
    path = path or TUNED_PATH
    if not os.path.exists(path):
        return None
    with open(path) as f:
        rec = json.load(f).get(app)
    if not rec:
        return None
    key = rec.get('learned_best' if which == 'learned' else 'exhaustive_best')
    return parse_key(key) if key else None
