#!/usr/bin/env python3
"""Benchmark: MPix/s of PolyMage pipelines on one MI355X.

Headline workload (``value``): the unsharp mask on an RGB float32
3x2848x4256 smooth random field (threshold 0.001, weight 3), one fused
group, with the schedule the learned autotuner recorded for this GPU.  A
step is one pass of the pipeline over the whole image with input and
output resident in HBM; buffer sets are rotated so each step reads cold
data.  After the timed region the last output is copied back and compared
bit for bit with the oracle's output on the same input, and the pipeline's
LDS-overflow status word is read; a mismatch makes the run fail.

The same JSON line carries ``pipelines``: every benchmark configuration
(Harris 2048^2, unsharp, bilateral grid 2560x1536, local Laplacian
2560x1536x3 uint16, the 48-frame camera batch) under the learned pick and
the exhaustive best the autotuner recorded, each timed the same way and
parity-checked the same way, with its roofline fraction, measured HBM
traffic (when the profiled kernel source matches), the tuning-time ratio
the autotuner recorded, and the CPU baseline on the host cores (all cores
the process may use, and one thread).

``roofline`` divides the ideal traffic (input bytes + output bytes) by the
time of one pipeline call: HIP events on the stream the kernels run on,
around back-to-back calls.  ``cpu_baseline`` times the C++/OpenMP backend
(the reference's code generator restated: the same pipeline and schedule
as OpenMP-parallel overlapped tiles) on the host cores over a bounded
sample; for Harris, the reference's own generated refHarris when it is
built.  The oracle is used only as the parity checker.
"""

import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = 'MPix/s per pipeline, 1 MI355X: learned vs exhaustive schedule; % HBM roofline'
CONFIG_APPS = ('harris', 'unsharp', 'bilateral', 'laplacian')


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=20)
    ap.add_argument('--warmup', type=int, default=5)
    ap.add_argument('--workload', default='unsharp',
                    choices=['unsharp', 'harris', 'bilateral', 'laplacian', 'campipe48',
                             'interpolate', 'blend'])
    ap.add_argument('--tiles', default=None, help='e.g. 32x256')
    ap.add_argument('--threads', type=int, default=None)
    ap.add_argument('--gpu-opt', action='append', default=[], metavar='KEY=VALUE',
                    help='extra code-generation option (e.g. no_stage=1, block=8,1)')
    ap.add_argument('--schedule', default='learned', choices=['learned', 'exhaustive', 'default'],
                    help='schedule recorded by the autotuner for this workload')
    ap.add_argument('--literal-mode', default='double', choices=['double', 'float12'],
                    help="float literals as C doubles (the reference's code generator) or as "
                         "12-digit float literals (its checked-in generated files)")
    ap.add_argument('--pipelines', default='all', choices=['all', 'none'],
                    help="also time every benchmark configuration under both recorded "
                         "schedules ('pipelines' in the JSON line)")
    ap.add_argument('--sustained-steps', type=int, default=100,
                    help='calls of the headline pipeline timed back to back for the sustained figure (0: skip)')
    ap.add_argument('--no-cpu', action='store_true', help='skip the CPU baselines')
    ap.add_argument('--no-size-classes', action='store_true',
                    help='campipe48: skip the size-classed batch (one launch per class)')
    ap.add_argument('--settle-s', type=float, default=None,
                    help='untimed launches before each run\'s first timing (default %.2f s)'
                    % DeviceRun.SETTLE_S)
    ap.add_argument('--cpu-seconds', type=float, default=4.0,
                    help='time budget of each all-core CPU baseline sample')
    ap.add_argument('--traffic', default=os.path.join(ROOT, 'profiles', 'traffic.json'),
                    help='JSON file with the HBM bytes per dispatch measured by rocprofv3 PMC runs')
    ap.add_argument('--copy-probe', type=int, default=0, metavar='MIB',
                    help='instead of the pipeline, run a plain device copy of MIB MiB '
                         '(calibrates the FETCH_SIZE / WRITE_SIZE counters)')
    return ap.parse_args()


def log(msg):
    print('[bench] ' + msg, file=sys.stderr, flush=True)


def copy_probe(mib, steps):
    """A plain device-to-device copy of known size, for scaling the PMC
    byte counters: FETCH_SIZE and WRITE_SIZE of one dispatch should read
    ``mib`` MiB each."""
    import torch
    n = mib * (1 << 20) // 4
    src = torch.rand(n, device='cuda:0')
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(steps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({'copy_probe_mib': mib, 'dispatches': steps}))


def device_copy_gbs(torch, dev, mib=512, reps=10):
    """Measured device-to-device copy bandwidth (read + write bytes per
    second) on this GPU: the practical ceiling next to the 8 TB/s spec."""
    n = mib * (1 << 20) // 4
    src = torch.rand(n, device=dev)
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del src, dst
    return round(2 * n * 4 / (ms / 1e3) / 1e9, 1)


# ---------------------------------------------------------------------------
# CPU side: host description, oracle outputs, CPU baselines
# ---------------------------------------------------------------------------

def host_cpu():
    model = None
    try:
        with open('/proc/cpuinfo') as f:
            for line in f:
                if line.startswith('model name'):
                    model = line.split(':', 1)[1].strip()
                    break
    except OSError:
        pass
    avail = len(os.sched_getaffinity(0)) if hasattr(os, 'sched_getaffinity') else os.cpu_count()
    threads = int(os.environ.get('OMP_NUM_THREADS') or avail)
    return {'model': model, 'logical_cpus': os.cpu_count(), 'cpus_allowed': avail,
            'omp_threads': min(threads, avail)}


def _omp():
    for name in ('libgomp.so.1', 'libgomp.so'):
        try:
            return ctypes.CDLL(name)
        except OSError:
            pass
    return None


def _time_cpu(fn, budget_s, max_runs=200, warm=True):
    """Min wall time of up to max_runs calls of fn within budget_s seconds
    (one warm-up call first unless ``warm`` is False)."""
    if warm:
        fn()
    best = None
    runs = 0
    t_start = time.perf_counter()
    while runs < max_runs and (runs == 0 or time.perf_counter() - t_start < budget_s):
        t0 = time.perf_counter()
        fn()
        dt = time.perf_counter() - t0
        best = dt if best is None else min(best, dt)
        runs += 1
    return best, runs, time.perf_counter() - t_start


def new_output(w):
    import numpy as np
    dt = w.out_func.typ.np_dtype()
    return np.full(w.out_shape, -7.0 if np.dtype(dt).kind == 'f' else 7, dt)


def oracle_output(w):
    """The oracle's output on the workload's input (the parity reference),
    from an output array pre-filled like the GPU's.  It is built with the
    oracle's parity flags (no floating-point contraction: the C semantics
    of the generated code); the CPU baseline's build, with the reference's
    -march=native, may contract a*b+c into fused multiply-adds."""
    from oracle import cemit
    o = new_output(w)
    mode = w.pipe.literal_mode
    cemit.build([w.out_func], 'bench_parity_%s_%s' % (w.name, mode), literal_mode=mode)(
        dict(w.values, **{w.out_func.name: o}))
    return o


def cpu_baseline(w, budget_s, cpu):
    """The CPU path on the host cores, on all cores the process may use
    (OMP_NUM_THREADS as set, at most the allowed CPUs) and on one thread:
    for Harris at its benchmark size the reference's own generated
    refHarris (oracle/_ref, built with the reference's compile flags);
    otherwise the oracle's C++/OpenMP port of the same pipeline (reference
    compile flags)."""
    import numpy as np
    omp = _omp()
    threads = cpu['omp_threads']
    ref_so = os.path.join(ROOT, 'oracle', '_ref', 'libref_harris_bench.so')
    if w.name == 'harris' and os.path.exists(ref_so):
        lib = ctypes.CDLL(ref_so)
        R, C = w.values['R'], w.values['C']
        img = w.values['img']
        out = np.zeros((R + 2, C + 2), np.float32)
        args = (ctypes.c_int(C), ctypes.c_int(R), ctypes.c_void_p(img.ctypes.data),
                ctypes.c_void_p(out.ctypes.data))
        fn = lambda: lib.ref_harris(*args)
        kind = 'reference'
        what = ("the reference's generated refHarris (ipp_bench/harris/harris_function.cpp, 32x256 "
                "tiles, float literals; g++ -O3 -fopenmp -march=x86-64-v3 -ftree-vectorize)")
    else:
        from polymage_amd import runtime
        from polymage_amd.cpu_emit import CPU_FLAGS
        cp = runtime.build_cpu(w.pipe)
        vals = dict(w.values)
        vals[w.out_func.name] = new_output(w)
        fn = lambda: cp.run_host(vals)
        kind = 'port'
        what = ("the C++/OpenMP backend's code for the same pipeline and schedule (the "
                "reference's code generator restated: OpenMP loop over overlapped tiles, "
                "per-thread scratch; g++ %s)" % ' '.join(f for f in CPU_FLAGS if f.startswith(('-O', '-f', '-m'))
                                                        and f not in ('-fPIC',)))
    if omp is not None:
        omp.omp_set_num_threads(threads)
    best, runs, spent = _time_cpu(fn, budget_s)
    res = {'value': round(w.mpix / best, 3), 'unit': 'MPix/s', 'cores': threads, 'kind': kind,
           'sample': '%d full-size runs (%.1f s) of %s, min time, %d OpenMP threads (OMP_PROC_BIND=%s)'
                     % (runs, spent, what, threads, os.environ.get('OMP_PROC_BIND', 'unset'))}
    if omp is not None:
        omp.omp_set_num_threads(1)
        b1, r1, s1 = _time_cpu(fn, 0.5 * budget_s, max_runs=3, warm=False)
        omp.omp_set_num_threads(threads)
        res['value_1thread'] = round(w.mpix / b1, 3)
        res['sample_1thread'] = '%d full-size runs (%.1f s), min time' % (r1, s1)
    return res


# ---------------------------------------------------------------------------
# GPU side
# ---------------------------------------------------------------------------

class DeviceRun(object):
    """A built workload with its inputs resident in HBM in several buffer
    sets (so the 256 MB Infinity Cache holds none of a step's data) and
    pre-marshalled launches, one per set."""

    def __init__(self, torch, dev, w, bp):
        import numpy as np
        self.w = w
        self.bp = bp
        self.torch = torch
        ins = list(bp.inputs)
        out = bp.outputs[0]
        self.out = out
        in_bytes = sum(w.values[f.name].nbytes for f in ins)
        self.out_bytes = int(np.prod(w.out_shape)) * out.typ.nbytes()
        self.nsets = max(2, min(10, int((1 << 30) // (in_bytes + self.out_bytes)) + 1))
        self.sets = []
        for _ in range(self.nsets):
            d = {}
            for f in ins:
                a = w.values[f.name]
                d[f.name] = torch.from_numpy(a.view(np.int16) if a.dtype == np.uint16 else a).to(dev)
            d[out.name] = torch.empty(self.out_bytes, dtype=torch.uint8, device=dev)
            self.sets.append(d)
        scal = dict((k, v) for k, v in w.values.items() if not hasattr(v, 'shape'))
        self.stream = torch.cuda.current_stream().cuda_stream
        self.launches = []
        for d in self.sets:
            vals = dict(scal)
            vals.update(d)
            self.launches.append(bp.bind_dev(vals, self.stream))
        self.last = 0

    def fill_outputs(self):
        """Output buffers pre-filled like the oracle's (-7.0 / 7), so points
        no Case covers compare too."""
        import numpy as np
        pat = new_output(self.w).reshape(-1).view(np.uint8)
        t = self.torch.from_numpy(pat.copy()).to(self.sets[0][self.out.name].device)
        for d in self.sets:
            d[self.out.name].copy_(t)

    # untimed launches before the first measurement of each run, so the
    # GPU clock has left the idle level it drops to while the host works
    # (CPU baselines, parity checks between pipelines), before warmup
    SETTLE_S = 0.25
    settled = False

    def time(self, steps, warmup):
        torch = self.torch
        if not self.settled:
            self.settled = True
            t_end = time.perf_counter() + DeviceRun.SETTLE_S
            k = 0
            while time.perf_counter() < t_end:
                for _ in range(10):
                    self.launches[k % self.nsets]()
                    k += 1
                torch.cuda.synchronize()
        for i in range(warmup):
            self.launches[i % self.nsets]()
        torch.cuda.synchronize()
        # HIP events on the launch stream around the timed region only: the
        # launches are pre-marshalled, so the host stays ahead of the GPU and
        # the kernels run back to back (an event between two kernels would
        # itself add a gap of several microseconds)
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        t0 = time.perf_counter()
        ev0.record()
        for i in range(steps):
            self.launches[i % self.nsets]()
        ev1.record()
        torch.cuda.synchronize()
        self.last = (steps - 1) % self.nsets if steps else 0
        return ev0.elapsed_time(ev1) / steps, time.perf_counter() - t0

    def last_output(self):
        import numpy as np
        buf = self.sets[self.last][self.out.name].cpu().numpy()
        return buf.view(self.w.out_func.typ.np_dtype()).reshape(self.w.out_shape)


def bits_equal(a, b):
    import numpy as np
    if a.dtype.kind == 'f':
        return bool(np.array_equal(a.view(np.uint32 if a.dtype == np.float32 else np.uint64),
                                   b.view(np.uint32 if b.dtype == np.float32 else np.uint64)))
    return bool(np.array_equal(a, b))


def parity_check(run, ref):
    """'bitexact', or a description of the mismatch: the last timed call's
    output against the oracle's output, and the LDS-overflow status word."""
    st = run.bp.status()
    if st != 0:
        return 'status %d (a tile exceeded its planned LDS box)' % st
    g = run.last_output()
    if bits_equal(g, ref):
        return 'bitexact'
    import numpy as np
    return 'MISMATCH (%d of %d values differ)' % (int(np.count_nonzero(g != ref)), g.size)


def traffic_of(args, w, bp):
    if not (args.traffic and os.path.exists(args.traffic)):
        return None
    with open(args.traffic) as f:
        rec = json.load(f).get(w.name)
    # a record applies only to the kernel source it was measured on
    if rec and rec.get('code_key') == bp.code.key:
        return rec.get('bytes_per_dispatch')
    return None


def build_workload(args, app, which, base=None):
    """The workload ``app`` under schedule ``which`` (explicit --tiles /
    --threads / --gpu-opt override it); ``base`` lends its input data."""
    from polymage_amd.apps import workloads
    gopts = {}
    for kv in args.gpu_opt:
        k, v = kv.split('=', 1)
        if v[:1] in '{[':
            gopts[k] = json.loads(v)        # e.g. group_tiles={"f": [8, 16]}
        else:
            gopts[k] = tuple(int(x) for x in v.split(',')) if ',' in v else int(v)
    if args.threads:
        gopts['threads'] = args.threads
    kw = {'literal_mode': args.literal_mode, 'with_data': base is None}
    if gopts:
        kw['gpu_options'] = gopts
    if args.tiles:
        tiles = tuple(int(t) for t in args.tiles.split('x'))
        w = workloads.BUILDERS[app](tiles=tiles, **kw)
        w.schedule = 'flags: T%s' % args.tiles
    else:
        w = workloads.tuned(app, which, **kw)
    if base is not None:
        w.values = dict(base.values)
    return w


def measure(args, torch, dev, app, which, base=None, ref=None, steps=None, warmup=None):
    """Time one workload under one schedule and check its last output."""
    from polymage_amd import runtime
    w = build_workload(args, app, which, base)
    bp = runtime.build(w.pipe)
    run = DeviceRun(torch, dev, w, bp)
    run.fill_outputs()
    ms, wall = run.time(steps or args.steps, args.warmup if warmup is None else warmup)
    if ref is None:
        ref = oracle_output(w)
    parity = parity_check(run, ref)
    achieved = w.ideal_bytes / (ms / 1e3) / 1e9
    nk = len(bp.code.kernels)
    traffic = traffic_of(args, w, bp)
    res = {'schedule': w.schedule, 'ms_per_call': round(ms, 5), 'mpix_s': round(w.mpix / (ms / 1e3), 3),
           'kernels': nk, 'frac': round(achieved / HBM_PEAK_GBS, 4), 'achieved_gbs': round(achieved, 2),
           'traffic': traffic, 'traffic_over_ideal': round(traffic / w.ideal_bytes, 3) if traffic else None,
           'code_key': bp.code.key, 'parity': parity, 'buffer_sets': run.nsets, 'wall_s': round(wall, 4)}
    del run
    torch.cuda.empty_cache()
    return w, res, ref


def recorded_tuning(app):
    from polymage_amd import tuner
    path = tuner.TUNED_PATH
    if not os.path.exists(path):
        return None
    with open(path) as f:
        rec = json.load(f).get(app)
    if not rec:
        return None
    keep = ('learned_best', 'exhaustive_best', 'learned_over_exhaustive', 'tuning_time_ratio',
            'gpu_time_ratio', 'candidates', 'timed')
    return dict((k, rec[k]) for k in keep if k in rec)


def app_entry(args, torch, dev, app, cpu, headline=None):
    """Both recorded schedules of one configuration, parity-checked, with
    its CPU baseline."""
    ent = {}
    w0, r0, ref = headline if headline else measure(args, torch, dev, app, 'learned')
    log('%s learned: %.4f ms, %s' % (app, r0['ms_per_call'], r0['parity']))
    ent['learned'] = r0
    from polymage_amd.apps import workloads
    if workloads.schedule(app, 'exhaustive')[3] != workloads.schedule(app, 'learned')[3] and not args.tiles:
        _, r1, _ = measure(args, torch, dev, app, 'exhaustive', base=w0, ref=ref)
        log('%s exhaustive: %.4f ms, %s' % (app, r1['ms_per_call'], r1['parity']))
        ent['exhaustive'] = r1
    else:
        ent['exhaustive'] = 'same schedule as learned'
        r1 = r0
    ent['learned_over_exhaustive_mpix'] = round(r1['ms_per_call'] / r0['ms_per_call'], 4)
    ent['recorded_tuning'] = recorded_tuning(app)
    ent['ideal_bytes'] = w0.ideal_bytes
    ent['mpix'] = round(w0.mpix, 4)
    if not args.no_cpu:
        try:
            ent['cpu_baseline'] = cpu_baseline(w0, args.cpu_seconds, cpu)
        except Exception as e:  # the CPU leg must not hide the GPU result
            ent['cpu_baseline'] = {'value': None, 'error': str(e)[:200]}
    return ent


def classed_batch(args, torch, w, recs, outs, refs, stream):
    """The same frames split into the autotuner's size classes, each class
    through its own compiled pipeline (the tiles tuned on a persistent
    launch over that class's frames): one persistent launch per class, each
    on a stream of its own so they overlap, and the classes' captured
    graphs likewise.  Every frame is compared with the oracle after each
    mode.  None when no classes are recorded."""
    import numpy as np
    from polymage_amd import runtime
    from polymage_amd.apps import workloads
    from polymage_amd.size_classes import ClassedBatch
    which = args.schedule if args.schedule in ('learned', 'exhaustive') else 'learned'
    sched = workloads.class_schedules('campipe', which)
    if not sched:
        return None
    built = []
    for (max_area, tiles, gs, gopts, label) in sched:
        cw = workloads.campipe(tiles=tiles, group_size=gs, gpu_options=gopts, with_data=False)
        built.append(runtime.build(cw.pipe))
    classes = [{'max_area': a} for (a, _, _, _, _) in sched]
    cb = ClassedBatch(built, classes, recs, lambda r: r['R'] * r['C'])
    # one side stream per class (default priority), ordered after the
    # timing stream and joined back into it, so the classes' persistent
    # launches overlap each other's tails
    main = torch.cuda.current_stream()
    side = [torch.cuda.Stream() for _ in sched]

    def run_all(mode):
        start = torch.cuda.Event()
        start.record(main)
        for s_ in side:
            s_.wait_event(start)
        cb.run(stream, mode, streams=[s_.cuda_stream for s_ in side])
        for s_ in side:
            done = torch.cuda.Event()
            done.record(s_)
            main.wait_event(done)

    res, parity = {}, {}
    try:
        for mode in ('persistent', 'graph'):
            for _ in range(args.warmup):
                run_all(mode)
            torch.cuda.synchronize()
            for d in outs:
                d.fill_(7)
            ev0 = torch.cuda.Event(enable_timing=True)
            ev1 = torch.cuda.Event(enable_timing=True)
            ev0.record()
            for _ in range(args.steps):
                run_all(mode)
            ev1.record()
            torch.cuda.synchronize()
            res[mode] = ev0.elapsed_time(ev1) / args.steps
            st = cb.status()
            bad = [i for i, (d, r) in enumerate(zip(outs, refs)) if not np.array_equal(d.cpu().numpy(), r)]
            parity[mode] = 'bitexact' if (st == 0 and not bad) else \
                'MISMATCH (status %d, frames %s)' % (st, bad[:8])
    finally:
        cb.destroy()
    return {'classes': [{'max_area': s[0], 'schedule': s[4], 'frames': len(g),
                         'code_key': b.code.key} for s, g, b in zip(sched, cb.groups, built)],
            'persistent_ms': round(res['persistent'], 4), 'graph_ms': round(res['graph'], 4),
            'parity': parity}


def batch_entry(args, torch, dev, cpu):
    """48 camera-pipeline frames of uneven sizes (uint16 Bayer in, uint8 RGB
    out): one persistent work-queue launch per kernel over all frames, the
    same launches replayed from a hipGraph, per-frame plain launches issued
    back to back, and per-frame plain calls one at a time (each waited
    for).  Every frame of the last persistent, graph and per-frame runs is
    compared with the oracle."""
    import numpy as np
    from polymage_amd import runtime
    from polymage_amd.apps import workloads, inputs
    w = build_workload(args, 'campipe', args.schedule if args.schedule != 'default' else 'default')
    w.values.pop('img', None)
    bp = runtime.build(w.pipe)
    from oracle import cemit
    check = cemit.build([w.out_func], 'bench_parity_campipe_%s' % w.pipe.literal_mode,
                        literal_mode=w.pipe.literal_mode)
    frames = inputs.campipe_frames(48, 5)
    m32 = torch.from_numpy(w.values['matrix_3200']).to(dev)
    m70 = torch.from_numpy(w.values['matrix_7000']).to(dev)
    recs, outs, refs = [], [], []
    mpix = 0.0
    ideal = 0
    for img in frames:
        R, C = img.shape
        vals = dict(w.values, R=R, C=C, img=img)
        o = np.full((3, R - 48, C - 64), 7, np.uint8)
        check(dict(vals, process=o))
        refs.append(o)
        d_img = torch.from_numpy(img.view(np.int16)).to(dev)
        d_out = torch.full((3, R - 48, C - 64), 7, dtype=torch.uint8, device=dev)
        outs.append(d_out)
        recs.append(dict(vals, img=d_img, process=d_out, matrix_3200=m32, matrix_7000=m70))
        mpix += (R - 48) * (C - 64) / 1e6
        ideal += R * C * 2 + 3 * (R - 48) * (C - 64)
    h = bp.batch_create(recs)
    stream = torch.cuda.current_stream().cuda_stream
    res = {}
    parity = {}
    for mode in ('frames', 'persistent', 'graph'):
        # each mode from a settled clock (the oracle checks above leave the
        # GPU idle for seconds)
        t_end = time.perf_counter() + DeviceRun.SETTLE_S
        while time.perf_counter() < t_end:
            bp.batch_run(h, stream, mode)
            torch.cuda.synchronize()
        for _ in range(args.warmup):
            bp.batch_run(h, stream, mode)
        torch.cuda.synchronize()
        for d in outs:
            d.fill_(7)
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        ev0.record()
        for _ in range(args.steps):
            bp.batch_run(h, stream, mode)
        ev1.record()
        torch.cuda.synchronize()
        res[mode] = ev0.elapsed_time(ev1) / args.steps
        st = bp.status()
        bad = [i for i, (d, r) in enumerate(zip(outs, refs)) if not np.array_equal(d.cpu().numpy(), r)]
        parity[mode] = 'bitexact' if (st == 0 and not bad) else \
            'MISMATCH (status %d, frames %s)' % (st, bad[:8])
    # the same frames through the plain entry point one call at a time, each
    # finished before the next starts (launches pre-marshalled, so the host
    # adds only launch and wait latency)
    one = [bp.bind_dev(r, stream) for r in recs]
    for fn in one[:args.warmup]:
        fn()
        torch.cuda.current_stream().synchronize()
    t0 = time.perf_counter()
    reps = max(1, args.steps // 4)
    for _ in range(reps):
        for fn in one:
            fn()
            torch.cuda.current_stream().synchronize()
    res['frames_sync'] = (time.perf_counter() - t0) * 1e3 / reps
    bp.batch_destroy(h)
    classed = None if args.no_size_classes else classed_batch(args, torch, w, recs, outs, refs, stream)
    ms = res['graph']
    achieved = ideal / (ms / 1e3) / 1e9
    ent = {'schedule': w.schedule, 'frames': len(frames), 'mpix': round(mpix, 3), 'ideal_bytes': ideal,
           'kernels': len(bp.code.kernels), 'code_key': bp.code.key,
           'graph_ms': round(res['graph'], 4), 'persistent_ms': round(res['persistent'], 4),
           'frames_ms': round(res['frames'], 4), 'frames_sync_ms': round(res['frames_sync'], 4),
           'mpix_s': round(mpix / (ms / 1e3), 3),
           'speedup_graph_vs_frames': round(res['frames'] / res['graph'], 3),
           'speedup_persistent_vs_frames': round(res['frames'] / res['persistent'], 3),
           'speedup_persistent_vs_frames_sync': round(res['frames_sync'] / res['persistent'], 3),
           'frac': round(achieved / HBM_PEAK_GBS, 4), 'achieved_gbs': round(achieved, 2),
           'parity': parity, 'recorded_tuning': recorded_tuning('campipe')}
    if classed is not None:
        classed['speedup_persistent_vs_frames'] = round(res['frames'] / classed['persistent_ms'], 3)
        classed['speedup_graph_vs_frames'] = round(res['frames'] / classed['graph_ms'], 3)
        ent['size_classes'] = classed
    if not args.no_cpu:
        try:
            omp = _omp()
            threads = cpu['omp_threads']
            sample = frames[:8]
            mp = sum((f.shape[0] - 48) * (f.shape[1] - 64) for f in sample) / 1e6
            outs_h = [np.zeros((3, f.shape[0] - 48, f.shape[1] - 64), np.uint8) for f in sample]
            cp = runtime.build_cpu(w.pipe)

            def go():
                for f, o in zip(sample, outs_h):
                    cp.run_host(dict(w.values, R=f.shape[0], C=f.shape[1], img=f, process=o))
            if omp is not None:
                omp.omp_set_num_threads(threads)
            best, runs, spent = _time_cpu(go, args.cpu_seconds)
            cb = {'value': round(mp / best, 3), 'unit': 'MPix/s', 'cores': threads, 'kind': 'port',
                  'sample': '%d runs (%.1f s) over the first 8 frames (%.2f MPix) of the C++/OpenMP '
                            'backend (same schedule), min time, %d threads'
                            % (runs, spent, mp, threads)}
            if omp is not None:
                omp.omp_set_num_threads(1)
                b1, r1, s1 = _time_cpu(go, 0.5 * args.cpu_seconds, max_runs=3, warm=False)
                omp.omp_set_num_threads(threads)
                cb['value_1thread'] = round(mp / b1, 3)
            ent['cpu_baseline'] = cb
        except Exception as e:
            ent['cpu_baseline'] = {'value': None, 'error': str(e)[:200]}
    return ent


def main():
    args = parse()
    if args.settle_s is not None:
        DeviceRun.SETTLE_S = args.settle_s
    if args.gpus != 1:
        print('bench.py measures one GPU; --gpus %d is not supported '
              '(no sharding across GPUs)' % args.gpus, file=sys.stderr)
        sys.exit(2)
    if args.copy_probe:
        copy_probe(args.copy_probe, args.steps)
        return
    # the host's cores as this process sees them, read before any OpenMP
    # runtime starts (a bound initial thread would report one CPU); the
    # CPU baseline's threads then stay on neighbouring cores
    cpu = host_cpu()
    os.environ.setdefault('OMP_PROC_BIND', 'close')
    import torch
    dev = torch.device('cuda:0')
    torch.cuda.set_device(dev)
    copy_gbs = device_copy_gbs(torch, dev)
    ok = True

    if args.workload == 'campipe48':
        ent = batch_entry(args, torch, dev, cpu)
        ok = all(v == 'bitexact' for v in ent['parity'].values())
        result = {
            'metric': METRIC, 'value': ent['mpix_s'], 'unit': 'MPix/s', 'n_gpus': 1,
            'steps': args.steps, 'warmup': args.warmup, 'ms_per_step': ent['graph_ms'],
            'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None, 'dtype': 'u16->u8',
            'data': 'synthetic (48 seeded 10-bit GRBG frames, sizes 32*U{16..81} x 32*U{12..61})',
            'config': {'workload': 'campipe x48 (%.1f MPix out)' % ent['mpix'], 'schedule': ent['schedule'],
                       'kernels': ent['kernels']},
            'parity': 'bitexact' if ok else 'MISMATCH',
            'batch': ent,
            'roofline': {'bound': 'hbm', 'achieved': ent['achieved_gbs'], 'peak': HBM_PEAK_GBS,
                         'unit': 'GB/s', 'frac': ent['frac'], 'traffic': None, 'device_copy_gbs': copy_gbs},
        }
        if 'cpu_baseline' in ent:
            result['cpu_baseline'] = ent['cpu_baseline']
        print(json.dumps(result))
        sys.exit(0 if ok else 3)

    app = args.workload
    log('headline %s (%s schedule)' % (app, args.schedule))
    w, head, ref = measure(args, torch, dev, app, args.schedule)
    ok = ok and head['parity'] == 'bitexact'
    ms = head['ms_per_call']
    result = {
        'metric': METRIC,
        'value': head['mpix_s'],
        'unit': 'MPix/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': ms,
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': {'f32': 'f32', 'u16': 'u16 in/out, f32 inside', 'u8': 'u16->u8'}.get(w.dtype, w.dtype),
        'data': 'synthetic (seeded smooth random field, rotated over %d HBM-resident buffer sets)'
                % head['buffer_sets'],
        'config': {'workload': '%s %s' % (w.name, 'x'.join(str(s) for s in w.out_shape)),
                   'schedule': head['schedule'], 'kernels': head['kernels'],
                   'literal_mode': args.literal_mode, 'code_key': head['code_key'],
                   'wall_s': head['wall_s'],
                   'clock_settle_s': DeviceRun.SETTLE_S},
        'parity': head['parity'],
        'roofline': {'bound': 'hbm', 'achieved': head['achieved_gbs'], 'peak': HBM_PEAK_GBS,
                     'unit': 'GB/s', 'frac': head['frac'], 'traffic': head['traffic'],
                     'kernel_ms': ms,
                     'kernel_ms_basis': 'HIP events around %d back-to-back pipeline calls '
                                        '(%d kernel%s each) / %d' % (args.steps, head['kernels'],
                                                                     's' if head['kernels'] > 1 else '',
                                                                     args.steps),
                     'device_copy_gbs': copy_gbs},
    }
    if args.sustained_steps:
        _, sus, _ = measure(args, torch, dev, app, args.schedule, base=w, ref=ref,
                            steps=args.sustained_steps, warmup=args.warmup)
        ok = ok and sus['parity'] == 'bitexact'
        result['sustained'] = {'steps': args.sustained_steps, 'ms_per_call': sus['ms_per_call'],
                               'mpix_s': sus['mpix_s'], 'frac': sus['frac'], 'parity': sus['parity']}
    if not args.no_cpu:
        try:
            result['cpu_baseline'] = cpu_baseline(w, args.cpu_seconds, cpu)
        except Exception as e:  # the CPU leg must not hide the GPU result
            result['cpu_baseline'] = {'value': None, 'error': str(e)[:200]}
    result['host_cpu'] = cpu
    if args.pipelines == 'all' and not args.tiles:
        pipes = {}
        for a in CONFIG_APPS:
            log('pipeline %s' % a)
            if a == app:
                ent = app_entry(args, torch, dev, a, cpu, headline=(w, head, ref))
                if 'cpu_baseline' in result:
                    ent['cpu_baseline'] = result['cpu_baseline']
            else:
                ent = app_entry(args, torch, dev, a, cpu)
            pipes[a] = ent
            ok = ok and ent['learned']['parity'] == 'bitexact' and \
                (not isinstance(ent['exhaustive'], dict) or ent['exhaustive']['parity'] == 'bitexact')
        log('pipeline campipe48')
        ent = batch_entry(args, torch, dev, cpu)
        ok = ok and all(v == 'bitexact' for v in ent['parity'].values())
        pipes['campipe48'] = ent
        result['pipelines'] = pipes
        if not ok:
            result['parity'] = 'MISMATCH (see pipelines)' if head['parity'] == 'bitexact' else head['parity']
    print(json.dumps(result))
    sys.exit(0 if ok else 3)


if __name__ == '__main__':
    main()
