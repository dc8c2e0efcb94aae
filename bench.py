#!/usr/bin/env python3
"""Benchmark: MPix/s of the PolyMage unsharp-mask pipeline on one MI355X.

Workload: unsharp mask on an RGB float32 3x2848x4256 smooth random field
(threshold 0.001, weight 3), the reference application's expressions, one
fused group, with the schedule the learned autotuner recorded for this GPU
(``--schedule exhaustive|default`` or ``--tiles`` to change it).  A step is
one pass of the pipeline over the whole image with input and output
resident in HBM; buffer sets are rotated so each step reads cold data.

Prints one JSON line.  ``value`` is output pixels over the timed region's
HIP-event time per step; ``roofline`` divides the ideal traffic (input
bytes + output bytes) by the kernel time of one pipeline call: HIP events on
the stream the kernels run on, around back-to-back calls, over the steps;
``cpu_baseline`` times the oracle's C++/OpenMP port of the same pipeline on
the host cores over a bounded sample.
"""

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


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
    ap.add_argument('--no-cpu', action='store_true', help='skip the CPU baseline')
    ap.add_argument('--cpu-seconds', type=float, default=10.0)
    ap.add_argument('--traffic', default=os.path.join(ROOT, 'profiles', 'traffic.json'),
                    help='JSON file with the HBM bytes per dispatch measured by rocprofv3 PMC runs')
    ap.add_argument('--copy-probe', type=int, default=0, metavar='MIB',
                    help='instead of the pipeline, run a plain device copy of MIB MiB '
                         '(calibrates the FETCH_SIZE / WRITE_SIZE counters)')
    return ap.parse_args()


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


def _time_cpu(fn, budget_s):
    """Min wall time of up to 200 calls of fn within budget_s seconds."""
    fn()     # warm-up
    best = None
    runs = 0
    t_start = time.perf_counter()
    while runs < 200 and time.perf_counter() - t_start < budget_s:
        t0 = time.perf_counter()
        fn()
        dt = time.perf_counter() - t0
        best = dt if best is None else min(best, dt)
        runs += 1
    return best, runs, time.perf_counter() - t_start


def cpu_baseline(w, budget_s):
    """The CPU path on the host cores: for Harris at its benchmark size,
    the reference's own generated refHarris (oracle/_ref, built with the
    reference's compile flags); otherwise the oracle's C++/OpenMP port of
    the same pipeline (reference compile flags)."""
    import numpy as np
    threads = min(os.cpu_count() or 1, 16)
    os.environ.setdefault('OMP_NUM_THREADS', str(threads))
    os.environ.setdefault('OMP_PROC_BIND', 'close')
    ref_so = os.path.join(ROOT, 'oracle', '_ref', 'libref_harris_bench.so')
    if w.name == 'harris' and os.path.exists(ref_so):
        import ctypes
        lib = ctypes.CDLL(ref_so)
        R, C = w.values['R'], w.values['C']
        img = w.values['img']
        out = np.zeros((R + 2, C + 2), np.float32)
        args = (ctypes.c_int(C), ctypes.c_int(R), ctypes.c_void_p(img.ctypes.data),
                ctypes.c_void_p(out.ctypes.data))
        best, runs, spent = _time_cpu(lambda: lib.ref_harris(*args), budget_s)
        return {'value': round(w.mpix / best, 3), 'unit': 'MPix/s',
                'cores': int(os.environ['OMP_NUM_THREADS']), 'kind': 'reference',
                'sample': '%d full-size runs (%.1f s) of the reference\'s generated refHarris '
                          '(ipp_bench/harris/harris_function.cpp, 32x256 tiles, float literals; '
                          'g++ -O3 -fopenmp -march=x86-64-v3 -ftree-vectorize), min time' % (runs, spent)}
    from oracle import cemit
    mode = w.pipe.literal_mode
    prog = cemit.build([w.out_func], 'bench_cpu_%s_%s' % (w.name, mode), flags=cemit.REFERENCE_FLAGS,
                       literal_mode=mode)
    vals = dict(w.values)
    vals[w.out_func.name] = np.zeros(w.out_shape, w.out_func.typ.np_dtype())
    best, runs, spent = _time_cpu(lambda: prog(vals), budget_s)
    return {'value': round(w.mpix / best, 3), 'unit': 'MPix/s', 'cores': int(os.environ['OMP_NUM_THREADS']),
            'kind': 'port',
            'sample': '%d full-size runs (%.1f s) of the oracle C++/OpenMP port of the '
                      'same pipeline (g++ -O3 -fopenmp -march=native), min time' % (runs, spent)}


def schedule_for(args):
    """Tiles / group size / threads for the workload: explicit flags, else
    the cost model's pick recorded by the autotuner (polymage_amd/tuned.json,
    written from a tuning run on this GPU), else the default tiles."""
    from polymage_amd.apps import workloads
    from polymage_amd import tuner
    if args.tiles or args.threads or args.gpu_opt or args.schedule == 'default':
        tiles = tuple(int(t) for t in args.tiles.split('x')) if args.tiles else \
            workloads.DEFAULT_TILES[args.workload]
        return {'tiles': tiles, 'threads': args.threads}, 'default' if not args.tiles else 'flags'
    rec = tuner.tuned_schedule(args.workload, args.schedule)
    if rec is None:
        return {'tiles': workloads.DEFAULT_TILES[args.workload]}, 'default (no tuning record)'
    return rec, '%s: %s' % (args.schedule, rec['key'])


def batch_bench(args):
    """48 camera-pipeline frames of uneven sizes (uint16 Bayer in, uint8 RGB
    out): one persistent work-queue launch per kernel over all frames,
    the same launches replayed from a hipGraph, per-frame plain launches
    issued back to back, and per-frame plain calls one at a time (each
    waited for).  ``value`` is the graph path's MPix/s (output pixels of all
    frames per second)."""
    import numpy as np
    import torch
    from polymage_amd import runtime
    from polymage_amd.apps import workloads, inputs
    args.workload = 'campipe'
    sched, sched_info = schedule_for(args)
    tiles = sched['tiles']
    gopts = {'threads': sched['threads']} if sched.get('threads') else None
    kw = {'group_size': sched['group_size']} if sched.get('group_size') else {}
    kw['literal_mode'] = args.literal_mode
    w = workloads.campipe(tiles=tiles, gpu_options=gopts, with_data=False, **kw)
    bp = runtime.build(w.pipe)
    dev = torch.device('cuda:0')
    torch.cuda.set_device(dev)
    frames = inputs.campipe_frames(48, 5)
    m32 = torch.from_numpy(w.values['matrix_3200']).to(dev)
    m70 = torch.from_numpy(w.values['matrix_7000']).to(dev)
    recs, keep = [], []
    mpix = 0.0
    ideal = 0
    for img in frames:
        R, C = img.shape
        d_img = torch.from_numpy(img.view(np.int16)).to(dev)
        d_out = torch.empty((3, R - 48, C - 64), dtype=torch.uint8, device=dev)
        keep += [d_img, d_out]
        recs.append({'R': R, 'C': C, 'colour_temp': 3700.0, 'gamma': 2.2, 'contrast': 50.0,
                     'img': d_img, 'process': d_out, 'matrix_3200': m32, 'matrix_7000': m70})
        mpix += (R - 48) * (C - 64) / 1e6
        ideal += R * C * 2 + 3 * (R - 48) * (C - 64)
    h = bp.batch_create(recs)
    stream = torch.cuda.current_stream().cuda_stream
    res = {}
    for mode in ('frames', 'persistent', 'graph'):
        for _ in range(args.warmup):
            bp.batch_run(h, stream, mode)
        torch.cuda.synchronize()
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        ev0.record()
        for _ in range(args.steps):
            bp.batch_run(h, stream, mode)
        ev1.record()
        torch.cuda.synchronize()
        res[mode] = ev0.elapsed_time(ev1) / args.steps
    # the same frames through the plain entry point one call at a time, each
    # finished before the next starts: what a caller processing one frame
    # per pipeline_campipe_dev call and waiting for it sees (launches
    # pre-marshalled, so the host adds only launch and wait latency)
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
    ms = res['graph']
    achieved = ideal / (ms / 1e3) / 1e9
    print(json.dumps({
        'metric': 'MPix/s per pipeline, 1 MI355X: learned vs exhaustive schedule; % HBM roofline',
        'value': round(mpix / (ms / 1e3), 3), 'unit': 'MPix/s', 'n_gpus': 1,
        'steps': args.steps, 'warmup': args.warmup, 'ms_per_step': round(ms, 5),
        'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None, 'dtype': 'u16->u8',
        'data': 'synthetic (48 seeded 10-bit GRBG frames, sizes 32*U{16..81} x 32*U{12..61})',
        'config': {'workload': 'campipe x48 (%.1f MPix out)' % mpix, 'tiles': list(tiles),
                   'kernels': len(bp.code.kernels), 'schedule': sched_info},
        'batch': {'frames_ms': round(res['frames'], 4), 'persistent_ms': round(res['persistent'], 4),
                  'graph_ms': round(res['graph'], 4),
                  'speedup_persistent_vs_frames': round(res['frames'] / res['persistent'], 3),
                  'speedup_graph_vs_frames': round(res['frames'] / res['graph'], 3),
                  'frames_sync_ms': round(res['frames_sync'], 4),
                  'speedup_persistent_vs_frames_sync': round(res['frames_sync'] / res['persistent'], 3)},
        'roofline': {'bound': 'hbm', 'achieved': round(achieved, 2), 'peak': HBM_PEAK_GBS,
                     'unit': 'GB/s', 'frac': round(achieved / HBM_PEAK_GBS, 4), 'traffic': None},
    }))


def main():
    args = parse()
    if args.gpus != 1:
        print('bench.py measures one GPU; --gpus %d is not supported '
              '(no sharding across GPUs)' % args.gpus, file=sys.stderr)
        sys.exit(2)
    if args.copy_probe:
        copy_probe(args.copy_probe, args.steps)
        return
    if args.workload == 'campipe48':
        batch_bench(args)
        return
    import numpy as np
    import torch
    from polymage_amd import runtime
    from polymage_amd.apps import workloads

    gopts = {}
    for kv in args.gpu_opt:
        k, v = kv.split('=', 1)
        gopts[k] = tuple(int(x) for x in v.split(',')) if ',' in v else int(v)
    sched, sched_info = schedule_for(args)
    tiles = sched['tiles']
    if sched.get('threads'):
        gopts['threads'] = sched['threads']
    kw = {'group_size': sched['group_size']} if sched.get('group_size') else {}
    kw['literal_mode'] = args.literal_mode
    w = getattr(workloads, args.workload)(tiles=tiles, gpu_options=gopts or None, **kw)
    bp = runtime.build(w.pipe)

    dev = torch.device('cuda:0')
    torch.cuda.set_device(dev)
    inputs = [f for f in bp.inputs]
    out = bp.outputs[0]
    in_bytes = sum(w.values[f.name].nbytes for f in inputs)
    out_bytes = int(np.prod(w.out_shape)) * out.typ.nbytes()
    # rotate buffer sets so the 256 MB Infinity Cache holds no step's data
    nsets = max(2, min(10, int((1 << 30) // (in_bytes + out_bytes)) + 1))
    sets = []
    for _ in range(nsets):
        d = {}
        for f in inputs:
            a = w.values[f.name]
            d[f.name] = torch.from_numpy(a.view(np.int16) if a.dtype == np.uint16 else a).to(dev)
        d[out.name] = torch.empty(out_bytes, dtype=torch.uint8, device=dev)
        sets.append(d)
    scal = dict((k, v) for k, v in w.values.items() if not hasattr(v, 'shape'))
    stream = torch.cuda.current_stream().cuda_stream
    launches = []
    for d in sets:
        vals = dict(scal)
        vals.update(d)
        launches.append(bp.bind_dev(vals, stream))

    for i in range(args.warmup):
        launches[i % nsets]()
    torch.cuda.synchronize()
    # HIP events on the launch stream around the timed region only: the
    # launches are pre-marshalled, so the host stays ahead of the GPU and
    # the kernels run back to back (an event between two kernels would
    # itself add a gap of several microseconds)
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record()
    for i in range(args.steps):
        launches[i % nsets]()
    ev1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    gpu_ms = ev0.elapsed_time(ev1)
    ms_per_step = gpu_ms / args.steps
    value = w.mpix / (ms_per_step / 1e3)
    t_kernel = ms_per_step

    nk = len(bp.code.kernels)
    achieved = w.ideal_bytes / (t_kernel / 1e3) / 1e9
    traffic = None
    if args.traffic and os.path.exists(args.traffic):
        with open(args.traffic) as f:
            rec = json.load(f).get(w.name)
        # a record applies only to the kernel source it was measured on
        if rec and rec.get('code_key') == bp.code.key:
            traffic = rec.get('bytes_per_dispatch')
    copy_gbs = device_copy_gbs(torch, dev)

    result = {
        'metric': 'MPix/s per pipeline, 1 MI355X: learned vs exhaustive schedule; % HBM roofline',
        'value': round(value, 3),
        'unit': 'MPix/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': round(ms_per_step, 5),
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'f32',
        'data': 'synthetic (seeded smooth random field, rotated over %d HBM-resident buffer sets)' % nsets,
        'config': {'workload': '%s %s' % (w.name, 'x'.join(str(s) for s in w.out_shape)),
                   'tiles': list(tiles), 'kernels': nk, 'schedule': sched_info,
                   'literal_mode': args.literal_mode,
                   'gpu_options': dict((k, list(v) if isinstance(v, tuple) else v)
                                       for k, v in gopts.items()),
                   'wall_s': round(wall, 4)},
        'roofline': {'bound': 'hbm', 'achieved': round(achieved, 2), 'peak': HBM_PEAK_GBS,
                     'unit': 'GB/s', 'frac': round(achieved / HBM_PEAK_GBS, 4),
                     'traffic': traffic, 'kernel_ms': round(t_kernel, 5),
                     'kernel_ms_basis': 'HIP events around %d back-to-back pipeline calls '
                                        '(%d kernel%s each) / %d' % (args.steps, nk,
                                                                     's' if nk > 1 else '', args.steps),
                     'device_copy_gbs': copy_gbs},
    }
    if not args.no_cpu:
        try:
            result['cpu_baseline'] = cpu_baseline(w, args.cpu_seconds)
        except Exception as e:  # the CPU leg must not hide the GPU result
            result['cpu_baseline'] = {'value': None, 'error': str(e)[:200]}
    print(json.dumps(result))


if __name__ == '__main__':
    main()
