"""Exhaustive vs learned schedule search on one GPU.

    python -m polymage_amd.autotune --apps unsharp,harris,... --out DIR

For every application:

1. exhaustive: compile every candidate of ``tuner.default_space(app)``
   (parallel hiprtc/g++ workers, fresh kernel cache) and time each on the
   GPU; record compile and timing seconds;
2. learned: featurise every candidate statically (planning, no
   compilation), rank them with an ensemble of MLP cost models (HIP
   kernels) trained on the timing records of the *other* applications,
   compile and time the top k, refit with those k measurements, time the
   next k, and so on (two rounds of 10 by default: 10 compiles fit the
   12 compile workers, so a round costs one compile latency), keep the fastest
   measured;
3. report, per application: the exhaustive winner, the learned winner, the
   learned winner's time *in the exhaustive run* over the exhaustive best
   (regret free of re-measurement noise) and as re-measured, and the
   tuning-time ratio (learned seconds / exhaustive seconds, wall clock
   including compilation and model training) and the fraction of
   candidates timed.

Results go to ``DIR/tuning.json`` and every measurement to
``DIR/timings.jsonl`` (the cost model's training data).  ``--record``
writes the learned and exhaustive schedules to ``polymage_amd/tuned.json``
for ``bench.py`` and the kernel-cache prebuild.
"""

import argparse
import json
import os
import tempfile
import time

import numpy as np


def _split(app):
    """'campipe@608x832' -> ('campipe', {'rows': 608, 'cols': 832}): an
    application tuned on a frame of a given size (a size class)."""
    if '@' not in app:
        return app, {}
    name, size = app.split('@')
    r, c = size.split('x')
    return name, {'rows': int(r), 'cols': int(c)}


def _workload(app):
    from .apps import workloads
    return getattr(workloads, _split(app)[0])


# size-class specs -> the frames of the class: such a spec is timed as one
# persistent launch over the class's frames (the batch it will run), not
# as a single frame's plain launch
_CLASS_FRAMES = {}


def _time(app, comp, runs, db, verbose=True):
    from . import tuner
    name, kw = _split(app)
    if app in _CLASS_FRAMES:
        return tuner.time_candidates_batch(_workload(app), comp, _CLASS_FRAMES[app],
                                           db=db, app=app, verbose=verbose)
    return tuner.time_candidates(_workload(app), comp, runs=runs, db=db, app=app,
                                 verbose=verbose, **kw)


def exhaustive(app, procs, db, runs=9, pool=None):
    from . import tuner
    name, kw = _split(app)
    cands = list(tuner.default_space(name))
    tuner.workload_data(_workload(app), **kw)
    t0 = time.time()
    comp = tuner.compile_all(_workload(app), cands, procs=procs, verbose=True, pool=pool, **kw)
    t_compile = time.time() - t0
    t1 = time.time()
    times = _time(app, comp, runs, db)
    t_time = time.time() - t1
    feats = dict((k, v[1]) for k, v in comp.items() if v[1] is not None)
    errors = dict((k, v[3]) for k, v in comp.items() if v[3])
    return {'times': times, 'features': feats, 'errors': errors,
            'compile_s': t_compile, 'time_s': t_time, 'n': len(cands)}


def learned(app, procs, train_records, k=10, rounds=2, runs=9, dup=20, pool=None, seed_k=5):
    """The learned search: rank every candidate by the cost model trained on
    the other applications' timings, compile and time the top k (of which
    the first ``seed_k`` are the analytic estimate's best: the prior the
    model corrects), refit with those k measurements of this application
    (weighted by repetition), time the model's next k, and keep the
    fastest measured.  Returns timings and the split of its wall-clock
    time."""
    from . import tuner, costmodel
    t0 = time.time()
    name, kw = _split(app)
    cands = list(tuner.default_space(name))
    feats = tuner.features_all(_workload(app), cands, procs=procs, pool=pool, **kw)
    feats = dict((key, f) for key, f in feats.items() if f is not None)
    t_feat = time.time() - t0
    keys = list(feats)
    Xc = np.stack([costmodel.vector(feats[kk]) for kk in keys])
    X0 = np.stack([costmodel.vector(r['features']) for r in train_records])
    y0 = costmodel.relative_targets(train_records)
    times = {}
    t_train = 0.0
    t_top = 0.0
    t_gpu = 0.0
    tops = []
    for rnd in range(rounds):
        X, y = X0, y0
        if times:
            mk = [kk for kk in keys if kk in times]
            ym = np.log(np.array([times[kk] for kk in mk]))
            ym = np.minimum(ym - np.median(ym), 1.0)
            X = np.vstack([X0] + [np.stack([costmodel.vector(feats[kk]) for kk in mk])] * dup)
            y = np.concatenate([y0] + [ym] * dup)
        t1 = time.time()
        model = costmodel.Ensemble()
        model.fit(X, y)
        pred = model.predict(Xc)
        t_train += time.time() - t1
        order = [keys[i] for i in np.argsort(pred) if keys[i] not in times]
        if rnd == 0 and seed_k:
            # the first round also times the analytic estimate's best
            # seed_k (a roofline prior: exact when one resource bounds the
            # schedule), the model fills the rest of the round
            est = np.array([feats[kk]['log_est_ns_px'] for kk in keys])
            seeds = [keys[i] for i in np.argsort(est)][:seed_k]
            order = seeds + [kk for kk in order if kk not in seeds]
        top_keys = order[:k]
        tops.append(top_keys)
        top = [c for c in cands if c.key in top_keys]
        t2 = time.time()
        comp = tuner.compile_all(_workload(app), top, procs=procs, verbose=True, pool=pool, **kw)
        t3 = time.time()
        times.update(_time(app, comp, runs, None))
        t_gpu += time.time() - t3
        t_top += time.time() - t2
    return {'times': times, 'top': [kk for t in tops for kk in t], 'rounds': tops,
            'features_s': t_feat, 'train_s': t_train, 'top_s': t_top, 'timing_s': t_gpu,
            'online_s': time.time() - t0, 'n_valid': len(feats)}


def _default_key(app):
    from .apps.workloads import DEFAULT_TILES
    from .tuner import Candidate
    app = _split(app)[0]
    gs = {'harris': 11, 'campipe': 40}.get(app, 5)
    return Candidate(DEFAULT_TILES.get(app, (32, 64)), gs, 256).key


def size_class_specs(spec):
    """'campipe:3' -> (app, classes, ['campipe@RxC', ...]): the size classes
    of the application's benchmark batch (the camera pipeline's 48 seeded
    frames) and the spec of each class's representative frame.  Each
    class's candidates are planned at the representative size and timed as
    a persistent batch launch over the class's own frames."""
    from .apps import inputs
    from .size_classes import frame_classes
    app, k = spec.split(':')
    if app != 'campipe':
        raise ValueError('size classes are defined for the campipe batch only')
    frames = inputs.campipe_frames(48, 5)
    classes = frame_classes([f.shape for f in frames], int(k))
    specs = ['%s@%dx%d' % (app, c['rows'], c['cols']) for c in classes]
    for spec, c in zip(specs, classes):
        _CLASS_FRAMES[spec] = [frames[i] for i in c['members']]
    return app, classes, specs


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--apps', default='unsharp,harris,bilateral,laplacian,campipe')
    ap.add_argument('--out', default='tune_out')
    ap.add_argument('--procs', type=int, default=12)
    ap.add_argument('--k', type=int, default=10)
    ap.add_argument('--rounds', type=int, default=2)
    ap.add_argument('--runs', type=int, default=9)
    ap.add_argument('--reuse-db', default=None, help='timings.jsonl from an earlier run')
    ap.add_argument('--record', action='store_true', help='write polymage_amd/tuned.json')
    ap.add_argument('--size-classes', default=None,
                    help="APP:K: also tune each of the K frame-size classes of APP's batch")
    args = ap.parse_args()
    from .tuner import TimingDB
    os.makedirs(args.out, exist_ok=True)
    # fresh kernel cache so compile time is really measured
    os.environ['PMX_KCACHE'] = tempfile.mkdtemp(prefix='pmx_kc_')
    from . import runtime
    runtime.CACHE = os.environ['PMX_KCACHE']
    db = TimingDB(os.path.join(args.out, 'timings.jsonl'))
    apps = [a for a in args.apps.split(',') if a]
    class_info = None
    if args.size_classes:
        capp, classes, specs = size_class_specs(args.size_classes)
        class_info = (capp, classes, specs)
        apps += specs
    from .tuner import make_pool
    pool = make_pool(args.procs)
    ex = {}
    for app in apps:
        print('exhaustive', app, flush=True)
        ex[app] = exhaustive(app, args.procs, db, runs=args.runs, pool=pool)
        best = min(ex[app]['times'].items(), key=lambda kv: kv[1])
        print('  %d candidates (%d timed), best %s %.4f ms, compile %.1fs, timing %.1fs' % (
            ex[app]['n'], len(ex[app]['times']), best[0], best[1], ex[app]['compile_s'],
            ex[app]['time_s']), flush=True)
    records = [r for r in db.load() if 'ms' in r]
    if args.reuse_db:
        with open(args.reuse_db) as f:
            records += [json.loads(l) for l in f if l.strip()]
    report = {}
    for app in apps:
        # (never the same application, at any frame size)
        train = [r for r in records if _split(r['app'])[0] != _split(app)[0] and r.get('features')]
        print('learned', app, 'trained on', len(train), 'records', flush=True)
        # fresh cache again so the learned path pays its own compiles
        runtime.CACHE = tempfile.mkdtemp(prefix='pmx_kc_')
        le = learned(app, args.procs, train, k=args.k, rounds=args.rounds, runs=args.runs,
                     pool=pool)
        t_train = le['train_s']
        w = _workload(app)(with_data=False, **_split(app)[1])
        ex_times = ex[app]['times']
        ex_best_key, ex_best = min(ex_times.items(), key=lambda kv: kv[1])
        le_best_key, le_best = min(le['times'].items(), key=lambda kv: kv[1])
        # the model's pick as the exhaustive run itself measured it
        in_ex = [ex_times[kk] for kk in le['top'] if kk in ex_times]
        le_same = min(in_ex) if in_ex else None
        ex_total = ex[app]['compile_s'] + ex[app]['time_s']
        rank_order = sorted(ex_times, key=ex_times.get)
        rep = {
            'candidates': ex[app]['n'], 'timed': len(ex_times),
            'exhaustive_best': ex_best_key, 'exhaustive_ms': ex_best,
            'exhaustive_mpix_s': w.mpix / (ex_best / 1e3),
            'learned_best': le_best_key, 'learned_ms': le_best,
            'learned_mpix_s': w.mpix / (le_best / 1e3),
            'learned_over_exhaustive': ex_best / le_best,
            'learned_over_exhaustive_same_run': (ex_best / le_same) if le_same else None,
            'learned_top': le['top'],
            'learned_top_ranks_in_exhaustive': [rank_order.index(kk) if kk in rank_order else None
                                                for kk in le['top']],
            'exhaustive_tuning_s': ex_total,
            'learned_s': le['online_s'], 'learned_train_s': t_train,
            'tuning_time_ratio': le['online_s'] / ex_total,
            'timed_ratio': len(le['times']) / float(max(1, len(ex_times))),
            'gpu_time_ratio': le['timing_s'] / max(1e-9, ex[app]['time_s']),
            'learned_detail': {'features_s': le['features_s'], 'train_s': le['train_s'],
                               'top_s': le['top_s'], 'n_valid': le['n_valid'],
                               'rounds': le['rounds']},
            'default_ms': ex_times.get(_default_key(app)),
            'all_times_ms': ex_times,
            'errors': ex[app]['errors'],
        }
        report[app] = rep
        print('  learned %s %.4f ms vs exhaustive %.4f ms -> %.3f (same run %s); '
              'tuning %.1fs (of which training %.1fs) vs %.1fs (%.3f); timed %d of %d' % (
                  le_best_key, le_best, ex_best, rep['learned_over_exhaustive'],
                  '%.3f' % rep['learned_over_exhaustive_same_run']
                  if rep['learned_over_exhaustive_same_run'] else '-',
                  le['online_s'], t_train, ex_total, rep['tuning_time_ratio'],
                  len(le['times']), len(ex_times)), flush=True)
        with open(os.path.join(args.out, 'tuning.json'), 'w') as f:
            json.dump(report, f, indent=1)
    pool.close()
    pool.join()
    if class_info is not None:
        capp, classes, specs = class_info
        report['%s:classes' % capp] = [dict((k, v) for k, v in c.items() if k != 'members')
                                       for c in classes]
        with open(os.path.join(args.out, 'tuning.json'), 'w') as f:
            json.dump(report, f, indent=1)
    if args.record:
        record(report)


def record(report, path=None):
    """Keep the chosen schedules (and the headline ratios) in tuned.json."""
    from .tuner import TUNED_PATH
    path = path or TUNED_PATH
    keep = {}
    if os.path.exists(path):
        with open(path) as f:
            keep = json.load(f)
    fields = ('exhaustive_best', 'exhaustive_ms', 'learned_best', 'learned_ms',
              'learned_over_exhaustive', 'learned_over_exhaustive_same_run',
              'tuning_time_ratio', 'timed_ratio', 'gpu_time_ratio', 'candidates', 'timed')
    for app, r in report.items():
        if app.endswith(':classes') or '@' in app:
            continue
        keep[app] = dict(keep.get(app, {}), **dict((k, r[k]) for k in fields))
    # size classes: the class bounds and each class's picks, kept with the
    # application (its single-size picks stay as they are)
    for app, classes in report.items():
        if not app.endswith(':classes'):
            continue
        name = app.split(':')[0]
        out = []
        for c in classes:
            spec = '%s@%dx%d' % (name, c['rows'], c['cols'])
            if spec in report:
                out.append(dict(c, **dict((k, report[spec][k]) for k in fields)))
        keep.setdefault(name, {})['classes'] = out
    with open(path, 'w') as f:
        json.dump(keep, f, indent=1, sort_keys=True)


if __name__ == '__main__':
    main()
