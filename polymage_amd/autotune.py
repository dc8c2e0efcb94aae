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


def _workload(app):
    from .apps import workloads
    return getattr(workloads, app)


def exhaustive(app, procs, db, runs=9, pool=None):
    from . import tuner
    cands = list(tuner.default_space(app))
    tuner.workload_data(_workload(app))
    t0 = time.time()
    comp = tuner.compile_all(_workload(app), cands, procs=procs, verbose=True, pool=pool)
    t_compile = time.time() - t0
    t1 = time.time()
    times = tuner.time_candidates(_workload(app), comp, runs=runs, db=db, app=app, verbose=True)
    t_time = time.time() - t1
    feats = dict((k, v[1]) for k, v in comp.items() if v[1] is not None)
    errors = dict((k, v[3]) for k, v in comp.items() if v[3])
    return {'times': times, 'features': feats, 'errors': errors,
            'compile_s': t_compile, 'time_s': t_time, 'n': len(cands)}


def learned(app, procs, train_records, k=10, rounds=2, runs=9, dup=20, pool=None):
    """The learned search: rank every candidate by the cost model trained on
    the other applications' timings, compile and time the top k, refit with
    those k measurements of this application (weighted by repetition), time
    the next k, and keep the fastest measured.  Returns timings and the
    split of its wall-clock time."""
    from . import tuner, costmodel
    t0 = time.time()
    cands = list(tuner.default_space(app))
    feats = tuner.features_all(_workload(app), cands, procs=procs, pool=pool)
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
        top_keys = order[:k]
        tops.append(top_keys)
        top = [c for c in cands if c.key in top_keys]
        t2 = time.time()
        comp = tuner.compile_all(_workload(app), top, procs=procs, verbose=True, pool=pool)
        t3 = time.time()
        times.update(tuner.time_candidates(_workload(app), comp, runs=runs, db=None, app=app,
                                           verbose=True))
        t_gpu += time.time() - t3
        t_top += time.time() - t2
    return {'times': times, 'top': [kk for t in tops for kk in t], 'rounds': tops,
            'features_s': t_feat, 'train_s': t_train, 'top_s': t_top, 'timing_s': t_gpu,
            'online_s': time.time() - t0, 'n_valid': len(feats)}


def _default_key(app):
    from .apps.workloads import DEFAULT_TILES
    from .tuner import Candidate
    gs = {'harris': 11, 'campipe': 40}.get(app, 5)
    return Candidate(DEFAULT_TILES.get(app, (32, 64)), gs, 256).key


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
    args = ap.parse_args()
    from .tuner import TimingDB
    os.makedirs(args.out, exist_ok=True)
    # fresh kernel cache so compile time is really measured
    os.environ['PMX_KCACHE'] = tempfile.mkdtemp(prefix='pmx_kc_')
    from . import runtime
    runtime.CACHE = os.environ['PMX_KCACHE']
    db = TimingDB(os.path.join(args.out, 'timings.jsonl'))
    apps = args.apps.split(',')
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
        train = [r for r in records if r['app'] != app and r.get('features')]
        print('learned', app, 'trained on', len(train), 'records', flush=True)
        # fresh cache again so the learned path pays its own compiles
        runtime.CACHE = tempfile.mkdtemp(prefix='pmx_kc_')
        le = learned(app, args.procs, train, k=args.k, rounds=args.rounds, runs=args.runs,
                     pool=pool)
        t_train = le['train_s']
        w = _workload(app)(with_data=False)
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
    for app, r in report.items():
        keep[app] = dict((k, r[k]) for k in (
            'exhaustive_best', 'exhaustive_ms', 'learned_best', 'learned_ms',
            'learned_over_exhaustive', 'learned_over_exhaustive_same_run',
            'tuning_time_ratio', 'timed_ratio', 'gpu_time_ratio', 'candidates', 'timed'))
    with open(path, 'w') as f:
        json.dump(keep, f, indent=1, sort_keys=True)


if __name__ == '__main__':
    main()
