"""Exhaustive vs learned schedule search on one GPU.

    python -m polymage_amd.autotune --apps unsharp,harris,... --out DIR

For every application:

1. exhaustive: compile every candidate of ``tuner.default_space(app)``
   (parallel hiprtc/g++ workers, fresh kernel cache) and time each on the
   GPU; record compile and timing seconds;
2. learned: train the MLP cost model (HIP kernels) on the timing records
   of the *other* applications only, featurise every candidate statically
   (planning, no compilation), compile and time only the model's top-k;
3. report MPix/s of both winners, their ratio, and the tuning-time ratio
   (learned / exhaustive, wall clock including compilation).

Results go to ``DIR/tuning.json`` and every measurement to
``DIR/timings.jsonl`` (the cost model's training data).
"""

import argparse
import json
import math
import os
import sys
import tempfile
import time

import numpy as np


def _workload(app):
    from .apps import workloads
    return getattr(workloads, app)


def exhaustive(app, procs, db, runs=7):
    from . import tuner
    cands = list(tuner.default_space(app))
    t0 = time.time()
    comp = tuner.compile_all(_workload(app), cands, procs=procs, verbose=True)
    t_compile = time.time() - t0
    t1 = time.time()
    times = tuner.time_candidates(_workload(app), comp, runs=runs, db=db, app=app, verbose=True)
    t_time = time.time() - t1
    feats = dict((k, v[1]) for k, v in comp.items() if v[1] is not None)
    errors = dict((k, v[3]) for k, v in comp.items() if v[3])
    return {'times': times, 'features': feats, 'errors': errors,
            'compile_s': t_compile, 'time_s': t_time, 'n': len(cands)}


def learned(app, procs, train_records, k=4, runs=7, db=None, seed=0):
    from . import tuner, costmodel
    from .apps import workloads
    t0 = time.time()
    cands = list(tuner.default_space(app))
    feats = {}
    for c in cands:
        try:
            w = tuner.build_candidate(_workload(app), c)
            feats[c.key] = costmodel.features(w.pipe)
        except Exception:
            continue
    t_feat = time.time() - t0
    X = np.stack([costmodel.vector(r['features']) for r in train_records])
    y = np.array([costmodel.target_of(r) for r in train_records])
    model = costmodel.MLP(seed=seed)
    t1 = time.time()
    losses = model.fit(X, y, epochs=600, lr=3e-3, seed=seed)
    t_train = time.time() - t1
    order, pred = costmodel.rank(model, feats)
    top = [c for c in cands if c.key in order[:k]]
    t2 = time.time()
    comp = tuner.compile_all(_workload(app), top, procs=procs, verbose=True)
    times = tuner.time_candidates(_workload(app), comp, runs=runs, db=None, app=app, verbose=True)
    t_top = time.time() - t2
    return {'times': times, 'top': order[:k], 'pred': pred, 'losses': losses,
            'features_s': t_feat, 'train_s': t_train, 'top_s': t_top,
            'total_s': time.time() - t0}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--apps', default='unsharp,harris,bilateral,laplacian,campipe')
    ap.add_argument('--out', default='gpurun_out/tune')
    ap.add_argument('--procs', type=int, default=12)
    ap.add_argument('--k', type=int, default=4)
    ap.add_argument('--runs', type=int, default=7)
    ap.add_argument('--reuse-db', default=None, help='timings.jsonl from an earlier run')
    args = ap.parse_args()
    from .tuner import TimingDB
    os.makedirs(args.out, exist_ok=True)
    # fresh kernel cache so compile time is really measured
    os.environ['PMX_KCACHE'] = tempfile.mkdtemp(prefix='pmx_kc_')
    from . import runtime
    runtime.CACHE = os.environ['PMX_KCACHE']
    db = TimingDB(os.path.join(args.out, 'timings.jsonl'))
    apps = args.apps.split(',')
    ex = {}
    for app in apps:
        print('exhaustive', app, flush=True)
        ex[app] = exhaustive(app, args.procs, db, runs=args.runs)
        best = min(ex[app]['times'].items(), key=lambda kv: kv[1])
        print('  %d candidates, best %s %.4f ms, compile %.1fs, timing %.1fs' % (
            ex[app]['n'], best[0], best[1], ex[app]['compile_s'], ex[app]['time_s']), flush=True)
    records = [r for r in db.load() if 'ms' in r]
    if args.reuse_db:
        records += [json.loads(l) for l in open(args.reuse_db) if l.strip()]
    report = {}
    for app in apps:
        train = [r for r in records if r['app'] != app and r.get('features')]
        print('learned', app, 'trained on', len(train), 'records', flush=True)
        # fresh cache again so the learned path pays its own compiles
        runtime.CACHE = tempfile.mkdtemp(prefix='pmx_kc_')
        le = learned(app, args.procs, train, k=args.k, runs=args.runs)
        w = _workload(app)(with_data=False)
        ex_best_key, ex_best = min(ex[app]['times'].items(), key=lambda kv: kv[1])
        le_best_key, le_best = min(le['times'].items(), key=lambda kv: kv[1])
        ex_total = ex[app]['compile_s'] + ex[app]['time_s']
        rep = {
            'candidates': ex[app]['n'],
            'exhaustive_best': ex_best_key, 'exhaustive_ms': ex_best,
            'exhaustive_mpix_s': w.mpix / (ex_best / 1e3),
            'learned_best': le_best_key, 'learned_ms': le_best,
            'learned_mpix_s': w.mpix / (le_best / 1e3),
            'learned_over_exhaustive': ex_best / le_best,
            'exhaustive_tuning_s': ex_total, 'learned_tuning_s': le['total_s'],
            'tuning_time_ratio': le['total_s'] / ex_total,
            'learned_top': le['top'],
            'rank_of_exhaustive_best_in_model': (le['top'].index(ex_best_key)
                                                 if ex_best_key in le['top'] else None),
            'losses': le['losses'],
            'default_ms': ex[app]['times'].get(_default_key(app)),
            'all_times_ms': ex[app]['times'],
            'errors': ex[app]['errors'],
        }
        report[app] = rep
        print('  learned %s %.4f ms vs exhaustive %.4f ms -> %.3f; tuning %.1fs vs %.1fs (%.3f)' % (
            le_best_key, le_best, ex_best, rep['learned_over_exhaustive'], le['total_s'],
            ex_total, rep['tuning_time_ratio']), flush=True)
        with open(os.path.join(args.out, 'tuning.json'), 'w') as f:
            json.dump(report, f, indent=1)


def _default_key(app):
    from .apps.workloads import DEFAULT_TILES
    from .tuner import Candidate
    gs = {'harris': 11, 'unsharp': 5, 'bilateral': 5, 'laplacian': 5, 'campipe': 5}[app]
    return Candidate(DEFAULT_TILES.get(app, (32, 64)), gs, 256).key


if __name__ == '__main__':
    main()
