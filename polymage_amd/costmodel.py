"""Learned cost model: rank candidate schedules so only the top-k are timed.

Features are computed statically from the group plans (no compilation, no
GPU): per pipeline, aggregated over its kernels,

  stage and kernel counts, tile sizes, overlap (redundant-compute) ratio,
  LDS bytes per workgroup, an occupancy estimate from LDS and threads,
  arithmetic intensity (expression operations per ideal HBM byte), HBM
  bytes per output pixel including live-outs exchanged between groups,
  idle-lane fraction of the sweeps, barriers per tile, and problem size.

The model is a small fully connected ReLU network (``dims`` below) whose
forward pass, mean-squared-error backward pass and Adam update run as the
hand-written HIP kernels of libpmx_rt (``pmx_mlp_forward`` /
``pmx_mlp_backward`` / ``pmx_mlp_adam``; the training loop of a whole
ensemble is one ``pmx_mlp_train_ensemble`` call).  It regresses log(ns per output
pixel) from standardised features.  Candidates are ranked by predicted
time; the exhaustive tuner remains the fallback and the yardstick for regret.
"""

import ctypes
import json
import math
import os

import numpy as np

from .analysis import classify_access, eval_bexpr
from .constructs import Reference
from .expr_ast import (AbstractBinaryOpNode, AbstractUnaryOpNode, Select,
                       Cast, InbuiltFunction)

FEATURES = [
    'n_kernels', 'n_stages', 'max_stages_per_kernel', 'log_tile0', 'log_tile1',
    'log_tile_area', 'log_total_tiles', 'overlap_ratio', 'log_lds_bytes',
    'occupancy', 'log_threads', 'ops_per_px', 'arith_intensity',
    'hbm_bytes_per_px', 'interm_bytes_per_px', 'idle_lane_frac', 'barriers',
    'log_mpix', 'inline_frac', 'reduction_kernels', 'log_est_ns_px',
    'waves_per_cu', 'block_pts', 'log_items_per_thread', 'tail_waste', 'tile_aspect',
    'lds_per_thread',
]

# analytic time estimate: vector ALU lanes per second and HBM bytes per
# second of one MI355X, and a per-launch overhead
_LANE_OPS = 256 * 64 * 2.4e9
_HBM = 8.0e12
_LAUNCH_NS = 3000.0


def _expr_ops(e):
    """Operation count of an expression tree."""
    n = 0
    stack = [e]
    while stack:
        x = stack.pop()
        if isinstance(x, (AbstractBinaryOpNode, AbstractUnaryOpNode)):
            n += 1
        elif isinstance(x, (Select,)):
            n += 2
            stack.append(x.condition.lhs if x.condition.conditional not in ('&&', '||') else x.true_expression)
        elif isinstance(x, InbuiltFunction):
            n += 8
        elif isinstance(x, Reference):
            n += 1
        if hasattr(x, 'children'):
            stack.extend(x.children())
    return n


def features(pipe):
    """Feature dict of a built Pipeline (plans its groups)."""
    from .hip_emit import plan_pipeline
    from .group_plan import GroupPlan
    from .reduction_plan import ReductionPlan
    plans = plan_pipeline(pipe)
    env = pipe.param_env
    outs = pipe.outputs
    # output pixels: the two largest extents of the first output (rows and
    # columns, wherever the channel dimension sits)
    exts = sorted((max(1, hi - lo + 1) for lo, hi in pipe.domain_box(outs[0])), reverse=True)
    out_px = 1
    for e in exts[:2]:
        out_px *= e
    in_bytes = 0
    for img in pipe.inputs:
        n = 1
        for lo, hi in pipe.domain_box(img):
            n *= (hi - lo + 1)
        in_bytes += n * img.typ.nbytes()
    out_bytes = 0
    for o in outs:
        n = 1
        for lo, hi in pipe.domain_box(o):
            n *= (hi - lo + 1)
        out_bytes += n * o.typ.nbytes()
    n_kernels = 0
    n_stages = 0
    max_st = 0
    tiles = [1, 1]
    tot_tiles = 0
    overlap = []
    lds = 0
    occ = []
    ops = 0
    interm = 0
    idle = []
    barriers = 0
    inline_n = 0
    red = 0
    waves_cu = []
    blk_w = blk_n = 0.0
    ipt = []
    tail = []
    threads = int(pipe.gpu_options.get('threads', 256))
    for p in plans:
        if isinstance(p, GroupPlan):
            n_kernels += 1
            n_stages += len(p.stages)
            max_st = max(max_st, len(p.stages))
            ts = [T for (d, T) in p.grid_dims if T > 1]
            for i, T in enumerate(ts[:2]):
                tiles[i] = max(tiles[i], T)
            tot_tiles += p.total_tiles
            tile_vol = 1
            for (d, T) in p.grid_dims:
                tile_vol *= T
            work = 0
            for s in p.stages:
                v = 1
                for e in s.static_ext:
                    v *= e
                for part in s.parts:
                    ops += _expr_ops(part.expr) * v * p.total_tiles
                if s.inline:
                    inline_n += 1
                    continue
                work += v
                U, V = getattr(s, 'block', (1, 1))
                items = int(math.ceil(v / float(U * V)))
                waves = int(math.ceil(items / float(threads)))
                idle.append(1.0 - items / float(waves * threads))
                blk_w += v * U * V
                blk_n += v
                ipt.append(math.log2(max(1.0, items / float(threads))))
            anchor_vol = 1
            for e in p.anchor.static_ext:
                anchor_vol *= e
            overlap.append(work / float(max(1, anchor_vol)) - 1.0)
            lds = max(lds, p.lds_bytes)
            occ.append(min(8, (160 * 1024) // max(1, p.lds_bytes)) if p.lds_bytes else 8)
            waves_cu.append(min(32, occ[-1] * threads // 64) * min(1.0, p.total_tiles / (256.0 * occ[-1])))
            slots = 256.0 * max(1, min(occ[-1], 2048 // threads))
            rounds = math.ceil(p.total_tiles / slots)
            tail.append(1.0 - p.total_tiles / (rounds * slots))
            barriers = max(barriers, p.nphases - 1 - (getattr(p, 'lane_pair', None) is not None))
            for s in p.stages:
                if s.liveout and s.func not in outs:
                    n = 1
                    for lo, hi in pipe.domain_box(s.func):
                        n *= (hi - lo + 1)
                    interm += 2 * n * s.func.typ.nbytes()
        elif isinstance(p, ReductionPlan):
            n_kernels += 1
            red += 1
            n_stages += 1
            # the accumulated array is written once and read by the next group
            n = 1
            for lo, hi in pipe.domain_box(p.func):
                n *= (hi - lo + 1)
            interm += 2 * n * p.func.typ.nbytes()
            # one pass over the reduction domain
            rn = 1
            for d in p.func.reductionDomain:
                rn *= max(1, eval_bexpr(d.upperBound, env) - eval_bexpr(d.lowerBound, env) + 1)
            for case in p.func.defn:
                e = case.expression if hasattr(case, 'expression') else case
                ops += _expr_ops(getattr(e, 'expression', e)) * rn
    ideal = in_bytes + out_bytes
    est_ns = max(ops / _LANE_OPS, (ideal + interm) / _HBM) * 1e9 + n_kernels * _LAUNCH_NS
    f = {
        'n_kernels': n_kernels,
        'n_stages': n_stages,
        'max_stages_per_kernel': max_st,
        'log_tile0': math.log2(tiles[0]),
        'log_tile1': math.log2(tiles[1]),
        'log_tile_area': math.log2(tiles[0] * tiles[1]),
        'log_total_tiles': math.log2(max(1, tot_tiles)),
        'overlap_ratio': float(np.mean(overlap)) if overlap else 0.0,
        'log_lds_bytes': math.log2(max(1, lds)),
        'occupancy': float(min(occ)) if occ else 8.0,
        'log_threads': math.log2(threads),
        'ops_per_px': ops / float(max(1, out_px)),
        'arith_intensity': ops / float(max(1, ideal)),
        'hbm_bytes_per_px': (ideal + interm) / float(max(1, out_px)),
        'interm_bytes_per_px': interm / float(max(1, out_px)),
        'idle_lane_frac': float(np.mean(idle)) if idle else 0.0,
        'barriers': barriers,
        'log_mpix': math.log2(max(1, out_px) / 1e6 + 1e-9),
        'inline_frac': inline_n / float(max(1, n_stages)),
        'reduction_kernels': red,
        'log_est_ns_px': math.log(est_ns / max(1, out_px)),
        'waves_per_cu': float(min(waves_cu)) if waves_cu else 32.0,
        'block_pts': blk_w / blk_n if blk_n else 1.0,
        'log_items_per_thread': float(np.mean(ipt)) if ipt else 0.0,
        'tail_waste': float(np.mean(tail)) if tail else 0.0,
        'tile_aspect': math.log2(tiles[0] / float(tiles[1])),
        'lds_per_thread': lds / float(threads),
    }
    return f


def vector(feats):
    # (records written before a feature existed read it as 0)
    return np.array([float(feats.get(k, 0.0)) for k in FEATURES], dtype=np.float32)


# ---------------------------------------------------------------------------
# MLP on the GPU (hand-written HIP kernels in libpmx_rt)
# ---------------------------------------------------------------------------

class MLP(object):
    """``dims = [len(FEATURES), 64, 64, 1]``; parameters live on the GPU."""

    def __init__(self, dims=None, seed=0, device='cuda:0'):
        import torch
        from . import runtime
        self.rt = runtime.rt()
        self.dims = list(dims or [len(FEATURES), 64, 64, 1])
        self.nl = len(self.dims) - 1
        self.cdims = (ctypes.c_int * len(self.dims))(*self.dims)
        cnt = ctypes.c_int()
        self.rt.pmx_mlp_param_count(self.cdims, self.nl, ctypes.byref(cnt))
        self.count = cnt.value
        g = np.random.default_rng(seed)
        host = []
        for l in range(self.nl):
            fan_in = self.dims[l]
            host.append(g.normal(0, math.sqrt(2.0 / fan_in), (self.dims[l + 1], fan_in)).ravel())
            host.append(np.zeros(self.dims[l + 1]))
        p = np.concatenate(host).astype(np.float32)
        assert p.size == self.count
        self.device = torch.device(device)
        self.params = torch.from_numpy(p).to(self.device)
        self.m = torch.zeros_like(self.params)
        self.v = torch.zeros_like(self.params)
        self.step_n = 0
        self.mu = None
        self.sd = None
        self.ymu = 0.0

    def _stream(self):
        import torch
        return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)

    def _acts(self, batch):
        import torch
        n = sum(batch * d for d in self.dims)
        return torch.empty(n, dtype=torch.float32, device=self.device)

    def forward_raw(self, x):
        """x: torch float32 [batch, dims[0]] on the device -> (out, acts)."""
        import torch
        x = x.contiguous()
        batch = x.shape[0]
        acts = self._acts(batch)
        rc = self.rt.pmx_mlp_forward(ctypes.c_void_p(self.params.data_ptr()), self.cdims, self.nl,
                                     ctypes.c_void_p(x.data_ptr()), batch,
                                     ctypes.c_void_p(acts.data_ptr()), self._stream())
        if rc:
            raise RuntimeError('pmx_mlp_forward: %d' % rc)
        off = sum(batch * d for d in self.dims[:-1])
        return acts[off:off + batch * self.dims[-1]].view(batch, self.dims[-1]), acts

    def backward_raw(self, acts, target):
        import torch
        batch = target.shape[0]
        grads = torch.empty_like(self.params)
        loss = torch.empty(1, dtype=torch.float32, device=self.device)
        target = target.contiguous()
        rc = self.rt.pmx_mlp_backward(ctypes.c_void_p(self.params.data_ptr()), self.cdims, self.nl,
                                      ctypes.c_void_p(acts.data_ptr()),
                                      ctypes.c_void_p(target.data_ptr()), batch,
                                      ctypes.c_void_p(grads.data_ptr()),
                                      ctypes.c_void_p(loss.data_ptr()), self._stream())
        if rc:
            raise RuntimeError('pmx_mlp_backward: %d' % rc)
        return grads, loss

    def adam(self, grads, lr=1e-3):
        self.step_n += 1
        rc = self.rt.pmx_mlp_adam(ctypes.c_void_p(self.params.data_ptr()),
                                  ctypes.c_void_p(grads.data_ptr()),
                                  ctypes.c_void_p(self.m.data_ptr()),
                                  ctypes.c_void_p(self.v.data_ptr()), self.count,
                                  ctypes.c_float(lr), ctypes.c_float(0.9), ctypes.c_float(0.999),
                                  ctypes.c_float(1e-8), self.step_n, self._stream())
        if rc:
            raise RuntimeError('pmx_mlp_adam: %d' % rc)

    # -- data handling ----------------------------------------------------------
    def _norm(self, X):
        return (X - self.mu) / self.sd

    def fit(self, X, y, epochs=400, lr=3e-3, seed=0):
        """X: [n, F] numpy features, y: [n] numpy targets (log ns/px)."""
        import torch
        X = np.asarray(X, np.float32)
        y = np.asarray(y, np.float32)
        self.mu = X.mean(0)
        self.sd = X.std(0) + 1e-3
        self.ymu = float(y.mean())
        xs = torch.from_numpy(self._norm(X).astype(np.float32)).to(self.device)
        ys = torch.from_numpy((y - self.ymu).reshape(-1, 1).astype(np.float32)).to(self.device)
        # the whole loop in one runtime call (kernels launched back to back)
        acts = self._acts(xs.shape[0])
        grads = torch.empty_like(self.params)
        loss = torch.empty(epochs, dtype=torch.float32, device=self.device)
        rc = self.rt.pmx_mlp_train(ctypes.c_void_p(self.params.data_ptr()), self.cdims, self.nl,
                                   ctypes.c_void_p(xs.data_ptr()), ctypes.c_void_p(ys.data_ptr()),
                                   xs.shape[0], epochs, ctypes.c_float(lr), ctypes.c_float(0.9),
                                   ctypes.c_float(0.999), ctypes.c_float(1e-8),
                                   ctypes.c_void_p(self.m.data_ptr()),
                                   ctypes.c_void_p(self.v.data_ptr()),
                                   ctypes.c_void_p(acts.data_ptr()),
                                   ctypes.c_void_p(grads.data_ptr()),
                                   ctypes.c_void_p(loss.data_ptr()), self.step_n, self._stream())
        if rc:
            raise RuntimeError('pmx_mlp_train: %d' % rc)
        self.step_n += epochs
        lh = loss.cpu().numpy()
        return [float(lh[e]) for e in range(epochs) if e % 50 == 0 or e == epochs - 1]

    def predict(self, X):
        import torch
        X = np.asarray(X, np.float32)
        xs = torch.from_numpy(self._norm(X).astype(np.float32)).to(self.device)
        out, _ = self.forward_raw(xs)
        return out.view(-1).cpu().numpy() + self.ymu

    def state(self):
        return {'dims': self.dims, 'params': self.params.cpu().numpy().tolist(),
                'mu': self.mu.tolist(), 'sd': self.sd.tolist(), 'ymu': self.ymu}


def target_of(rec):
    """log(ns per output pixel) of a timing record."""
    return math.log(rec['ms'] * 1e6 / (rec['mpix'] * 1e6))


def relative_targets(records, clip=1.0):
    """Per-application targets: log time minus the application's median
    log time, clipped at ``clip`` (e^1 = 2.7x slower than the median).
    Only the order of one application's schedules matters, so the
    application's overall speed is removed, and very slow schedules do not
    dominate the squared error."""
    y = np.array([target_of(r) for r in records])
    apps = [r['app'] for r in records]
    for a in set(apps):
        idx = [i for i, x in enumerate(apps) if x == a]
        y[idx] = np.minimum(y[idx] - np.median(y[idx]), clip)
    return y


class Ensemble(object):
    """A few MLPs from different initial weights; predictions averaged
    (the single-network ranking varies a lot with the seed when an
    application is left out of the training set)."""

    def __init__(self, n=5, epochs=400):
        self.n = n
        self.epochs = epochs
        self.models = []

    def fit(self, X, y, lr=3e-3):
        """Train all members in one runtime call (``pmx_mlp_train_ensemble``:
        every kernel covers all members, blockIdx.y = member); each member
        ends exactly as ``MLP(seed=s).fit(X, y, epochs, lr)`` would."""
        import torch
        members = [MLP(seed=s) for s in range(self.n)]
        m0 = members[0]
        X = np.asarray(X, np.float32)
        y = np.asarray(y, np.float32)
        mu, sd, ymu = X.mean(0), X.std(0) + 1e-3, float(y.mean())
        for m in members:
            m.mu, m.sd, m.ymu = mu, sd, ymu
        xs = torch.from_numpy(m0._norm(X).astype(np.float32)).to(m0.device)
        ys = torch.from_numpy((y - ymu).reshape(-1, 1).astype(np.float32)).to(m0.device)
        batch = xs.shape[0]
        params = torch.stack([m.params for m in members]).contiguous()
        mm = torch.zeros_like(params)
        vv = torch.zeros_like(params)
        grads = torch.empty_like(params)
        hidden = torch.empty(self.n * sum(batch * d for d in m0.dims[1:]), dtype=torch.float32,
                             device=m0.device)
        loss = torch.empty((self.n, self.epochs), dtype=torch.float32, device=m0.device)
        rc = m0.rt.pmx_mlp_train_ensemble(
            ctypes.c_void_p(params.data_ptr()), m0.cdims, m0.nl, self.n,
            ctypes.c_void_p(xs.data_ptr()), ctypes.c_void_p(ys.data_ptr()), batch, self.epochs,
            ctypes.c_float(lr), ctypes.c_float(0.9), ctypes.c_float(0.999), ctypes.c_float(1e-8),
            ctypes.c_void_p(mm.data_ptr()), ctypes.c_void_p(vv.data_ptr()),
            ctypes.c_void_p(hidden.data_ptr()), ctypes.c_void_p(grads.data_ptr()),
            ctypes.c_void_p(loss.data_ptr()), 0, m0._stream())
        if rc:
            raise RuntimeError('pmx_mlp_train_ensemble: %d' % rc)
        for i, m in enumerate(members):
            m.params, m.m, m.v = params[i], mm[i], vv[i]
            m.step_n = self.epochs
        self.models = members
        lh = loss.cpu().numpy()
        return [[float(lh[i, e]) for e in range(self.epochs) if e % 50 == 0 or e == self.epochs - 1]
                for i in range(self.n)]

    def predict(self, X):
        return sum(m.predict(X) for m in self.models) / float(len(self.models))


def rank(model, feats_by_key):
    keys = list(feats_by_key)
    X = np.stack([vector(feats_by_key[k]) for k in keys])
    pred = model.predict(X)
    order = np.argsort(pred)
    return [keys[i] for i in order], dict(zip(keys, pred.tolist()))
