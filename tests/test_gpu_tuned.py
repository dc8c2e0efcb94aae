"""The kernels the benchmark times, checked at the configured sizes.

Every schedule recorded in polymage_amd/tuned.json (the learned pick and
the exhaustive best of each application) is built through
``workloads.tuned``, the same function bench.py builds its pipelines with,
so the code checked here is the code timed there (same source, same code
key).  Each runs at full size on its seeded benchmark input through the C
ABI (``pipeline_<name>_host``, which also reads the LDS-overflow status
word) and is compared with the oracle bit for bit, untouched output points
included.  The 48 seeded camera frames run through the batch API in its
persistent, hipGraph and per-frame modes, each frame compared with the
oracle.
"""
import numpy as np
import pytest

from oracle import cemit
from polymage_amd import tuner
from polymage_amd.apps import workloads, inputs

from helpers import bits_equal, gpu_build, emulating

pytestmark = pytest.mark.gpu

APPS = ('harris', 'unsharp', 'bilateral', 'laplacian', 'campipe', 'interpolate', 'blend')


def _picks():
    out = []
    for app in APPS:
        keys = {}
        for which in ('learned', 'exhaustive'):
            rec = tuner.tuned_schedule(app, which)
            if rec is not None and rec['key'] not in keys.values():
                keys[which] = rec['key']
                out.append((app, which))
    return out


def _fill(shape, typ):
    dt = typ.np_dtype()
    return np.full(shape, -7.0 if np.dtype(dt).kind == 'f' else 7, dt)


@pytest.mark.parametrize("app,which", _picks())
def test_tuned_schedule_full_size_bitexact(gpu, app, which):
    if emulating():
        pytest.skip("full-size inputs are too slow for the kernel emulator")
    w = workloads.tuned(app, which)
    print("case tuned", app, w.schedule, w.out_shape, flush=True)
    g = _fill(w.out_shape, w.out_func.typ)
    o = g.copy()
    gpu_build(w.pipe).run_host(dict(w.values, **{w.out_func.name: g}))
    orc = cemit.build([w.out_func], 'tuned_%s_%s' % (app, w.pipe.literal_mode),
                      literal_mode=w.pipe.literal_mode)
    orc(dict(w.values, **{w.out_func.name: o}))
    assert bits_equal(g, o)


def test_campipe_48_frames_tuned_all_modes(gpu):
    """The benchmark's batch: 48 seeded frames of uneven sizes, the tuned
    camera-pipeline schedule, persistent launch, graph replays and
    per-frame launches, every frame equal to the oracle's."""
    if emulating():
        pytest.skip("graph capture and device tensors are not emulated")
    import torch
    w = workloads.tuned('campipe', 'learned', with_data=False)
    print("case campipe48", w.schedule, flush=True)
    bp = gpu_build(w.pipe)
    orc = cemit.build([w.out_func], 'tuned_campipe_%s' % w.pipe.literal_mode,
                      literal_mode=w.pipe.literal_mode)
    dev = torch.device('cuda:0')
    m32 = torch.from_numpy(w.values['matrix_3200']).to(dev)
    m70 = torch.from_numpy(w.values['matrix_7000']).to(dev)
    recs, outs, refs = [], [], []
    for img in inputs.campipe_frames(48, 5):
        R, C = img.shape
        vals = dict(w.values, R=R, C=C, img=img)
        o = np.zeros((3, R - 48, C - 64), np.uint8)
        orc(dict(vals, process=o))
        refs.append(o)
        d_img = torch.from_numpy(img.view(np.int16)).to(dev)
        d_out = torch.zeros((3, R - 48, C - 64), dtype=torch.uint8, device=dev)
        outs.append(d_out)
        recs.append(dict(vals, img=d_img, process=d_out, matrix_3200=m32, matrix_7000=m70))
    h = bp.batch_create(recs)
    stream = torch.cuda.current_stream().cuda_stream
    try:
        for mode in ('persistent', 'graph', 'graph', 'frames'):
            for d in outs:
                d.fill_(7)
            bp.batch_run(h, stream, mode)
            torch.cuda.synchronize()
            assert bp.status() == 0, mode
            for i, (d, r) in enumerate(zip(outs, refs)):
                assert np.array_equal(d.cpu().numpy(), r), (mode, i, r.shape)
    finally:
        bp.batch_destroy(h)
