"""Frame-size classes of a batch (polymage_amd/size_classes.py): the
split of the camera pipeline's 48 seeded frames, the class of a frame, and
a classed batch (each class through its own compiled pipeline with its own
tiles) giving the oracle's output for every frame, run on the kernel
emulator."""
import ctypes

import numpy as np

from oracle import cemit
from polymage_amd.apps import inputs
from polymage_amd.apps.campipe import camera_pipe, MATRIX_3200, MATRIX_7000
from polymage_amd.size_classes import frame_classes, class_of, ClassedBatch

from helpers import build
import kernel_emu


def test_frame_classes_split_the_batch_by_area():
    sizes = [f.shape for f in inputs.campipe_frames(48, 5)]
    cls = frame_classes(sizes, 3)
    assert [len(c['members']) for c in cls] == [16, 16, 16]
    assert sorted(i for c in cls for i in c['members']) == list(range(48))
    assert cls[0]['max_area'] < cls[1]['max_area'] < cls[2]['max_area']
    for k, c in enumerate(cls):
        for i in c['members']:
            assert class_of(sizes[i][0] * sizes[i][1], cls) == k
        # the representative frame is one of the class's own
        assert (c['rows'], c['cols']) in [sizes[i] for i in c['members']]
    # larger than every class: the last one
    assert class_of(10 ** 9, cls) == 2
    assert frame_classes([], 3) == [] and len(frame_classes([(64, 64)], 3)) == 1


class _Emu(object):
    """runtime.BuiltPipeline's batch calls on the kernel emulator."""

    def __init__(self, ep):
        self.ep = ep

    def batch_create(self, recs):
        return self.ep.batch_create(recs)

    def batch_run(self, h, stream=0, mode='persistent'):
        self.ep.batch_run(h, mode)

    def batch_destroy(self, h):
        self.ep.batch_destroy(h)

    def status(self):
        fn = self.ep.lib['pipeline_%s_status' % self.ep.name]
        fn.restype = ctypes.c_int
        return fn()


def test_classed_batch_matches_the_oracle():
    pd = {}
    out = camera_pipe(pd, variant='batch')
    orc = cemit.build([out], 'cam_oracle_batch')
    g = np.random.default_rng(11)
    sizes = [(128, 192), (160, 256), (96, 160), (192, 256), (128, 160)]
    recs, refs = [], []
    for (R, C) in sizes:
        img = g.integers(0, 1024, (R, C)).astype(np.uint16)
        vals = {'R': R, 'C': C, 'colour_temp': 3700.0, 'gamma': 2.2, 'contrast': 50.0, 'img': img,
                'matrix_3200': np.array(MATRIX_3200, np.float32),
                'matrix_7000': np.array(MATRIX_7000, np.float32)}
        o = np.zeros((3, R - 48, C - 64), np.uint8)
        orc(dict(vals, process=o))
        refs.append(o)
        recs.append(dict(vals, process=np.full_like(o, 7)))
    classes = frame_classes(sizes, 2)
    built = []
    for tiles in ([16, 32], [32, 64]):
        pipe = build(out, pd, {'R': 192, 'C': 256}, 'campipe', tiles, group_size=40)
        built.append(_Emu(kernel_emu.build(pipe)))
    cb = ClassedBatch(built, classes, recs, lambda r: r['R'] * r['C'])
    assert sorted(i for grp in cb.groups for i in grp) == list(range(len(sizes)))
    assert all(grp for grp in cb.groups)
    cb.run(0, 'persistent')
    assert cb.status() == 0
    cb.destroy()
    for rec, r in zip(recs, refs):
        assert np.array_equal(rec['process'], r)
