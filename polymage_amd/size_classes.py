"""Frames of uneven sizes in size classes, each with its own schedule.

A batch of frames (the camera pipeline's 48 frames of 384x512 to
1952x2592) is split by area into a few classes of about equal frame count.
The autotuner searches each class's schedule on a representative frame
size (``autotune --size-classes campipe:3``) and records it in
``tuned.json`` under the application's ``classes``.  ``ClassedBatch`` then
runs the frames of each class through a persistent work-queue launch of
that class's own compiled pipeline (or the class's captured graph), one
launch per class on one stream, so every class uses the tiles tuned for
frames of its size.
"""

import numpy as np


def frame_classes(sizes, k):
    """Split frames into at most ``k`` classes of about equal count by area
    (rows * cols).  Returns a list of dicts, smallest areas first:
    ``max_area`` (largest frame area of the class: frames up to it belong
    to it), ``rows``/``cols`` (the frame of median area, the size the
    class's schedule is tuned on) and ``members`` (frame indices)."""
    sizes = [(int(r), int(c)) for (r, c) in sizes]
    if not sizes:
        return []
    order = sorted(range(len(sizes)), key=lambda i: (sizes[i][0] * sizes[i][1], i))
    k = max(1, min(int(k), len(sizes)))
    out = []
    for part in np.array_split(np.array(order), k):
        members = [int(i) for i in part]
        if not members:
            continue
        mid = members[len(members) // 2]
        out.append({'max_area': max(sizes[i][0] * sizes[i][1] for i in members),
                    'rows': sizes[mid][0], 'cols': sizes[mid][1],
                    'members': sorted(members)})
    return out


def class_of(area, classes):
    """Index of the class a frame of ``area`` belongs to: the first whose
    ``max_area`` is at least ``area`` (the last class takes larger ones)."""
    for i, c in enumerate(classes):
        if area <= c['max_area']:
            return i
    return len(classes) - 1


class ClassedBatch(object):
    """One batch handle per size class, over that class's compiled pipeline.

    ``built`` is a list of BuiltPipelines (one per class, in class order),
    ``classes`` the class list (``max_area`` per class) and ``records`` the
    frame records of ``BuiltPipeline.batch_create``.  Frames are assigned
    by area; a class with no frame gets no handle."""

    def __init__(self, built, classes, records, area_of):
        self.built = list(built)
        self.classes = list(classes)
        self.groups = [[] for _ in self.classes]
        for i, rec in enumerate(records):
            self.groups[class_of(area_of(rec), self.classes)].append(i)
        self.handles = []
        try:
            for bp, idx in zip(self.built, self.groups):
                self.handles.append(bp.batch_create([records[i] for i in idx]) if idx else None)
        except Exception:
            self.destroy()
            raise

    def run(self, stream=0, mode='persistent', streams=None):
        """Every class's frames, one launch (or graph) per class, in class
        order on ``stream``; with ``streams`` (one stream handle per class)
        class i goes to streams[i], so the classes' launches can overlap
        (ordering against other work on those streams is the caller's)."""
        for i, (bp, h) in enumerate(zip(self.built, self.handles)):
            if h is not None:
                bp.batch_run(h, streams[i] if streams else stream, mode)

    def status(self):
        """The first nonzero device status of the class pipelines, or 0."""
        for bp, h in zip(self.built, self.handles):
            if h is not None:
                st = bp.status()
                if st:
                    return st
        return 0

    def destroy(self):
        for bp, h in zip(self.built, self.handles):
            if h is not None:
                bp.batch_destroy(h)
        self.handles = [None] * len(self.handles)
