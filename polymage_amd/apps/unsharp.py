"""Unsharp mask (sandbox/apps/python/img_proc/unsharp_mask/
polymage_unsharp_mask.py:16-77): 5-tap binomial blur along x then y,
sharpen = img*(1+weight) - blur*weight, keep img where |img-blur| <
threshold.

``unsharp_mask`` is the reference application verbatim, including its
Case(cond) which leaves blurx columns 0 and C+1..C+3 undefined, so mask
is only defined for y in [3, C-2].  ``unsharp_mask_full`` uses the same
expressions without the Case, like video_demo/unsharp_naive.cpp, so every
output point is defined; it is the benchmark workload."""

from ..constructs import (Parameter, Variable, Interval, Image, Function,
                          Case, Condition, Select, Abs, Int, Float)


def _build(use_case, pipe_data):
    R = Parameter(Int, "R")
    C = Parameter(Int, "C")
    threshold = Parameter(Float, "threshold")
    weight = Parameter(Float, "weight")
    pipe_data.update(R=R, C=C, threshold=threshold, weight=weight)
    x = Variable(Int, "x")
    y = Variable(Int, "y")
    c = Variable(Int, "c")
    img = Image(Float, "img", [3, R + 4, C + 4])
    pipe_data['img'] = img
    cr = Interval(Int, 0, 2)
    xrow = Interval(Int, 2, R + 1)
    xcol = Interval(Int, 0, C + 3)
    yrow = Interval(Int, 2, R + 1)
    ycol = Interval(Int, 2, C + 1)
    cond = Condition(x, '>=', 1) & Condition(x, '<=', R) & \
        Condition(y, '<=', C) & Condition(y, '>=', 1)

    def defn(e):
        return [Case(cond, e)] if use_case else [e]

    blurx = Function(([c, x, y], [cr, xrow, xcol]), Float, "blurx")
    blurx.defn = defn((img(c, x - 2, y) + img(c, x - 1, y) * 4.0 + img(c, x, y) * 6.0 +
                       img(c, x + 1, y) * 4.0 + img(c, x + 2, y)) * 0.0625)
    blury = Function(([c, x, y], [cr, yrow, ycol]), Float, "blury")
    blury.defn = defn((blurx(c, x, y - 2) + blurx(c, x, y - 1) * 4.0 + blurx(c, x, y) * 6.0 +
                       blurx(c, x, y + 1) * 4.0 + blurx(c, x, y + 2)) * 0.0625)
    sharpen = Function(([c, x, y], [cr, yrow, ycol]), Float, "sharpen")
    sharpen.defn = defn(img(c, x, y) * (1 + weight) - blury(c, x, y) * weight)
    masked = Function(([c, x, y], [cr, yrow, ycol]), Float, "mask")
    masked.defn = defn(Select(Condition(Abs(img(c, x, y) - blury(c, x, y)), '<', threshold),
                              img(c, x, y), sharpen(c, x, y)))
    return masked


def unsharp_mask(pipe_data=None):
    return _build(True, {} if pipe_data is None else pipe_data)


def unsharp_mask_full(pipe_data=None):
    return _build(False, {} if pipe_data is None else pipe_data)
