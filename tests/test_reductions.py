"""General reductions (Min, Max, Mul, Sum; histograms) on the GPU.

Cases, after the reference's tests/test_hist.py:10-63 (a per-row Min /
Max / Sum of a float image into ``hist(c, x)``) and the histograms it
implies:

* rows: the reference's test itself.  The reduction variable y is in no
  target index, so each owner thread (one per (c, x)) walks y in order:
  the float results are the serial loop's, bit for bit.
* int histogram: ``hist(bin(img(x, y))) += 1`` and a weighted variant;
  integer updates commute, so one thread per pixel with device atomics.
* int Min / Max / Mul per bin, unsigned Sum.
* float histogram into data-dependent bins with no owned dimension: the
  serial single-thread loop (order matters for float sums).

Every case is compared with the oracle (which honours the operator), on
the CPU emulator in the CPU suite and on the device in the GPU suite.
"""
import numpy as np
import pytest

from oracle import cemit
from polymage_amd import buildPipeline
from polymage_amd.constructs import (Parameter, Variable, Interval, Image, Condition, Case,
                                     Reduction, Reduce, Op, Cast, Min, Max, Int, UInt, Float)
from polymage_amd.hip_emit import plan_pipeline
from polymage_amd.reduction_plan import ReductionPlan

from helpers import bits_equal, gpu_build
import kernel_emu


def rows_hist(default=None):
    """tests/test_hist.py:10-63: hist(c, x) = Min / Max / Sum over y of
    img(x, y) for c = 0 / 1 / 2 (``default`` replaces the initial 0, so the
    Min is not 0 everywhere)."""
    R = Parameter(Int, "R")
    C = Parameter(Int, "C")
    x = Variable(Int, "x")
    y = Variable(Int, "y")
    c = Variable(Int, "c")
    row = Interval(Int, 0, R + 1)
    col = Interval(Int, 0, C + 1)
    cr = Interval(Int, 0, 2)
    img = Image(Float, "img", [R + 2, C + 2])
    hist = Reduction(([c, x], [cr, row]), ([c, x, y], [cr, row, col]), Float, "hist")
    hist.defn = [Case(Condition(c, '==', 0), Reduce(hist(c, x), img(x, y), Op.Min)),
                 Case(Condition(c, '==', 1), Reduce(hist(c, x), img(x, y), Op.Max)),
                 Case(Condition(c, '==', 2), Reduce(hist(c, x), img(x, y), Op.Sum))]
    if default is not None:
        hist.default = default
    return hist, {'R': R, 'C': C}, 'img', (lambda R_, C_: (3, R_ + 2))


def bin_hist(op, typ=Int, weight=None, nbins=64):
    """hist(b) op= value over every pixel, b = (int)(img * (nbins - 1))."""
    R = Parameter(Int, "R")
    C = Parameter(Int, "C")
    x = Variable(Int, "x")
    y = Variable(Int, "y")
    b = Variable(Int, "b")
    img = Image(Float, "img", [R, C])
    hist = Reduction(([b], [Interval(Int, 0, nbins - 1)]),
                     ([x, y], [Interval(Int, 0, R - 1), Interval(Int, 0, C - 1)]), typ, "hist")
    idx = Min(Max(Cast(Int, img(x, y) * float(nbins - 1)), 0), nbins - 1)
    if weight == 'x':
        val = Cast(typ, x % 7 + 1)
    elif weight == 'float':
        val = img(x, y)
    else:
        val = Cast(typ, 1) if typ is not Int else 1
    hist.defn = [Reduce(hist(idx), val, op)]
    if op in (Op.Min, Op.Mul):
        hist.default = 1 if op is Op.Mul else (1 << 30)
    return hist, {'R': R, 'C': C}, 'img', (lambda R_, C_: (nbins,))


CASES = {
    'rows': (rows_hist, 'owner'),
    'rows_half': (lambda: rows_hist(0.5), 'owner'),
    'int_sum': (lambda: bin_hist(Op.Sum), 'atomic'),
    'int_weighted_sum': (lambda: bin_hist(Op.Sum, weight='x'), 'atomic'),
    'int_max': (lambda: bin_hist(Op.Max, weight='x'), 'atomic'),
    'int_min': (lambda: bin_hist(Op.Min, weight='x'), 'atomic'),
    'int_mul': (lambda: bin_hist(Op.Mul, weight='x', nbins=8), 'atomic'),
    'uint_sum': (lambda: bin_hist(Op.Sum, typ=UInt, weight='x'), 'atomic'),
    'float_sum_bins': (lambda: bin_hist(Op.Sum, typ=Float, weight='float', nbins=16), 'serial'),
}


def _build(case, R, C):
    out, params, img_name, oshape = CASES[case][0]()
    est = [(params['R'], R), (params['C'], C)]
    pipe = buildPipeline([out], param_estimates=est, pipe_name='red_' + case, tile_sizes=[16, 32])
    return out, pipe, img_name, oshape(R, C)


def _inputs(case, R, C, seed):
    rng = np.random.default_rng(seed)
    if case.startswith('rows'):
        img = rng.random((R + 2, C + 2)).astype(np.float32)
    else:
        img = rng.random((R, C)).astype(np.float32)
    return img


def _kind(pipe):
    for plan in plan_pipeline(pipe):
        if isinstance(plan, ReductionPlan):
            return 'owner' if plan.owner else ('atomic' if plan.atomic else 'serial')


@pytest.mark.parametrize("case", sorted(CASES))
def test_reduction_is_planned_as_expected(case):
    out, pipe, _, _ = _build(case, 40, 50)
    assert _kind(pipe) == CASES[case][1]


def _run(case, R, C, seed, runner):
    out, pipe, img_name, oshape = _build(case, R, C)
    img = _inputs(case, R, C, seed)
    dt = out.typ.np_dtype()
    o = np.full(oshape, 7, dt)
    g = o.copy()
    cemit.build([out], 'red_oracle_' + case)({'R': R, 'C': C, img_name: img, 'hist': o})
    runner(pipe).run_host({'R': R, 'C': C, img_name: img, 'hist': g})
    return g, o


@pytest.mark.parametrize("case", sorted(CASES))
def test_reduction_emulated_bitexact(case):
    g, o = _run(case, 23, 37, 3, kernel_emu.build)
    assert bits_equal(g, o)


@pytest.mark.gpu
@pytest.mark.parametrize("case,R,C", [(c, 61, 75) for c in sorted(CASES)] +
                         [('rows', 1030, 517), ('int_sum', 2048, 2048), ('int_weighted_sum', 777, 1313),
                          ('uint_sum', 2048, 1536)])
def test_reduction_bitexact(gpu, case, R, C):
    print("case reduction", case, R, C, flush=True)
    g, o = _run(case, R, C, R + C, gpu_build)
    assert bits_equal(g, o)


def _bins(img, nbins):
    # (int)(img * (nbins - 1)) in C: the float product is computed in double
    # (the literal is a double), then truncated toward zero; then clamped
    b = np.trunc(img.astype(np.float64) * float(nbins - 1)).astype(np.int64)
    return np.clip(b, 0, nbins - 1)


def test_reduction_oracle_closed_forms():
    """The oracle's reductions against numpy on the same inputs: per-row
    min / max / sequential float32 sum from the default value, and
    bincount / min / max / product (mod 2^32) per bin."""
    R, C = 23, 37
    for case in ('rows', 'rows_half'):
        out, pipe, _, oshape = _build(case, R, C)
        img = _inputs(case, R, C, 3)
        o = np.full(oshape, 7, np.float32)
        cemit.build([out], 'red_oracle_' + case)({'R': R, 'C': C, 'img': img, 'hist': o})
        d = np.float32(0.5 if case == 'rows_half' else 0.0)
        assert np.array_equal(o[0], np.minimum(d, img.min(axis=1)))
        assert np.array_equal(o[1], np.maximum(d, img.max(axis=1)))
        seq = np.cumsum(np.concatenate([np.full((R + 2, 1), d, np.float32), img], axis=1),
                        axis=1, dtype=np.float32)[:, -1]
        assert bits_equal(o[2], seq)
    img = _inputs('int_sum', R, C, 3)
    x = np.arange(R)[:, None] + np.zeros((1, C), np.int64)
    wgt = (x % 7 + 1).ravel()
    for case, nb in (('int_sum', 64), ('int_weighted_sum', 64), ('int_max', 64), ('int_min', 64),
                     ('int_mul', 8), ('uint_sum', 64)):
        out, pipe, _, oshape = _build(case, R, C)
        o = np.full(oshape, 7, out.typ.np_dtype())
        cemit.build([out], 'red_oracle_' + case)({'R': R, 'C': C, 'img': img, 'hist': o})
        b = _bins(img, nb).ravel()
        if case == 'int_sum':
            want = np.bincount(b, minlength=nb)
        elif case in ('int_weighted_sum', 'uint_sum'):
            want = np.bincount(b, weights=wgt, minlength=nb).astype(np.int64)
        elif case == 'int_max':
            want = np.zeros(nb, np.int64)
            np.maximum.at(want, b, wgt)
        elif case == 'int_min':
            want = np.full(nb, 1 << 30, np.int64)
            np.minimum.at(want, b, wgt)
        else:
            want = np.ones(nb, np.int64)
            for bi, wi in zip(b, wgt):
                want[bi] = (want[bi] * int(wi)) % (1 << 32)
            want = want.astype(np.uint32).view(np.int32).astype(np.int64)
        assert np.array_equal(o.astype(np.int64), want), case
