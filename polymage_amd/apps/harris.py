"""Harris corner detection (sandbox/apps/python/img_proc/harris/
polymage_harris.py:13-78): Sobel Ix/Iy with 1/12 weights, products,
3x3 box sums, response det - 0.04*trace^2 on [2,R-1]x[2,C-1]."""

from ..constructs import (Parameter, Variable, Interval, Image, Function,
                          Case, Condition, Int, Float)


def harris_pipe(pipe_data=None):
    if pipe_data is None:
        pipe_data = {}
    R = Parameter(Int, "R")
    C = Parameter(Int, "C")
    x = Variable(Int, "x")
    y = Variable(Int, "y")
    pipe_data['R'] = R
    pipe_data['C'] = C

    row = Interval(Int, 0, R + 1)
    col = Interval(Int, 0, C + 1)
    cond = Condition(x, '>=', 1) & Condition(x, '<=', R) & \
        Condition(y, '<=', C) & Condition(y, '>=', 1)
    condRed = Condition(x, '>=', 2) & Condition(x, '<=', R - 1) & \
        Condition(y, '<=', C - 1) & Condition(y, '>=', 2)

    img = Image(Float, "img", [R + 2, C + 2])
    pipe_data['img'] = img

    def F(name):
        return Function(([x, y], [row, col]), Float, name)

    Iy = F("Iy")
    Iy.defn = [Case(cond, img(x - 1, y - 1) * (-1.0 / 12) + img(x - 1, y + 1) * (1.0 / 12) +
                    img(x, y - 1) * (-2.0 / 12) + img(x, y + 1) * (2.0 / 12) +
                    img(x + 1, y - 1) * (-1.0 / 12) + img(x + 1, y + 1) * (1.0 / 12))]
    Ix = F("Ix")
    Ix.defn = [Case(cond, img(x - 1, y - 1) * (-1.0 / 12) + img(x + 1, y - 1) * (1.0 / 12) +
                    img(x - 1, y) * (-2.0 / 12) + img(x + 1, y) * (2.0 / 12) +
                    img(x - 1, y + 1) * (-1.0 / 12) + img(x + 1, y + 1) * (1.0 / 12))]
    Ixx = F("Ixx")
    Ixx.defn = [Case(cond, Ix(x, y) * Ix(x, y))]
    Iyy = F("Iyy")
    Iyy.defn = [Case(cond, Iy(x, y) * Iy(x, y))]
    Ixy = F("Ixy")
    Ixy.defn = [Case(cond, Ix(x, y) * Iy(x, y))]

    def box(f):
        return (f(x - 1, y - 1) + f(x - 1, y) + f(x - 1, y + 1) +
                f(x, y - 1) + f(x, y) + f(x, y + 1) +
                f(x + 1, y - 1) + f(x + 1, y) + f(x + 1, y + 1))
    Sxx = F("Sxx")
    Sxx.defn = [Case(condRed, box(Ixx))]
    Syy = F("Syy")
    Syy.defn = [Case(condRed, box(Iyy))]
    Sxy = F("Sxy")
    Sxy.defn = [Case(condRed, box(Ixy))]
    det = F("det")
    det.defn = [Case(condRed, Sxx(x, y) * Syy(x, y) - Sxy(x, y) * Sxy(x, y))]
    trace = F("trace")
    trace.defn = [Case(condRed, Sxx(x, y) + Syy(x, y))]
    harris = F("harris")
    harris.defn = [Case(condRed, det(x, y) - 0.04 * trace(x, y) * trace(x, y))]
    return harris
