"""DSL re-statements of the reference's checked-in generated video-demo
pipelines (sandbox/video_demo/*_naive.cpp), including their I/O
conversions, so the oracle and the GPU path can be compared with those
files bit for bit.  All are meant for literal_mode='float12' (the checked-in
files print float literals as %.12g + 'f')."""

from polymage_amd.constructs import (Parameter, Variable, Interval, Image, Function,
                                     Case, Condition, Select, Abs, Cast, Int, Float,
                                     UChar)


def vd_harris(pd):
    """video_demo/harris_naive.cpp:10-98: uint8 BGR interleaved input,
    gray = (b*0.114 + g*0.587 + r*0.299)/255, Harris, output x 65536 x 256."""
    R, C = Parameter(Int, "R"), Parameter(Int, "C")
    x, y = Variable(Int, "x"), Variable(Int, "y")
    pd.update(R=R, C=C)
    row, col = Interval(Int, 0, R + 1), Interval(Int, 0, C + 1)
    img = Image(UChar, "img", [R + 2, C + 2, 3])
    cond = Condition(x, '>=', 1) & Condition(x, '<=', R) & Condition(y, '<=', C) & Condition(y, '>=', 1)
    condRed = Condition(x, '>=', 2) & Condition(x, '<=', R - 1) & \
        Condition(y, '<=', C - 1) & Condition(y, '>=', 2)

    def F(n):
        return Function(([x, y], [row, col]), Float, n)
    gray = F("gray")
    gray.defn = [(img(x, y, 0) * 0.114 + img(x, y, 1) * 0.587 + img(x, y, 2) * 0.299) / 255.0]
    Ix = F("Ix")
    Ix.defn = [Case(cond, gray(x - 1, y - 1) * (-1.0 / 12) + gray(x + 1, y - 1) * (1.0 / 12) +
                    gray(x - 1, y) * (-2.0 / 12) + gray(x + 1, y) * (2.0 / 12) +
                    gray(x - 1, y + 1) * (-1.0 / 12) + gray(x + 1, y + 1) * (1.0 / 12))]
    Iy = F("Iy")
    Iy.defn = [Case(cond, gray(x - 1, y - 1) * (-1.0 / 12) + gray(x - 1, y + 1) * (1.0 / 12) +
                    gray(x, y - 1) * (-2.0 / 12) + gray(x, y + 1) * (2.0 / 12) +
                    gray(x + 1, y - 1) * (-1.0 / 12) + gray(x + 1, y + 1) * (1.0 / 12))]
    Ixx, Iyy, Ixy = F("Ixx"), F("Iyy"), F("Ixy")
    Ixx.defn = [Case(cond, Ix(x, y) * Ix(x, y))]
    Iyy.defn = [Case(cond, Iy(x, y) * Iy(x, y))]
    Ixy.defn = [Case(cond, Ix(x, y) * Iy(x, y))]

    def box(f):
        return (f(x - 1, y - 1) + f(x - 1, y) + f(x - 1, y + 1) + f(x, y - 1) + f(x, y) +
                f(x, y + 1) + f(x + 1, y - 1) + f(x + 1, y) + f(x + 1, y + 1))
    Sxx, Syy, Sxy = F("Sxx"), F("Syy"), F("Sxy")
    Sxx.defn = [Case(condRed, box(Ixx))]
    Syy.defn = [Case(condRed, box(Iyy))]
    Sxy.defn = [Case(condRed, box(Ixy))]
    det, trace = F("det"), F("trace")
    det.defn = [Case(condRed, Sxx(x, y) * Syy(x, y) - Sxy(x, y) * Sxy(x, y))]
    trace.defn = [Case(condRed, Sxx(x, y) + Syy(x, y))]
    harris = F("harris")
    harris.defn = [Case(condRed, (det(x, y) - 0.04 * trace(x, y) * trace(x, y)) * 65536.0 * 256.0)]
    return harris


def vd_unsharp(pd):
    """video_demo/unsharp_naive.cpp:10-85: uint8 interleaved input scaled
    by 1/255 into planar float, [1 4 6 4 1]/16 blurs written as
    0.0625/0.25/0.375 taps, sharpen = img*(1+w) + blury*-(w), output
    interleaved float (x, y, c)."""
    R, C = Parameter(Int, "R"), Parameter(Int, "C")
    threshold, weight = Parameter(Float, "threshold"), Parameter(Float, "weight")
    pd.update(R=R, C=C, threshold=threshold, weight=weight)
    x, y, c = Variable(Int, "x"), Variable(Int, "y"), Variable(Int, "c")
    img_flip = Image(UChar, "img", [R + 4, C + 4, 3])
    cr = Interval(Int, 0, 2)
    img = Function(([c, x, y], [cr, Interval(Int, 0, R + 3), Interval(Int, 0, C + 3)]), Float, "imgf")
    img.defn = [Cast(Float, img_flip(x, y, c)) / 255.0]
    blurx = Function(([c, x, y], [cr, Interval(Int, 2, R + 1), Interval(Int, 0, C + 3)]), Float, "blurx")
    blurx.defn = [img(c, x - 2, y) * 0.0625 + img(c, x - 1, y) * 0.25 + img(c, x, y) * 0.375 +
                  img(c, x + 1, y) * 0.25 + img(c, x + 2, y) * 0.0625]
    yr, yc = Interval(Int, 2, R + 1), Interval(Int, 2, C + 1)
    blury = Function(([c, x, y], [cr, yr, yc]), Float, "blury")
    blury.defn = [blurx(c, x, y - 2) * 0.0625 + blurx(c, x, y - 1) * 0.25 + blurx(c, x, y) * 0.375 +
                  blurx(c, x, y + 1) * 0.25 + blurx(c, x, y + 2) * 0.0625]
    sharpen = Function(([c, x, y], [cr, yr, yc]), Float, "sharpen")
    sharpen.defn = [img(c, x, y) * (1 + weight) + blury(c, x, y) * -weight]
    mask = Function(([x, y, c], [yr, yc, cr]), Float, "mask")
    mask.defn = [Select(Condition(Abs(img(c, x, y) - blury(c, x, y)), '<', threshold),
                        img(c, x, y), sharpen(c, x, y))]
    return mask
