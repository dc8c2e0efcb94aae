"""Multiscale interpolation in the PolyMage DSL
(sandbox/apps/python/img_proc/interpolate/polymage_interpolate.py:13-189).

Input: RGBA float ``img[4, R+2, C+2]`` (colour and alpha in [0, 1], a
one-pixel ghost border).  Output: ``interpolate[3, R, C]``, the colour
channels of a pyramid interpolation that fills low-alpha regions from
coarser levels:

  downsampled_L0  premultiplied colour img(c) * img(3)
  Dx_l / D_l      [1 2 1]/4 downsampling along x then y, levels 1..L-1
  Ux / interp_l   averaging upsample along x then y of the coarser
                  interpolated level, blended as d + (1 - d_alpha) * up
  interpolate     interp_0(c) / interp_0(alpha) over the interior

Every level's functions are zero on their one-point border (the
reference's border Cases).  Stage expressions are kept as in the reference
so results are bit-identical under its C semantics.
"""

from ..constructs import (Parameter, Variable, Interval, Image, Function,
                          Case, Condition, Int, Float)


def interpolate(pipe_data=None, levels=10):
    if pipe_data is None:
        pipe_data = {}
    R = Parameter(Int, "R")
    C = Parameter(Int, "C")
    pipe_data.update(R=R, C=C)
    x, y, c = Variable(Int, "x"), Variable(Int, "y"), Variable(Int, "c")
    img = Image(Float, "img", [4, R + 2, C + 2])
    pipe_data['img'] = img
    rgba = Interval(Int, 0, 3)
    rgb = Interval(Int, 0, 2)

    def box_cases(body, xhi, yhi):
        """``body`` on [1, xhi] x [1, yhi], zero on the one-point frame."""
        inner = Condition(x, '>=', 1) & Condition(x, '<=', xhi) & \
            Condition(y, '>=', 1) & Condition(y, '<=', yhi)
        return [Case(inner, body),
                Case(Condition(y, '<=', 0), 0),
                Case(Condition(y, '>=', yhi + 1), 0),
                Case(Condition(x, '>=', xhi + 1), 0),
                Case(Condition(x, '<=', 0), 0)]

    def down(f, l, tag):
        fine, coarse = 1 << (l - 1), 1 << l
        dx = Function(([c, x, y], [rgba, Interval(Int, 0, R / coarse + 1),
                                   Interval(Int, 0, C / fine + 1)]),
                      Float, "Dx_%d_%s" % (l, tag))
        dx.defn = box_cases((f(c, 2 * x - 2, y) + f(c, 2 * x - 1, y) * 2.0 + f(c, 2 * x, y)) * 0.25,
                            R / coarse, C / fine)
        d = Function(([c, x, y], [rgba, Interval(Int, 0, R / coarse + 1),
                                  Interval(Int, 0, C / coarse + 1)]),
                     Float, "D_%d_%s" % (l, tag))
        d.defn = box_cases((dx(c, x, 2 * y - 2) + dx(c, x, 2 * y - 1) * 2.0 + dx(c, x, 2 * y)) * 0.25,
                           R / coarse, C / coarse)
        return d

    def up_blend(coarse_f, d, l, name):
        here, below = 1 << l, 1 << (l + 1)
        ux = Function(([c, x, y], [rgba, Interval(Int, 0, R / here + 1),
                                   Interval(Int, 0, C / below + 1)]),
                      Float, "Ux_" + name)
        ux.defn = box_cases((coarse_f(c, (x + 1) / 2, y) + coarse_f(c, (x + 2) / 2, y)) / 2.0,
                            R / here, C / below)
        up = (ux(c, x, (y + 1) / 2) + ux(c, x, (y + 2) / 2)) / 2.0
        out = Function(([c, x, y], [rgba, Interval(Int, 0, R / here + 1),
                                    Interval(Int, 0, C / here + 1)]), Float, name)
        out.defn = box_cases(d(c, x, y) + (1.0 - d(3, x, y)) * up, R / here, C / here)
        return out

    pyramid = [Function(([c, x, y], [rgba, Interval(Int, 0, R + 1), Interval(Int, 0, C + 1)]),
                        Float, "downsampled_L0")]
    pyramid[0].defn = [img(c, x, y) * img(3, x, y)]
    for l in range(1, levels):
        pyramid.append(down(pyramid[l - 1], l, "downsampled"))

    filled = pyramid[levels - 1]
    for l in range(levels - 2, -1, -1):
        filled = up_blend(filled, pyramid[l], l, "interpolated_L%d" % l)

    result = Function(([c, x, y], [rgb, Interval(Int, 0, R - 1), Interval(Int, 0, C - 1)]),
                      Float, "interpolate")
    result.defn = [filled(c, x + 1, y + 1) / filled(3, x + 1, y + 1)]
    return result
