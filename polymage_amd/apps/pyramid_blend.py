"""Laplacian-pyramid blending in the PolyMage DSL
(sandbox/apps/python/img_proc/pyramid_blend/polymage_blending.py:13-196).

Two float RGB images ``img1``, ``img2`` ``[3, R, C]`` and a float mask
``[R, C]`` are blended over L = 4 pyramid levels:

  Dx_l / Dy_l   [1 4 6 4 1]/16 downsampling along x then y: the Gaussian
                pyramids of both images and of the mask
  Ux_l / Uy_l   Laplacian levels: G_l - upsample(G_{l+1}), the upsample
                [1 6 1]/8 at even and [4 4]/8 at odd positions
  Res_l         lap1 * mask_l + lap2 * (1 - mask_l)
  Col_l / blend collapse from the coarsest level: Res_l + upsample(Col_{l+1})

Each level is defined on the window the reference gives it, which shrinks
towards the borders; points outside the final window are not written.
Stage expressions are kept as in the reference so results are
bit-identical under its C semantics.
"""

from ..constructs import (Parameter, Variable, Interval, Image, Function,
                          Condition, Select, Int, Float)


def pyramid_blend(pipe_data=None, levels=4):
    if pipe_data is None:
        pipe_data = {}
    R = Parameter(Int, "R")
    C = Parameter(Int, "C")
    pipe_data.update(R=R, C=C)
    x, y, c = Variable(Int, "x"), Variable(Int, "y"), Variable(Int, "c")
    L = levels
    img1 = Image(Float, "img1", [3, R, C])
    img2 = Image(Float, "img2", [3, R, C])
    mask = Image(Float, "mask", [R, C])
    pipe_data.update(img1=img1, img2=img2, mask=mask)
    chans = Interval(Int, 0, 2)

    def down(f, l, colour, tag):
        coarse, fine = 1 << l, 1 << (l - 1)
        rows = Interval(Int, l, R / coarse - 2)
        cols_c = Interval(Int, l, C / coarse - 2)
        cols_f = Interval(Int, l - 1, C / fine - 2)
        pre = ([c], [chans]) if colour else ([], [])

        def at(g, *idx):
            return g(c, *idx) if colour else g(*idx)
        dx = Function((pre[0] + [x, y], pre[1] + [rows, cols_f]), Float, "Dx_%d_%s" % (l, tag))
        dx.defn = [(1 * at(f, 2 * x - 2, y) + 4 * at(f, 2 * x - 1, y) + 6 * at(f, 2 * x, y) +
                    4 * at(f, 2 * x + 1, y) + 1 * at(f, 2 * x + 2, y)) * 0.0625]
        dy = Function((pre[0] + [x, y], pre[1] + [rows, cols_c]), Float, "Dy_%d_%s" % (l, tag))
        dy.defn = [(1 * at(dx, x, 2 * y - 2) + 4 * at(dx, x, 2 * y - 1) + 6 * at(dx, x, 2 * y) +
                    4 * at(dx, x, 2 * y + 1) + 1 * at(dx, x, 2 * y + 2)) * 0.0625]
        return dy

    def windows(l):
        """(rows, cols, coarse cols) windows of level l's Laplacian/collapse
        stages."""
        inner = (1 << (L - l + 1)) - 1
        rows = Interval(Int, inner, R / (1 << l) - (1 << (L - l + 1)) + 2)
        cols = Interval(Int, inner, C / (1 << l) - (1 << (L - l + 1)) + 2)
        cols_c = Interval(Int, (1 << (L - l)) - 1, C / (1 << (l + 1)) - (1 << (L - l)) + 2)
        return rows, cols, cols_c

    def upsample_x(fd, rows, cols_c, name):
        ux = Function(([c, x, y], [chans, rows, cols_c]), Float, name)
        even = (1 * fd(c, (x / 2) - 1, y) + 6 * fd(c, (x / 2), y) + 1 * fd(c, (x / 2) + 1, y)) * 0.125
        odd = (4 * fd(c, (x - 1) / 2, y) + 4 * fd(c, (x + 1) / 2, y)) * 0.125
        ux.defn = [Select(Condition(x % 2, "==", 0), even, odd)]
        return ux

    def up_y(ux, c_, x_, y_):
        even = (1 * ux(c_, x_, (y_ / 2) - 1) + 6 * ux(c_, x_, (y_ / 2)) + 1 * ux(c_, x_, (y_ / 2) + 1)) * 0.125
        odd = (4 * ux(c_, x_, (y_ - 1) / 2) + 4 * ux(c_, x_, (y_ + 1) / 2)) * 0.125
        return even, odd

    def laplacian_level(fd, f, l, tag):
        rows, cols, cols_c = windows(l)
        ux = upsample_x(fd, rows, cols_c, "Ux_%d_%s" % (l, tag))
        even, odd = up_y(ux, c, x, y)
        uy = Function(([c, x, y], [chans, rows, cols]), Float, "Uy_%d_%s" % (l, tag))
        uy.defn = [Select(Condition(y % 2, "==", 0), f(c, x, y) - even, f(c, x, y) - odd)]
        return uy

    def collapse(fd, f, l, name):
        rows, cols, cols_c = windows(l)
        ux = upsample_x(fd, rows, cols_c, name + '_x')
        even, odd = up_y(ux, c, x, y)
        uy = Function(([c, x, y], [chans, rows, cols]), Float, name)
        uy.defn = [Select(Condition(y % 2, "==", 0), f(c, x, y) + even, f(c, x, y) + odd)]
        return uy

    g1, g2, gm = [img1], [img2], [mask]
    for l in range(1, L):
        g1.append(down(g1[l - 1], l, True, "img1"))
        g2.append(down(g2[l - 1], l, True, "img2"))
        gm.append(down(gm[l - 1], l, False, "mask"))
    lap1 = [laplacian_level(g1[l + 1], g1[l], l, "img1") for l in range(L - 1)] + [g1[L - 1]]
    lap2 = [laplacian_level(g2[l + 1], g2[l], l, "img2") for l in range(L - 1)] + [g2[L - 1]]

    res = []
    for l in range(L):
        rows, cols, _ = windows(l)
        r = Function(([c, x, y], [chans, rows, cols]), Float, "Res_%d" % l)
        r.defn = [lap1[l](c, x, y) * gm[l](x, y) + lap2[l](c, x, y) * (1 - gm[l](x, y))]
        res.append(r)

    out = res[L - 1]
    for l in range(L - 2, 0, -1):
        out = collapse(out, res[l], l, "Col_%d" % l)
    return collapse(out, res[0], 0, "blend")
