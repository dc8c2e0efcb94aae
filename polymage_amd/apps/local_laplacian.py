"""Local Laplacian filter in the PolyMage DSL.

The reference ships no DSL source for it, only the generated C++ of its
video demo (sandbox/video_demo/laplacian_naive.cpp:10-931, K=4 intensity
levels, J=5 pyramid levels, uint8 BGR in and out).  This module restates
that program as a DSL pipeline, parametrised by K, J and the I/O type:

  img_colour  input widened to uint16 (x256 for uint8) with edge-replicated
              padding of 2^J - 1 on the top/left (naive.cpp:118-196)
  img         gray = (0.299 r + 0.587 g + 0.114 b) / 65535
  remapLUT    alpha * t * exp(-t*t/2), t = i/256, i in [-(K-1)*256, (K-1)*256)
  gPyramid_L0 beta*(img - k/(K-1)) + k/(K-1) + remapLUT(clamp(int(img*(K-1)*256),
              0, (K-1)*256) - 256 k), one plane per level k
  D*_L{j}     [1 3 3 1]/8 downsampling along x (Dx_) then y (D_), for the
              per-level pyramid and for the gray input pyramid
  U_lPyramid  Laplacian levels: D_gPyramid_L{j} - upsample(D_gPyramid_L{j+1}),
              upsample = 0.25/0.75 bilinear along x (Ux_) then y
  outLPyramid blend of the two levels bracketing the input gray value
  outGPyramid collapse from the top: outL + upsample(outG of the next level)
  laplacian   colour = clamp(result * (c/65535 + 0.01) / (img + 0.01), 0, 1)
              scaled to the output type

Domains follow the file: level L (0 = full resolution) spans
[2^(J-L) - 1, Rp/2^L - 2^(J+1-L) + 2] for the output pyramids and
[1, Rp/2^L - 2] for the Gaussian ones, where Rp is the padded size.
"""

from ..constructs import (Parameter, Variable, Interval, Image, Function,
                          Case, Condition, Select, Cast, Max, Min, Exp, Int,
                          Float, UChar, UShort)


def local_laplacian(pipe_data=None, K=4, J=5, io='u8', lut_hi_inclusive=False):
    """``io`` 'u8' reproduces the video demo (x256 widening, uint8 output
    of value*255); 'u16' takes uint16 input and writes uint16 value*65535.
    ``lut_hi_inclusive`` extends the remap table by one entry so that a
    gray value of exactly 1.0 (possible with uint16 input) indexes inside
    it; the reference's 1536-entry table would be read one past its end."""
    if pipe_data is None:
        pipe_data = {}
    R = Parameter(Int, "R")           # padded rows
    C = Parameter(Int, "C")           # padded cols
    alpha = Parameter(Float, "alpha")
    beta = Parameter(Float, "beta")
    pipe_data.update(R=R, C=C, alpha=alpha, beta=beta)
    x, y, c, k, i = (Variable(Int, n) for n in ("x", "y", "c", "k", "i"))
    pad_lo = (1 << J) - 1
    pad_hi = (1 << (J + 1)) - 3
    # unpadded input
    in_t = UChar if io == 'u8' else UShort
    img_in = Image(in_t, "img_colour", [R - (pad_lo + pad_hi), C - (pad_lo + pad_hi), 3])
    pipe_data['img_colour'] = img_in
    pipe_data['pad'] = (pad_lo, pad_hi)
    Rin, Cin = R - (pad_lo + pad_hi), C - (pad_lo + pad_hi)

    rowsP = Interval(Int, 0, R - 1)
    colsP = Interval(Int, 0, C - 1)
    chan = Interval(Int, 0, 2)
    levels = Interval(Int, 0, K - 1)

    # edge-replicated, widened input
    cx = Min(Max(x - pad_lo, 0), Rin - 1)
    cy = Min(Max(y - pad_lo, 0), Cin - 1)
    colour = Function(([x, y, c], [rowsP, colsP, chan]), UShort, "colour")
    if io == 'u8':
        colour.defn = [Cast(UShort, img_in(cx, cy, c)) * 256]
    else:
        colour.defn = [Cast(UShort, img_in(cx, cy, c))]

    img = Function(([x, y], [rowsP, colsP]), Float, "img")
    img.defn = [(0.299 * colour(x, y, 2) + 0.587 * colour(x, y, 1) + 0.114 * colour(x, y, 0)) / 65535.0]

    span = (K - 1) * 256
    lut_hi = span if lut_hi_inclusive else span - 1
    remap = Function(([i], [Interval(Int, -span, lut_hi)]), Float, "remapLUT", "const")
    t = Cast(Float, i) / 256.0
    remap.defn = [alpha * t * Exp(-t * t / 2.0)]

    step = 1.0 / (K - 1)
    idx = Min(Max(Cast(Int, img(x, y) * Cast(Float, float(span))), 0), span)
    g0 = Function(([k, x, y], [levels, rowsP, colsP]), Float, "gPyramid_L0")
    g0.defn = [beta * (img(x, y) - k * step) + k * step + remap(idx - 256 * k)]

    def gauss_rows(j):
        return Interval(Int, 1, R / (1 << j) - 2)

    def gauss_cols(j):
        return Interval(Int, 1, C / (1 << j) - 2)

    def out_rows(L):
        return Interval(Int, (1 << (J - L)) - 1, R / (1 << L) - (1 << (J + 1 - L)) + 2)

    def out_cols(L):
        return Interval(Int, (1 << (J - L)) - 1, C / (1 << L) - (1 << (J + 1 - L)) + 2)

    def down_x(f, name, j, with_k):
        cols = gauss_cols(j - 1)
        if with_k:
            fn = Function(([k, x, y], [levels, gauss_rows(j), cols]), Float, name)
            fn.defn = [(f(k, 2 * x - 1, y) + 3.0 * f(k, 2 * x, y) + 3.0 * f(k, 2 * x + 1, y) +
                        f(k, 2 * x + 2, y)) * 0.125]
        else:
            fn = Function(([x, y], [gauss_rows(j), cols]), Float, name)
            fn.defn = [(f(2 * x - 1, y) + 3.0 * f(2 * x, y) + 3.0 * f(2 * x + 1, y) +
                        f(2 * x + 2, y)) * 0.125]
        return fn

    def down_y(f, name, j, with_k):
        if with_k:
            fn = Function(([k, x, y], [levels, gauss_rows(j), gauss_cols(j)]), Float, name)
            fn.defn = [(f(k, x, 2 * y - 1) + 3.0 * f(k, x, 2 * y) + 3.0 * f(k, x, 2 * y + 1) +
                        f(k, x, 2 * y + 2)) * 0.125]
        else:
            fn = Function(([x, y], [gauss_rows(j), gauss_cols(j)]), Float, name)
            fn.defn = [(f(x, 2 * y - 1) + 3.0 * f(x, 2 * y) + 3.0 * f(x, 2 * y + 1) +
                        f(x, 2 * y + 2)) * 0.125]
        return fn

    # Gaussian pyramids
    gP = [g0]
    inG = [img]
    for j in range(1, J):
        dx = down_x(gP[-1], "Dx_gPyramid_L%d" % j, j, True)
        gP.append(down_y(dx, "D_gPyramid_L%d" % j, j, True))
        dxi = down_x(inG[-1], "Dx_inGPyramid_L%d" % j, j, False)
        inG.append(down_y(dxi, "D_inGPyramid_L%d" % j, j, False))

    def up_x(f, name, L, with_k, cols):
        # rows of level L from rows of level L+1 (0.25/0.75 bilinear)
        if with_k:
            fn = Function(([k, x, y], [levels, out_rows(L), cols]), Float, name)
            fn.defn = [Select(Condition(x % 2, '==', 0),
                              0.25 * f(k, x / 2 - 1, y) + 0.75 * f(k, x / 2, y),
                              0.25 * f(k, x / 2 + 1, y) + 0.75 * f(k, x / 2, y))]
        else:
            fn = Function(([x, y], [out_rows(L), cols]), Float, name)
            fn.defn = [Select(Condition(x % 2, '==', 0),
                              0.25 * f(x / 2 - 1, y) + 0.75 * f(x / 2, y),
                              0.25 * f(x / 2 + 1, y) + 0.75 * f(x / 2, y))]
        return fn

    def up_y_k(f, y_):
        return Select(Condition(y_ % 2, '==', 0),
                      0.25 * f(k, x, y_ / 2 - 1) + 0.75 * f(k, x, y_ / 2),
                      0.25 * f(k, x, y_ / 2 + 1) + 0.75 * f(k, x, y_ / 2))

    def up_y(f, y_):
        return Select(Condition(y_ % 2, '==', 0),
                      0.25 * f(x, y_ / 2 - 1) + 0.75 * f(x, y_ / 2),
                      0.25 * f(x, y_ / 2 + 1) + 0.75 * f(x, y_ / 2))

    # Laplacian pyramid of every level plane
    lP = [None] * J
    lP[J - 1] = gP[J - 1]
    for L in range(J - 2, -1, -1):
        ux = up_x(gP[L + 1], "Ux_lPyramid_L%d" % L, L, True, out_cols(L + 1))
        u = Function(([k, x, y], [levels, out_rows(L), out_cols(L)]), Float, "U_lPyramid_L%d" % L)
        u.defn = [gP[L](k, x, y) - up_y_k(ux, y)]
        lP[L] = u

    # blend the two bracketing levels per output pixel
    outL = []
    for L in range(J):
        din = inG[L]
        li = Min(Max(Cast(Int, din(x, y) * Cast(Float, K - 1)), 0), K - 2)
        a = din(x, y) * Cast(Float, K - 1) - li
        fn = Function(([x, y], [out_rows(L), out_cols(L)]), Float, "outLPyramid_L%d" % L)
        fn.defn = [(1.0 - a) * lP[L](li, x, y) + a * lP[L](li + 1, x, y)]
        outL.append(fn)

    # collapse
    outG = [None] * J
    outG[J - 1] = outL[J - 1]
    for L in range(J - 2, -1, -1):
        ux = up_x(outG[L + 1], "Ux_outGPyramid_L%d" % L if L else "Ux_result_ref_gray", L, False,
                  out_cols(L + 1))
        g = Function(([x, y], [out_rows(L), out_cols(L)]), Float,
                     "outGPyramid_L%d" % L if L else "result_ref_gray")
        g.defn = [outL[L](x, y) + up_y(ux, y)]
        outG[L] = g

    out_t = UChar if io == 'u8' else UShort
    scale = 255.0 if io == 'u8' else 65535.0
    res = outG[0]
    ratio = res(x, y) * (colour(x, y, c) / 65535.0 + 0.01) / (img(x, y) + 0.01)
    lap = Function(([x, y, c], [out_rows(0), out_cols(0), chan]), out_t, "laplacian")
    lap.defn = [Cast(out_t, Min(Max(ratio, 0.0), 1.0) * scale)]
    return lap
