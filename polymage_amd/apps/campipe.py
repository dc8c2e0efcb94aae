"""Camera pipeline (sandbox/apps/python/img_proc/campipe/
polymage_campipe.py:13-304): hot-pixel suppression, Bayer GRBG
deinterleave, demosaic, interleave, colour matrix interpolated for the
colour temperature (Q8.8 fixed point), tone curve look-up table.

``variant='reference'`` is the application verbatim (Short input, a
65536-entry curve over [-32768, 32767]).  ``variant='batch'`` is the
batched benchmark form: uint16 10-bit input, a 1024-entry curve indexed by
the corrected value clamped to [0, 1023], and a 3x3 sharpen ([1 2 1]^T
[1 2 1]/16 blur, 2*v - blur, saturated to uint8) on every pixel whose 3x3
neighbourhood lies inside the curve's domain; the one-pixel frame keeps
the curved value.  Output is uint8 [3, R-48, C-64] in both variants.
"""

from ..constructs import (Parameter, Variable, Interval, Image, Function,
                          Case, Condition, Select, Cast, Max, Min, Powf, Int,
                          Float, Short, UShort, UChar)


def camera_pipe(pipe_data=None, variant='reference'):
    if pipe_data is None:
        pipe_data = {}
    batch = (variant == 'batch')
    x, y, c, v = (Variable(Int, n) for n in ("x", "y", "c", "v"))
    R = Parameter(Int, "R")
    C = Parameter(Int, "C")
    kelvin = Parameter(Float, "colour_temp")
    gamma = Parameter(Float, "gamma")
    contrast = Parameter(Float, "contrast")
    pipe_data.update(R=R, C=C, colour_temp=kelvin, gamma=gamma, contrast=contrast)

    ghost_x, ghost_y = 12, 16
    ghost_zone_2x = Interval(Int, (ghost_x - 4), (R - 24 - ghost_x + 6) - 1)
    ghost_zone_2y = Interval(Int, (ghost_y - 4), (C - 32 - ghost_y + 6) - 1)
    ghost_zone_0x = Interval(Int, (ghost_x - 0), (R - 24 - ghost_x + 0) - 1)
    ghost_zone_0y = Interval(Int, (ghost_y - 0), (C - 32 - ghost_y + 0) - 1)
    half_ghost_zone_2x = Interval(Int, (ghost_x // 2) - 2, ((R - 24 - ghost_x) // 2 + 2))
    half_ghost_zone_2y = Interval(Int, (ghost_y // 2) - 2, ((C - 32 - ghost_y) // 2 + 2))
    half_ghost_zone_1x = Interval(Int, (ghost_x // 2) - 1, ((R - 24 - ghost_x) // 2 + 1))
    half_ghost_zone_1y = Interval(Int, (ghost_y // 2) - 1, ((C - 32 - ghost_y) // 2 + 1))
    half_ghost_zone_0x = Interval(Int, (ghost_x // 2) - 0, ((R - 24 - ghost_x) // 2 + 0))
    half_ghost_zone_0y = Interval(Int, (ghost_y // 2) - 0, ((C - 32 - ghost_y) // 2 + 0))
    rgb = Interval(Int, 0, 2)
    grbg = Interval(Int, 0, 3)

    matrix_3200 = Image(Float, "matrix_3200", [3, 4])
    matrix_7000 = Image(Float, "matrix_7000", [3, 4])
    img = Image(UShort if batch else Short, "img", [R, C])
    pipe_data.update(img=img, matrix_3200=matrix_3200, matrix_7000=matrix_7000)

    # 1. hot pixel suppression
    x_max = Max(img(x - 2, y), img(x + 2, y))
    y_max = Max(img(x, y - 2), img(x, y + 2))
    x_min = Min(img(x - 2, y), img(x + 2, y))
    y_min = Min(img(x, y - 2), img(x, y + 2))
    max_ = Max(x_max, y_max)
    min_ = Min(x_min, y_min)
    clamp = Min(max_, img(x, y))
    clamp = Max(min_, clamp)
    denoised = Function(([x, y], [ghost_zone_2x, ghost_zone_2y]), Short, "denoised")
    denoised.defn = [clamp]

    # 2. deinterleave
    deinterleaved = Function(([c, x, y], [grbg, half_ghost_zone_2x, half_ghost_zone_2y]),
                             Short, "deinterleaved")
    deinterleaved.defn = [Case(Condition(c, '==', 0), denoised(2 * x, 2 * y)),
                          Case(Condition(c, '==', 1), denoised(2 * x, 2 * y + 1)),
                          Case(Condition(c, '==', 2), denoised(2 * x + 1, 2 * y)),
                          Case(Condition(c, '==', 3), denoised(2 * x + 1, 2 * y + 1))]

    # 3. demosaic
    def plane(name, ch):
        f = Function(([x, y], [half_ghost_zone_2x, half_ghost_zone_2y]), Short, name)
        f.defn = [deinterleaved(ch, x, y)]
        return f
    g_gr = plane("g_gr", 0)
    r_r = plane("r_r", 1)
    b_b = plane("b_b", 2)
    g_gb = plane("g_gb", 3)

    def absd(a, b):
        return Select(Condition(a, '>', b), a - b, b - a)

    gv_r = (g_gb(x - 1, y) + g_gb(x, y)) / 2
    gvd_r = absd(g_gb(x - 1, y), g_gb(x, y))
    gh_r = (g_gr(x, y + 1) + g_gr(x, y)) / 2
    ghd_r = absd(g_gr(x, y + 1), g_gr(x, y))
    g_r = Function(([x, y], [half_ghost_zone_1x, half_ghost_zone_1y]), Short, "g_r")
    g_r.defn = [Select(Condition(ghd_r, '<', gvd_r), gh_r, gv_r)]

    gv_b = (g_gr(x + 1, y) + g_gr(x, y)) / 2
    gvd_b = absd(g_gr(x + 1, y), g_gr(x, y))
    gh_b = (g_gb(x, y - 1) + g_gb(x, y)) / 2
    ghd_b = absd(g_gb(x, y - 1), g_gb(x, y))
    g_b = Function(([x, y], [half_ghost_zone_1x, half_ghost_zone_1y]), Short, "g_b")
    g_b.defn = [Select(Condition(ghd_b, '<', gvd_b), gh_b, gv_b)]

    def zone0(name):
        return Function(([x, y], [half_ghost_zone_0x, half_ghost_zone_0y]), Short, name)

    correction = g_gr(x, y) - (g_r(x, y) + g_r(x, y - 1)) / 2
    r_gr = zone0("r_gr")
    r_gr.defn = [correction + (r_r(x, y - 1) + r_r(x, y)) / 2]
    correction = g_gr(x, y) - (g_b(x, y) + g_b(x - 1, y)) / 2
    b_gr = zone0("b_gr")
    b_gr.defn = [correction + (b_b(x, y) + b_b(x - 1, y)) / 2]
    correction = g_gb(x, y) - (g_r(x, y) + g_r(x + 1, y)) / 2
    r_gb = zone0("r_gb")
    r_gb.defn = [correction + (r_r(x, y) + r_r(x + 1, y)) / 2]
    correction = g_gb(x, y) - (g_b(x, y) + g_b(x, y + 1)) / 2
    b_gb = zone0("b_gb")
    b_gb.defn = [correction + (b_b(x, y) + b_b(x, y + 1)) / 2]

    correction = g_b(x, y) - (g_r(x, y) + g_r(x + 1, y - 1)) / 2
    rp_b = correction + (r_r(x, y) + r_r(x + 1, y - 1)) / 2
    rpd_b = absd(r_r(x, y), r_r(x + 1, y - 1))
    correction = g_b(x, y) - (g_r(x, y - 1) + g_r(x + 1, y)) / 2
    rn_b = correction + (r_r(x, y - 1) + r_r(x + 1, y)) / 2
    rnd_b = absd(r_r(x, y - 1), r_r(x + 1, y))
    r_b = zone0("r_b")
    r_b.defn = [Select(Condition(rpd_b, '<', rnd_b), rp_b, rn_b)]

    correction = g_r(x, y) - (g_b(x, y) + g_b(x - 1, y + 1)) / 2
    bp_r = correction + (b_b(x, y) + b_b(x - 1, y + 1)) / 2
    bpd_r = absd(b_b(x, y), b_b(x - 1, y + 1))
    correction = g_r(x, y) - (g_b(x, y + 1) + g_b(x - 1, y)) / 2
    bn_r = correction + (b_b(x, y + 1) + b_b(x - 1, y)) / 2
    bnd_r = absd(b_b(x, y + 1), b_b(x - 1, y))
    b_r = zone0("b_r")
    b_r.defn = [Select(Condition(bpd_r, '<', bnd_r), bp_r, bn_r)]

    # 4. interleave
    def interleave_x(a, b, name):
        out = Function(([x, y], [ghost_zone_0x, ghost_zone_0y]), Short, name)
        out.defn = [Select(Condition((x % 2), '==', 0), a(x / 2, y), b(x / 2, y))]
        return out

    def interleave_y(a, b, name):
        out = Function(([x, y], [half_ghost_zone_0x, ghost_zone_0y]), Short, name)
        out.defn = [Select(Condition((y % 2), '==', 0), a(x, y / 2), b(x, y / 2))]
        return out

    red = interleave_x(interleave_y(r_gr, r_r, "red_x1"),
                       interleave_y(r_b, r_gb, "red_x2"), "red")(x, y)
    green = interleave_x(interleave_y(g_gr, g_r, "green_x1"),
                         interleave_y(g_b, g_gb, "green_x2"), "green")(x, y)
    blue = interleave_x(interleave_y(b_gr, b_r, "blue_x1"),
                        interleave_y(b_b, b_gb, "blue_x2"), "blue")(x, y)

    # 5. colour correction (Q8.8 matrix interpolated in 1/kelvin)
    alpha = (1.0 / kelvin - 1.0 / 3200) / (1.0 / 7000 - 1.0 / 3200)
    mat = Function(([x, y], [rgb, grbg]), Int, "matrix", "const")
    val = (matrix_3200(x, y) * alpha + matrix_7000(x, y) * (1 - alpha))
    mat.defn = [Cast(Int, val * 256.0)]
    r = mat(0, 3) + mat(0, 0) * red + mat(0, 1) * green + mat(0, 2) * blue
    g = mat(1, 3) + mat(1, 0) * red + mat(1, 1) * green + mat(1, 2) * blue
    b = mat(2, 3) + mat(2, 0) * red + mat(2, 1) * green + mat(2, 2) * blue
    corrected = Function(([c, x, y], [rgb, ghost_zone_0x, ghost_zone_0y]), Short, "corrected")
    corrected.defn = [Case(Condition(c, '==', 2), r / 256),
                      Case(Condition(c, '==', 1), g / 256),
                      Case(Condition(c, '==', 0), b / 256)]

    # 6. tone curve
    def lut_value(anything):
        xf = Cast(Float, anything / 1024.0)
        xf = Min(1.0, xf)
        xf = Max(0.0, xf)
        g_ = Powf(xf, 1.0 / gamma)
        b_ = 2.0 - Powf(2.0, contrast / 100.0)
        a_ = 2.0 - 2.0 * b_
        z = Select(Condition(g_, '>', 0.5),
                   1.0 - (a_ * (1.0 - g_) * (1.0 - g_) + b_ * (1.0 - g_)),
                   a_ * g_ * g_ + b_ * g_)
        return Cast(UChar, Min(Max(z * 256.0, 0.0), 255.0))

    lut_range = Interval(Int, 0, 1023) if batch else Interval(Int, -32768, 32767)
    curve = Function(([v], [lut_range]), UChar, "curveLUT", "const")
    curve.defn = [lut_value(v)]

    name_curved = "curved" if batch else "process"
    curved = Function(([c, x, y], [rgb, ghost_zone_0x, ghost_zone_0y]), UChar, name_curved)
    if batch:
        curved.defn = [curve(Min(Max(Cast(Int, corrected(c, x, y)), 0), 1023))]
    else:
        curved.defn = [curve(corrected(c, x, y))]
        return curved

    # 7. 3x3 sharpen (batch form)
    inner = Condition(x, '>=', ghost_x + 1) & Condition(x, '<=', R - 24 - ghost_x - 2) & \
        Condition(y, '>=', ghost_y + 1) & Condition(y, '<=', C - 32 - ghost_y - 2)
    frame = Condition(x, '<', ghost_x + 1) | Condition(x, '>', R - 24 - ghost_x - 2) | \
        Condition(y, '<', ghost_y + 1) | Condition(y, '>', C - 32 - ghost_y - 2)
    blur = (curved(c, x - 1, y - 1) + 2 * curved(c, x - 1, y) + curved(c, x - 1, y + 1) +
            2 * curved(c, x, y - 1) + 4 * curved(c, x, y) + 2 * curved(c, x, y + 1) +
            curved(c, x + 1, y - 1) + 2 * curved(c, x + 1, y) + curved(c, x + 1, y + 1)) / 16
    sharp = 2 * curved(c, x, y) - blur
    process = Function(([c, x, y], [rgb, ghost_zone_0x, ghost_zone_0y]), UChar, "process")
    process.defn = [Case(inner, Cast(UChar, Min(Max(sharp, 0), 255))),
                    Case(frame, curved(c, x, y))]
    return process


# Halide/FCam calibration matrices the reference application uses
# (campipe/init.py:25-40)
MATRIX_3200 = [[1.6697, -0.2693, -0.4004, -42.4346],
               [-0.3576, 1.0615, 1.5949, -37.1158],
               [-0.2175, -1.8751, 6.9640, -26.6970]]
MATRIX_7000 = [[2.2997, -0.4478, 0.1706, -39.0923],
               [-0.3826, 1.5906, -0.2080, -25.4311],
               [-0.0888, -0.7344, 2.2832, -20.0826]]
