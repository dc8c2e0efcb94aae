"""Bilateral grid (sandbox/apps/python/img_proc/bilateral_grid/
polymage_bilateral_grid.py:16-130): a (value, weight) scatter Reduction
into a padded (rows/8+4) x (cols/8+4) x 14 grid (sigma_s = 8, 1/sigma_r =
10), 5-tap [1 4 6 4 1] blurs along z, x and y, trilinear slicing and
normalisation."""

from ..constructs import (Parameter, Variable, Interval, Image, Function,
                          Reduction, Reduce, Op, Case, Condition, Select, Cast,
                          Int, Float)


def bilateral_grid(pipe_data=None, defined_borders=False):
    """The reference application.  Its blurs are defined only inside
    ``cond`` yet read one to two cells outside it, and the slice reads
    blury at ``2+xi+1`` and ``2+zi+1`` beyond ``cond`` as well, so border
    outputs depend on unwritten memory.  With ``defined_borders`` the three
    blurs are also defined as 0 outside ``cond`` (every grid cell read is
    then written), which is the benchmark workload."""
    if pipe_data is None:
        pipe_data = {}
    rows = Parameter(Int, "rows")
    cols = Parameter(Int, "cols")
    pipe_data['R'] = rows
    pipe_data['C'] = cols
    pipe_data['rows'] = rows
    pipe_data['cols'] = cols
    x, y, z, c = (Variable(Int, n) for n in ("x", "y", "z", "c"))
    rowr = Interval(Int, 0, rows - 1)
    colr = Interval(Int, 0, cols - 1)
    img = Image(Float, "input", [rows, cols])
    pipe_data['input'] = img
    invSigma_r = 10
    sigma_s = 8
    gridrowr = Interval(Int, 0, rows / sigma_s + 3)
    gridcolr = Interval(Int, 0, cols / sigma_s + 3)
    gridintr = Interval(Int, 0, invSigma_r + 3)
    cr = Interval(Int, 0, 1)
    cd = Interval(Int, 0, 0)
    cond0 = Condition(c, '==', 0)
    cond1 = Condition(c, '==', 1)

    grid = Reduction(([c, x, y, z], [cr, gridrowr, gridcolr, gridintr]),
                     ([x, y], [rowr, colr]), Float, "grid")
    intaccess = Cast(Int, (img(x, y) * invSigma_r))
    grid.defn = [Case(cond0, Reduce(grid(0, 2 + (x / sigma_s), 2 + (y / sigma_s), 2 + intaccess),
                                    img(x, y), Op.Sum)),
                 Case(cond1, Reduce(grid(1, 2 + (x / sigma_s), 2 + (y / sigma_s), 2 + intaccess),
                                    1, Op.Sum))]
    grid.default = 0

    cond = Condition(x, '>=', 2) & Condition(x, '<=', rows / sigma_s + 1) & \
        Condition(y, '>=', 2) & Condition(y, '<=', cols / sigma_s + 1) & \
        Condition(z, '>=', 2) & Condition(z, '<=', invSigma_r + 1)
    outside = Condition(x, '<', 2) | Condition(x, '>', rows / sigma_s + 1) | \
        Condition(y, '<', 2) | Condition(y, '>', cols / sigma_s + 1) | \
        Condition(z, '<', 2) | Condition(z, '>', invSigma_r + 1)

    def cases(e):
        if defined_borders:
            return [Case(cond, e), Case(outside, Cast(Float, 0))]
        return [Case(cond, e)]

    dom4 = ([c, x, y, z], [cr, gridrowr, gridcolr, gridintr])
    blurz = Function(dom4, Float, "blurz")
    blurz.defn = cases(grid(c, x, y, z - 2) + grid(c, x, y, z - 1) * 4 + grid(c, x, y, z) * 6 +
                       grid(c, x, y, z + 1) * 4 + grid(c, x, y, z + 2))
    blurx = Function(dom4, Float, "blurx")
    blurx.defn = cases(blurz(c, x - 2, y, z) + blurz(c, x - 1, y, z) * 4 + blurz(c, x, y, z) * 6 +
                       blurz(c, x + 1, y, z) * 4 + blurz(c, x + 2, y, z))
    blury = Function(dom4, Float, "blury")
    blury.defn = cases(blurx(c, x, y - 2, z) + blurx(c, x, y - 1, z) * 4 + blurx(c, x, y, z) * 6 +
                       blurx(c, x, y + 1, z) * 4 + blurx(c, x, y + 2, z))

    zv = img(x, y) * invSigma_r
    zi = Cast(Int, zv)
    zf = zv - zi
    xf = Cast(Float, x % sigma_s) / sigma_s
    yf = Cast(Float, y % sigma_s) / sigma_s
    xi = x / sigma_s
    yi = y / sigma_s

    def lerp(a, b, w):
        return a + w * (b - a)

    xintery0z0 = lerp(blury(c, 2 + xi, 2 + yi, 2 + zi), blury(c, 2 + xi + 1, 2 + yi, 2 + zi), xf)
    xintery1z0 = lerp(blury(c, 2 + xi, 2 + yi + 1, 2 + zi), blury(c, 2 + xi + 1, 2 + yi + 1, 2 + zi), xf)
    xinteryinterz0 = lerp(xintery0z0, xintery1z0, yf)
    xintery0z1 = lerp(blury(c, 2 + xi, 2 + yi, 2 + zi + 1), blury(c, 2 + xi + 1, 2 + yi, 2 + zi + 1), xf)
    xintery1z1 = lerp(blury(c, 2 + xi, 2 + yi + 1, 2 + zi + 1),
                      blury(c, 2 + xi + 1, 2 + yi + 1, 2 + zi + 1), xf)
    xinteryinterz1 = lerp(xintery0z1, xintery1z1, yf)

    interpolated = Function(([c, x, y], [cr, rowr, colr]), Float, "interpolated")
    interpolated.defn = [lerp(xinteryinterz0, xinteryinterz1, zf)]
    filterexpr = Select(Condition(interpolated(1, x, y), '>', 0),
                        interpolated(0, x, y) / interpolated(1, x, y), img(x, y))
    filtered = Function(([c, x, y], [cd, rowr, colr]), Float, "filtered")
    filtered.defn = [filterexpr]
    return filtered


def bilateral_grid_full(pipe_data=None):
    return bilateral_grid(pipe_data, defined_borders=True)
