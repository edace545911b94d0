"""Code-stream order of the interpolation decomposition, enumerated from its specification with
closed-form loops (no predictor or quantizer): the flat index of each coded point, in emission order.

Level rule (InterpolationDecomposition.hpp:286-339, :166-176): levels = max_d ceil(log2 D_d); anchors
are stored iff some D_d > anchor stride (> 0), then levels = min(levels, log2(stride) + 1) - 1.
Passes per level at stride st = 2^(level-1) (InterpolationDecomposition.hpp:2294-2509):
  3-D direction 0: z-pass (y, x step 2st), y-pass (z step st, x step 2st), x-pass (z, y step st)
  3-D direction != 0: x-pass (z, y step 2st), y-pass (z step 2st, x step st), z-pass (y, x step st)
  2-D direction 0: y-pass (x step 2st), x-pass (y step st); direction 1: x-pass, then y-pass
  1-D: one line.
Row passes (3d_z, 3d_y, 2d_y) loop the outer non-pass axis outermost, then the odd grid indices
g of the pass axis, then the row along the fastest remaining axis. For cubic interpolation with an
even grid count n > 2, the last odd index n-1 is coded after the whole outer loop (3-D), or last (2-D).
Line passes (3d_x, 2d_x, 1-D) code the odd grid indices of each line in increasing order.
"""
import math


def levels_and_anchor(shape, anchor):
    lv = max(int(math.ceil(math.log2(d))) if d > 1 else 0 for d in shape)
    use = anchor > 0 and any(d > anchor for d in shape)
    if use:
        lv = min(lv, int(math.log2(anchor)) + 1) - 1
    return lv, (anchor if use else 0)


def _grid(d, st):
    return (d - 1) // st + 1


def _row_pass_order(n, cubic):
    """Odd grid indices coded inside the outer loop, and the deferred one (or None)."""
    if not cubic:
        return list(range(1, n, 2)), None
    inner = [g for g in range(1, n, 2) if g < n - 1 or (n == 2 and g == 1)]
    deferred = n - 1 if (n % 2 == 0 and n > 2) else None
    return inner, deferred


def order(shape, cubic, direction, anchor):
    N = len(shape)
    lv, a = levels_and_anchor(shape, anchor)
    strides_el = [1] * N
    for i in range(N - 2, -1, -1):
        strides_el[i] = strides_el[i + 1] * shape[i + 1]

    def flat(*c):
        return sum(ci * s for ci, s in zip(c, strides_el))

    out = []
    if not a:
        out.append(0)
    for level in range(lv, 0, -1):
        st = 1 << (level - 1)
        if N == 1:
            n = _grid(shape[0], st)
            out += [flat(g * st) for g in range(1, n, 2)]
        elif N == 2:
            Dy, Dx = shape

            def ypass(xstep):
                n = _grid(Dy, st)
                inner, dfr = _row_pass_order(n, cubic)
                for g in inner + ([dfr] if dfr is not None else []):
                    out.extend(flat(g * st, x) for x in range(0, Dx, xstep))

            def xpass(ystep):
                n = _grid(Dx, st)
                for y in range(0, Dy, ystep):
                    out.extend(flat(y, g * st) for g in range(1, n, 2))

            if direction == 0:
                ypass(2 * st)
                xpass(st)
            else:
                xpass(2 * st)
                ypass(st)
        else:
            Dz, Dy, Dx = shape

            def zpass(ystep, xstep):
                n = _grid(Dz, st)
                inner, dfr = _row_pass_order(n, cubic)
                for y in range(0, Dy, ystep):
                    for g in inner:
                        out.extend(flat(g * st, y, x) for x in range(0, Dx, xstep))
                if dfr is not None:
                    for y in range(0, Dy, ystep):
                        out.extend(flat(dfr * st, y, x) for x in range(0, Dx, xstep))

            def ypass(zstep, xstep):
                n = _grid(Dy, st)
                inner, dfr = _row_pass_order(n, cubic)
                for z in range(0, Dz, zstep):
                    for g in inner:
                        out.extend(flat(z, g * st, x) for x in range(0, Dx, xstep))
                if dfr is not None:
                    for z in range(0, Dz, zstep):
                        out.extend(flat(z, dfr * st, x) for x in range(0, Dx, xstep))

            def xpass(zstep, ystep):
                n = _grid(Dx, st)
                for z in range(0, Dz, zstep):
                    for y in range(0, Dy, ystep):
                        out.extend(flat(z, y, g * st) for g in range(1, n, 2))

            if direction == 0:
                zpass(2 * st, 2 * st)
                ypass(st, 2 * st)
                xpass(st, st)
            else:
                xpass(2 * st, 2 * st)
                ypass(2 * st, st)
                zpass(st, st)
    return out, a


def anchors(shape, a):
    """Row-major anchor grid (step a in every dimension), as flat indices."""
    import itertools
    N = len(shape)
    strides_el = [1] * N
    for i in range(N - 2, -1, -1):
        strides_el[i] = strides_el[i + 1] * shape[i + 1]
    return [sum(c * s for c, s in zip(cs, strides_el)) for cs in itertools.product(*[range(0, d, a) for d in shape])]


def _line_order(n, cubic):
    """Grid indices interpolation_1d codes on one line of n grid points (IDH:374-419): odd indices
    upward (the last of an even count by its own rule); cubic with n >= 5: 3, 5, ... while i + 3 < n,
    then 1, then the first i with i + 3 >= n, then n - 1 for even n."""
    if n <= 1:
        return []
    if not cubic or n < 5:
        return list(range(1, n - 1, 2)) + ([n - 1] if n % 2 == 0 else [])
    bulk = list(range(3, n - 3, 2))
    last = bulk[-1] + 2 if bulk else 3
    return bulk + [1, last] + ([n - 1] if n % 2 == 0 else [])


def _fdf_targets(n, cubic):
    """Pass-axis index sets of the foreach boxes interpolation_1d_fastest_dim_first runs, in order
    (IDH:436-528): linear: odd 1..n-2, then n-1 for even n; cubic: odd 3..n-4, then the boundary rows
    1, n-2 (odd n > 3), n-3 (even n > 4), n-1 (even n > 2), one box each."""
    if n <= 1:
        return []
    if not cubic:
        return [list(range(1, n - 1, 2))] + ([[n - 1]] if n % 2 == 0 else [])
    boxes = [list(range(3, n - 3, 2))]
    bl = [1]
    if n % 2 == 1 and n > 3:
        bl.append(n - 2)
    if n % 2 == 0 and n > 4:
        bl.append(n - 3)
    if n % 2 == 0 and n > 2:
        bl.append(n - 1)
    return boxes + [[b] for b in bl]


def generic_order(shape, cubic, direction, anchor):
    """Code order of the generic traversal (integer elements, 4-D fields; IDH:2314-2519): per level,
    1-D one line; 2-D every 2*stride-th line along dims[0], then every stride-th line along dims[1];
    3-D / 4-D foreach boxes along dims[0], dims[1], ... with the axes not yet passed at step 2*stride
    and those passed at step stride, every box looping axis 0 outermost. dims is the direction-th
    permutation in lexicographic order."""
    import itertools
    N = len(shape)
    lv, a = levels_and_anchor(shape, anchor)
    el = [1] * N
    for i in range(N - 2, -1, -1):
        el[i] = el[i + 1] * shape[i + 1]
    dims = list(itertools.permutations(range(N)))[direction]
    out = [] if a else [0]
    for level in range(lv, 0, -1):
        st = 1 << (level - 1)
        if N == 1:
            out += [g * st for g in _line_order(_grid(shape[0], st), cubic)]
        elif N == 2:
            d0, d1 = dims
            for j in range(0, shape[d1], 2 * st):
                out += [j * el[d1] + g * st * el[d0] for g in _line_order(_grid(shape[d0], st), cubic)]
            for i in range(0, shape[d0], st):
                out += [i * el[d0] + g * st * el[d1] for g in _line_order(_grid(shape[d1], st), cubic)]
        else:
            step = {d: 2 * st for d in range(N)}
            for k, pd in enumerate(dims):
                if k > 0:
                    step[dims[k - 1]] = st
                for targets in _fdf_targets(_grid(shape[pd], st), cubic):
                    axes = [[g * st for g in targets] if d == pd else list(range(0, shape[d], step[d]))
                            for d in range(N)]
                    out += [sum(c * e for c, e in zip(cs, el)) for cs in itertools.product(*axes)]
    return out, a
