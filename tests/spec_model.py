"""A per-target model of the interpolation decomposition written from its specification, for small
fields: for every coded point it decides the predictor formula, its neighbours, the compare class and
whether the reconstruction is written back, then quantizes through the single-point functions.

The two leaf operations it calls are pinned elsewhere against the reference's own build: the
interpolator formulas (fo_interp_scalar_*) and the point quantizer (fo_quantize_point_*).
What this module restates independently of the oracle's loops is the placement:

Row passes (InterpolationDecomposition.hpp 3d_z :819, 3d_y :1138, 2d_y :604), grid count n:
  linear  g odd < n-1: lin(g-1, g+1); g = n-1 (n even): n < 3 -> eq(n-2), else
          linear1(n-1-D, n-2) with D = 2, or 3 on the last pass of level 1 (:852, :942)
  cubic   g = 1: q1(0,2,4) if n >= 5, lin(0,2) if n in {3,4}, eq(0) if n = 2; 3 <= g < n-3: cubic;
          g = n-2 (n odd > 3): q2(n-5, n-3, n-1); g = n-3 (n even > 4): q2(n-6, n-4, n-2);
          g = n-1 (n even > 2): linear1(n-4, n-2), after the whole outer loop in 3-D (scalar
          compare class there, :1095-1132, :1335-1372), in order in 2-D (vector class, :755-765)
Line passes (Interpolation_quantizer.inl :32-180 linear, :184-377 cubic), odd_len = n/2,
even_len = n - odd_len:
  linear  index 2t+1, t < odd_len-1: lin(2t, 2t+2); last odd index 2 odd_len - 1: eq(2(even_len-1))
          if n < 3 else linear1(2(even_len-2), 2(even_len-1)); vector class for
          t < 8 #{i = 0, 8, ..: i + 8 < odd_len}
  cubic   index 1: eq(0) / lin(0,2) / q1(0,2,4) by even_len < 2 / < 3 / else (scalar class);
          index 2t+3 while t+3 < even_len: cubic(2t, .., 2t+6), vector class for
          t < 8 #{i = 0, 8, ..: i + 10 < even_len}; n odd (odd_len > 1): index n-2 q2(n-5, n-3, n-1);
          n even: index n-3 q2(n-6, n-4, n-2) if odd_len > 2, then index n-1 linear1(n-4, n-2)
Quantizer: LinearQuantizer.hpp:49-90; the first point without anchors uses the constructor's
reciprocal 0.5/eb (:31), every level calls set_eb (reciprocal 1/(2 eb), :39-43, IDH:176-195).
"""
import ctypes as C

import numpy as np

import traversal_spec as spec

LIN, LIN1, Q1, Q2, CUB, EQ = 0, 1, 2, 3, 4, 5


class Model:
    def __init__(self, orc, dtype, simd):
        self.L = orc.L
        self.dt = np.dtype(dtype)
        self.simd = simd
        if self.dt == np.float32:
            self.interp, self.qp, self.ct = self.L.fo_interp_scalar_f32, self.L.fo_quantize_point_f32, C.c_float
        else:
            self.interp, self.qp, self.ct = self.L.fo_interp_scalar_f64, self.L.fo_quantize_point_f64, C.c_double

    def pred(self, kind, x):
        if kind == EQ:
            return x[0]
        args = list(x) + [0.0] * (4 - len(x))
        return self.dt.type(self.interp(kind, *[float(a) for a in args]))

    def run(self, data, cubic, direction, anchor, eb, alpha=1.0, beta=1.0):
        d = np.array(data, dtype=self.dt).reshape(-1).copy()
        shape = data.shape
        N = len(shape)
        lv, a = spec.levels_and_anchor(shape, anchor)
        el = [1] * N
        for i in range(N - 2, -1, -1):
            el[i] = el[i + 1] * shape[i + 1]
        codes, unpred = [], []
        if a:
            unpred += [d[i] for i in spec.anchors(shape, a)]
        else:
            codes.append(self.first_point(d, eb, unpred))
        self.d, self.codes, self.unpred = d, codes, unpred
        for level in range(lv, 0, -1):
            st = 1 << (level - 1)
            self.cur_eb = level_bound(eb, level, alpha, beta)
            passes = self.passes(N, direction, st)
            for k, (kind, ax, steps) in enumerate(passes):
                skip = level == 1 and k == len(passes) - 1
                if kind == "row":
                    self.row_pass(shape, el, ax, steps, st, cubic, skip, N)
                else:
                    self.line_pass(shape, el, ax, steps, st, cubic, skip)
        return np.array(codes, np.int16), np.array(unpred, self.dt), d.reshape(shape)

    def first_point(self, d, eb, unpred):
        # quantize_and_overwrite(data[0], 0) before any set_eb: reciprocal 0.5 / eb
        ori = d[0]
        diff = self.dt.type(ori - self.dt.type(0))
        q = np.rint(np.float64(diff) * (0.5 / eb))
        if not (-32768 < q < 32768):
            unpred.append(ori)
            return -32768
        dec = self.dt.type(np.float64(q) * (2 * eb))  # fma(q, 2eb, 0) rounds once, like the product
        err = self.dt.type(dec - ori)
        if -eb <= np.float64(err) <= eb:
            d[0] = dec
            return int(q)
        unpred.append(ori)
        return -32768

    @staticmethod
    def passes(N, direction, st):
        """(kind, pass axis, {other axis: step}) in execution order."""
        if N == 1:
            return [("line", 0, {})]
        if N == 2:
            y, x = ("row", 0, {1: 2 * st}), ("line", 1, {0: st})
            if direction == 0:
                return [y, x]
            return [("line", 1, {0: 2 * st}), ("row", 0, {1: st})]
        if direction == 0:
            return [("row", 0, {1: 2 * st, 2: 2 * st}), ("row", 1, {0: st, 2: 2 * st}), ("line", 2, {0: st, 1: st})]
        return [("line", 2, {0: 2 * st, 1: 2 * st}), ("row", 1, {0: 2 * st, 2: st}), ("row", 0, {1: st, 2: st})]

    def point(self, pos, pred, vclass, skip):
        out, esc = self.ct(0), C.c_int(0)
        ori = self.d[pos]
        q = self.qp(ori, pred, self.cur_eb, 1 if (vclass and self.simd) else 0, C.byref(out), C.byref(esc))
        self.codes.append(q)
        if q == -32768:
            self.unpred.append(ori)
        elif not skip:
            self.d[pos] = out.value

    def row_pass(self, shape, el, ax, steps, st, cubic, skip, N):
        n = (shape[ax] - 1) // st + 1
        others = sorted(steps)  # outer axis first, then the row (fastest remaining) axis
        outer, inner = others[0], others[-1]
        outer_vals = range(0, shape[outer], steps[outer]) if N == 3 else [0]
        row_vals = range(0, shape[inner], steps[inner])

        def at(g, o, r):
            c = [0] * len(shape)
            c[ax] = g * st
            c[inner] = r
            if N == 3:
                c[outer] = o
            return sum(ci * e for ci, e in zip(c, el))

        def emit(g, o, kind, nb, vclass):
            for r in row_vals:
                self.point(at(g, o, r), self.pred(kind, [self.d[at(h, o, r)] for h in nb]), vclass, skip)

        deferred = []
        for o in outer_vals:
            for g in range(1, n, 2):
                if not cubic:
                    if g < n - 1:
                        emit(g, o, LIN, [g - 1, g + 1], True)
                    elif n < 3:
                        emit(g, o, EQ, [n - 2], True)
                    else:
                        dl = 3 if skip else 2
                        emit(g, o, LIN1, [n - 1 - dl, n - 2], True)
                    continue
                if g == 1 and (n == 2 or g < n - 1):
                    if n >= 5:
                        emit(g, o, Q1, [0, 2, 4], True)
                    elif n >= 3:
                        emit(g, o, LIN, [0, 2], True)
                    else:
                        emit(g, o, EQ, [0], True)
                elif 3 <= g < n - 3:
                    emit(g, o, CUB, [g - 3, g - 1, g + 1, g + 3], True)
                elif n % 2 == 1 and n > 3 and g == n - 2:
                    emit(g, o, Q2, [n - 5, n - 3, n - 1], True)
                elif n % 2 == 0 and n > 4 and g == n - 3:
                    emit(g, o, Q2, [n - 6, n - 4, n - 2], True)
                elif n % 2 == 0 and n > 2 and g == n - 1:
                    if N == 2:
                        emit(g, o, LIN1, [n - 4, n - 2], True)
                    else:
                        deferred.append(o)
                else:
                    raise AssertionError(f"no cubic rule for g={g} n={n}")
        for o in deferred:
            emit(n - 1, o, LIN1, [n - 4, n - 2], False)

    def line_pass(self, shape, el, ax, steps, st, cubic, skip):
        n = (shape[ax] - 1) // st + 1
        if n <= 1:
            return
        import itertools
        others = sorted(steps)
        ranges = [range(0, shape[a], steps[a]) for a in others]
        odd_len, even_len = n // 2, n - n // 2
        for combo in itertools.product(*ranges):
            base = sum(c * el[a] for a, c in zip(others, combo))
            x = lambda i: self.d[base + i * st * el[ax]]  # noqa: E731

            def put(i, kind, nb, vclass):
                self.point(base + i * st * el[ax], self.pred(kind, [x(h) for h in nb]), vclass, skip)

            if not cubic:
                mv = 8 * len([i for i in range(0, odd_len, 8) if i + 8 < odd_len])
                for t in range(odd_len - 1):
                    put(2 * t + 1, LIN, [2 * t, 2 * t + 2], t < mv)
                last = 2 * odd_len - 1
                if n < 3:
                    put(last, EQ, [2 * (even_len - 1)], False)
                else:
                    put(last, LIN1, [2 * (even_len - 2), 2 * (even_len - 1)], False)
                continue
            if even_len < 2:
                put(1, EQ, [0], False)
            elif even_len < 3:
                put(1, LIN, [0, 2], False)
            else:
                put(1, Q1, [0, 2, 4], False)
            mv = 8 * len([i for i in range(0, even_len, 8) if i + 10 < even_len])
            t = 0
            while t + 3 < even_len:
                put(2 * t + 3, CUB, [2 * t, 2 * t + 2, 2 * t + 4, 2 * t + 6], t < mv)
                t += 1
            if odd_len > 1:
                if odd_len < even_len:
                    put(2 * odd_len - 1, Q2, [2 * (even_len - 3), 2 * (even_len - 2), 2 * (even_len - 1)], False)
                else:
                    if odd_len > 2:
                        put(2 * odd_len - 3, Q2, [2 * (even_len - 3), 2 * (even_len - 2), 2 * (even_len - 1)],
                            False)
                    put(2 * odd_len - 1, LIN1, [2 * (even_len - 2), 2 * (even_len - 1)], False)


def level_bound(eb, level, alpha, beta):
    """Bound of one level (InterpolationDecomposition.hpp:176-195)."""
    if alpha < 0:
        return eb * 0.5 if level >= 3 else eb
    if alpha >= 1:
        return eb / min(alpha ** (level - 1), beta)
    return eb


def model_compress(orc, data, cubic, direction, anchor, eb, simd, alpha=1.0, beta=1.0):
    """Codes, unpredictable list and mutated field of the specified decomposition."""
    return Model(orc, data.dtype, simd).run(data, cubic, direction, anchor, eb, alpha, beta)
