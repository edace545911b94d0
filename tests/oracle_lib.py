"""ctypes bindings for the checkers: the C restatement (oracle/liboracle.so) and, where it was
built, the reference's own entropy coder / quantizer sources (oracle/_ref/libfalcons_ref.so).

Only tests, the smoke check and the CPU baseline of bench.py use this module. The product never
imports it."""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libfalcons_ref.so")

# element type codes of Falcons_FLOAT .. Falcons_INT64 (Config.hpp:29-38)
DTYPE_CODE = {np.dtype(np.float32): 0, np.dtype(np.float64): 1, np.dtype(np.uint8): 2, np.dtype(np.int8): 3,
              np.dtype(np.uint16): 4, np.dtype(np.int16): 5, np.dtype(np.uint32): 6, np.dtype(np.int32): 7,
              np.dtype(np.uint64): 8, np.dtype(np.int64): 9}


def dtype_code(dt):
    return DTYPE_CODE[np.dtype(dt)]


INT_CTYPES = {"u8": C.c_uint8, "i8": C.c_int8, "u16": C.c_uint16, "i16": C.c_int16, "u32": C.c_uint32,
              "i32": C.c_int32, "u64": C.c_uint64, "i64": C.c_int64}

_u8p = C.POINTER(C.c_uint8)
_i16p = C.POINTER(C.c_int16)
_szp = C.POINTER(C.c_size_t)


class FoConfig(C.Structure):
    _fields_ = [
        ("N", C.c_int32), ("dims", C.c_uint64 * 4), ("num", C.c_uint64),
        ("cmpr_algo", C.c_int32), ("eb_mode", C.c_int32),
        ("abs_eb", C.c_double), ("rel_eb", C.c_double), ("psnr_eb", C.c_double), ("l2norm_eb", C.c_double),
        ("lorenzo", C.c_int32), ("lorenzo2", C.c_int32), ("regression", C.c_int32), ("regression2", C.c_int32),
        ("openmp", C.c_int32), ("quantbin_cnt", C.c_int32), ("block_size", C.c_int32),
        ("pred_dim", C.c_int32), ("data_type", C.c_int32),
        ("interp_algo", C.c_int32), ("interp_direction", C.c_int32), ("interp_anchor_stride", C.c_int32),
        ("interp_alpha", C.c_double), ("interp_beta", C.c_double),
    ]


def _ptr(a, t):
    return a.ctypes.data_as(t)


class Oracle:
    def __init__(self, path=ORACLE_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} is not built (run make -C oracle)")
        L = C.CDLL(path)
        self.L = L
        L.fo_rle_encode.restype = C.c_size_t
        L.fo_rle_encode.argtypes = [_i16p, C.c_size_t, _u8p]
        L.fo_rle_decode.restype = C.c_int
        L.fo_rle_decode.argtypes = [_u8p, C.c_size_t, C.c_size_t, _i16p, _szp]
        L.fo_rle_preprocess.argtypes = [_i16p, C.c_size_t, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.fo_fse_compress_pub.restype = C.c_size_t
        L.fo_fse_compress_pub.argtypes = [_u8p, C.c_size_t, _u8p, C.c_size_t]
        L.fo_fse_decompress_pub.restype = C.c_size_t
        L.fo_fse_decompress_pub.argtypes = [_u8p, C.c_size_t, _u8p, C.c_size_t]
        L.fo_fse_normalize_pub.restype = C.c_size_t
        L.fo_fse_normalize_pub.argtypes = [C.POINTER(C.c_short), C.c_uint, C.POINTER(C.c_uint), C.c_size_t, C.c_uint]
        L.fo_interp_compress.restype = C.c_int
        L.fo_interp_compress.argtypes = [C.c_int, C.c_int, _szp, C.c_int, C.c_int, C.c_long, C.c_double, C.c_double,
                                         C.c_double, C.c_void_p, _i16p, _szp, C.c_void_p, C.c_size_t, _szp, _szp]
        L.fo_interp_decompress.restype = C.c_int
        L.fo_interp_decompress.argtypes = [C.c_int, C.c_int, _szp, C.c_int, C.c_int, C.c_long, C.c_double, C.c_double,
                                           C.c_double, _i16p, C.c_size_t, C.c_void_p, C.c_size_t, C.c_void_p]
        L.fo_interp_plan.argtypes = [_szp, C.c_int, C.c_long, _szp, C.POINTER(C.c_int)]
        L.fo_interp_scalar_f32.restype = C.c_float
        L.fo_interp_scalar_f32.argtypes = [C.c_int, C.c_float, C.c_float, C.c_float, C.c_float]
        L.fo_interp_scalar_f64.restype = C.c_double
        L.fo_interp_scalar_f64.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double]
        L.fo_quantize_point_f32.restype = C.c_int
        L.fo_quantize_point_f32.argtypes = [C.c_float, C.c_float, C.c_double, C.c_int, C.POINTER(C.c_float),
                                            C.POINTER(C.c_int)]
        L.fo_quantize_point_f64.restype = C.c_int
        L.fo_quantize_point_f64.argtypes = [C.c_double, C.c_double, C.c_double, C.c_int, C.POINTER(C.c_double),
                                            C.POINTER(C.c_int)]
        L.fo_config_default.argtypes = [C.POINTER(FoConfig), C.c_int, C.POINTER(C.c_uint64)]
        L.fo_config_save.restype = C.c_size_t
        L.fo_config_save.argtypes = [C.POINTER(FoConfig), _u8p]
        L.fo_config_load.restype = C.c_int
        L.fo_config_load.argtypes = [C.POINTER(FoConfig), _u8p, C.c_size_t, _szp]
        L.fo_compress.restype = C.c_size_t
        L.fo_compress.argtypes = [C.POINTER(FoConfig), C.c_int, C.c_void_p, _u8p, C.c_size_t, _u8p, C.c_size_t, _szp]
        L.fo_decompress.restype = C.c_int
        L.fo_decompress.argtypes = [_u8p, C.c_size_t, C.c_int, C.c_void_p, C.c_size_t, C.POINTER(FoConfig)]
        L.fo_size_bound.restype = C.c_size_t
        L.fo_size_bound.argtypes = [C.POINTER(FoConfig), C.c_int]
        L.fo_last_error.restype = C.c_char_p
        L.fo_data_range.restype = C.c_double
        L.fo_data_range.argtypes = [C.c_int, C.c_void_p, C.c_size_t]
        L.fo_set_simd.argtypes = [C.c_int]
        L.fo_tune_interp.restype = C.c_int
        L.fo_tune_interp.argtypes = [C.POINTER(FoConfig), C.c_int, C.c_void_p, C.c_int]
        L.fo_tune_report.restype = C.c_int
        L.fo_tune_report.argtypes = [C.POINTER(C.c_double), C.c_int]
        L.fo_tune_sample.restype = C.c_long
        L.fo_tune_sample.argtypes = [C.POINTER(FoConfig), C.c_int, C.c_void_p, C.POINTER(C.c_size_t),
                                     C.POINTER(C.c_int), C.POINTER(C.c_size_t), C.c_size_t,
                                     C.POINTER(C.c_size_t), C.POINTER(C.c_size_t), C.c_size_t]
        L.fo_o1_histogram.argtypes = [_i16p, C.c_size_t, C.POINTER(C.c_uint32), _i16p, _szp]
        L.fo_o1_ctx.restype = C.c_int
        L.fo_o1_sym.restype = C.c_int

    # ---- entropy coder ----
    def rle_encode(self, codes):
        codes = np.ascontiguousarray(codes, dtype=np.int16)
        n = codes.size
        out = np.zeros(3 * n + 64 * 1024 + 16 * 4096, dtype=np.uint8)
        ln = self.L.fo_rle_encode(_ptr(codes, _i16p), n, _ptr(out, _u8p))
        return out[:ln].tobytes()

    def rle_decode(self, blob, n):
        b = np.frombuffer(blob, dtype=np.uint8).copy()
        out = np.zeros(n + 32, dtype=np.int16)
        used = C.c_size_t(0)
        rc = self.L.fo_rle_decode(_ptr(b, _u8p), b.size, n, _ptr(out, _i16p), C.byref(used))
        if rc:
            raise ValueError("oracle entropy decode failed")
        return out[:n]

    def rle_preprocess(self, codes):
        codes = np.ascontiguousarray(codes, dtype=np.int16)
        m, d = C.c_int(0), C.c_int(0)
        self.L.fo_rle_preprocess(_ptr(codes, _i16p), codes.size, C.byref(m), C.byref(d))
        return m.value, d.value

    def fse_compress(self, src):
        src = np.ascontiguousarray(src, dtype=np.uint8)
        out = np.zeros(65536 + 1024, dtype=np.uint8)
        r = self.L.fo_fse_compress_pub(_ptr(out, _u8p), out.size, _ptr(src, _u8p), src.size)
        return r, out[: r if r > 1 else 0].tobytes()

    def fse_decompress(self, blob, n):
        b = np.frombuffer(blob, dtype=np.uint8).copy()
        out = np.zeros(n, dtype=np.uint8)
        self.L.fo_fse_decompress_pub(_ptr(out, _u8p), n, _ptr(b, _u8p), b.size)
        return out

    def fse_normalize(self, count, table_log, max_sym):
        cnt = np.ascontiguousarray(count, dtype=np.uint32)
        norm = np.zeros(256, dtype=np.int16)
        r = self.L.fo_fse_normalize_pub(norm.ctypes.data_as(C.POINTER(C.c_short)), table_log,
                                        cnt.ctypes.data_as(C.POINTER(C.c_uint)), int(cnt.sum()), max_sym)
        return r, norm[: max_sym + 1]

    # ---- decomposition ----
    def plan(self, dims, anchor):
        d = (C.c_size_t * 4)(*dims)
        a, lv = C.c_size_t(0), C.c_int(0)
        self.L.fo_interp_plan(d, len(dims), anchor, C.byref(a), C.byref(lv))
        return a.value, lv.value

    def interp_compress(self, data, cubic, direction, anchor, alpha, beta, eb):
        """Returns (codes in stream order, unpredictable list, mutated data, anchor used)."""
        data = np.ascontiguousarray(data).copy()
        dtype = dtype_code(data.dtype)
        dims = (C.c_size_t * 4)(*data.shape)
        n = data.size
        codes = np.zeros(n + 16, dtype=np.int16)
        unpred = np.zeros(n + 16, dtype=data.dtype)
        nq, nun, ao = C.c_size_t(0), C.c_size_t(0), C.c_size_t(0)
        rc = self.L.fo_interp_compress(dtype, data.ndim, dims, cubic, direction, anchor, alpha, beta, eb,
                                       data.ctypes.data, _ptr(codes, _i16p), C.byref(nq), unpred.ctypes.data,
                                       unpred.size, C.byref(nun), C.byref(ao))
        if rc:
            raise ValueError("oracle interp compress failed")
        return codes[: nq.value], unpred[: nun.value], data, ao.value

    def interp_decompress(self, shape, dtype, codes, unpred, cubic, direction, anchor, alpha, beta, eb):
        out = np.zeros(shape, dtype=dtype)
        dims = (C.c_size_t * 4)(*shape)
        codes = np.ascontiguousarray(codes, dtype=np.int16)
        unpred = np.ascontiguousarray(unpred, dtype=dtype)
        rc = self.L.fo_interp_decompress(dtype_code(dtype), len(shape), dims, cubic,
                                         direction, anchor, alpha, beta, eb, _ptr(codes, _i16p), codes.size,
                                         unpred.ctypes.data, unpred.size, out.ctypes.data)
        if rc:
            raise ValueError("oracle interp decompress failed")
        return out

    # ---- whole stream ----
    def config(self, shape, **kw):
        c = FoConfig()
        dims = (C.c_uint64 * 4)(*shape)
        self.L.fo_config_default(C.byref(c), len(shape), dims)
        for k, v in kw.items():
            setattr(c, k, v)
        return c

    def gen_compress(self, data, cubic, direction, anchor, alpha, beta, eb):
        """The generic traversal alone (fo_gen_compress): (codes, unpredictables, anchor used)."""
        data = np.ascontiguousarray(data).copy()
        dims = (C.c_size_t * 4)(*data.shape)
        codes = np.zeros(data.size + 16, dtype=np.int16)
        un = np.zeros(data.size + 16, dtype=data.dtype)
        nq, nun, au = C.c_size_t(0), C.c_size_t(0), C.c_size_t(0)
        self.L.fo_gen_compress.restype = C.c_int
        self.L.fo_gen_compress.argtypes = [C.c_int, C.c_int, _szp, C.c_int, C.c_int, C.c_long, C.c_double,
                                           C.c_double, C.c_double, C.c_void_p, _i16p, _szp, C.c_void_p, C.c_size_t,
                                           _szp, _szp]
        rc = self.L.fo_gen_compress(dtype_code(data.dtype), data.ndim, dims, cubic, direction, anchor, alpha, beta,
                                    eb, data.ctypes.data, _ptr(codes, _i16p), C.byref(nq), un.ctypes.data,
                                    un.size, C.byref(nun), C.byref(au))
        assert rc == 0, rc
        return codes[: nq.value], un[: nun.value], au.value

    def tune_sample(self, data, conf):
        """The parameter search's sampler: (block side, profiling branch, profiled starts, selected
        starts), starts as (count, 3) arrays; None when the field is not tuned."""
        data = np.ascontiguousarray(data)
        dtype = dtype_code(data.dtype)
        cap = data.size // 8 + 16
        st = np.zeros((cap, 4), dtype=np.uint64)
        sel = np.zeros((cap, 4), dtype=np.uint64)
        sbs, prof, nst = C.c_size_t(0), C.c_int(0), C.c_size_t(0)
        sp = C.POINTER(C.c_size_t)
        n = self.L.fo_tune_sample(C.byref(conf), dtype, data.ctypes.data, C.byref(sbs), C.byref(prof),
                                  st.ctypes.data_as(sp), cap, C.byref(nst), sel.ctypes.data_as(sp), cap)
        if n < 0:
            return None
        assert nst.value <= cap and n <= cap
        return sbs.value, prof.value, st[: nst.value].copy(), sel[:n].copy()

    def compress(self, data, conf, want_pre=False):
        data = np.ascontiguousarray(data).copy()
        dtype = dtype_code(data.dtype)
        cap = self.L.fo_size_bound(C.byref(conf), dtype)
        out = np.zeros(cap, dtype=np.uint8)
        pre = np.zeros(cap + data.nbytes // 2 + (1 << 20), dtype=np.uint8) if want_pre else None
        pl = C.c_size_t(0)
        ln = self.L.fo_compress(C.byref(conf), dtype, data.ctypes.data, _ptr(out, _u8p), cap,
                                _ptr(pre, _u8p) if want_pre else None, pre.size if want_pre else 0, C.byref(pl))
        if ln == 0:
            raise ValueError(self.L.fo_last_error().decode())
        res = out[:ln].tobytes()
        if want_pre:
            return res, data, pre[: pl.value].tobytes()
        return res, data

    def decompress(self, blob, dtype, num=None):
        b = np.frombuffer(blob, dtype=np.uint8).copy()
        conf = FoConfig()
        if num is None:
            hl = C.c_size_t(0)
            if self.L.fo_config_load(C.byref(conf), _ptr(b, _u8p), b.size, C.byref(hl)):
                raise ValueError(self.L.fo_last_error().decode())
            num = conf.num
        out = np.zeros(num, dtype=dtype)
        rc = self.L.fo_decompress(_ptr(b, _u8p), b.size, dtype_code(dtype), out.ctypes.data,
                                  num, C.byref(conf))
        if rc:
            raise ValueError(self.L.fo_last_error().decode())
        shape = tuple(int(conf.dims[i]) for i in range(conf.N))
        return out.reshape(shape), conf


class Ref:
    """The reference's own sources (entropy coder, FSE, scalar quantizer, interpolators)."""

    def __init__(self, path=REF_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        L = C.CDLL(path)
        self.L = L
        L.ref_rlefse_encode.restype = C.c_long
        L.ref_rlefse_encode.argtypes = [_i16p, C.c_size_t, _u8p, C.c_size_t]
        L.ref_rlefse_decode.restype = C.c_int
        L.ref_rlefse_decode.argtypes = [_u8p, C.c_size_t, C.c_size_t, _i16p]
        L.ref_fse_compress.restype = C.c_size_t
        L.ref_fse_compress.argtypes = [_u8p, C.c_size_t, _u8p, C.c_size_t]
        L.ref_fse_iserror.restype = C.c_uint
        L.ref_fse_iserror.argtypes = [C.c_size_t]
        L.ref_fse_normalize.restype = C.c_size_t
        L.ref_fse_normalize.argtypes = [C.POINTER(C.c_short), C.c_uint, C.POINTER(C.c_uint), C.c_size_t, C.c_uint]
        L.ref_quantize_f32.restype = C.c_int
        L.ref_quantize_f32.argtypes = [C.c_float, C.c_float, C.c_double, C.c_int, C.POINTER(C.c_float),
                                       C.POINTER(C.c_int)]
        L.ref_quantize_f64.restype = C.c_int
        L.ref_quantize_f64.argtypes = [C.c_double, C.c_double, C.c_double, C.c_int, C.POINTER(C.c_double),
                                       C.POINTER(C.c_int)]
        L.ref_recover_f32.restype = C.c_float
        L.ref_recover_f32.argtypes = [C.c_float, C.c_int, C.c_double]
        L.ref_recover_f64.restype = C.c_double
        L.ref_recover_f64.argtypes = [C.c_double, C.c_int, C.c_double]
        L.ref_interp_f32.restype = C.c_float
        L.ref_interp_f32.argtypes = [C.c_int, C.c_float, C.c_float, C.c_float, C.c_float]
        L.ref_interp_f64.restype = C.c_double
        L.ref_interp_f64.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double]
        L.ref_sample.restype = C.c_long
        L.ref_sample.argtypes = [C.c_int, C.c_int, C.POINTER(C.c_size_t), C.c_void_p, C.c_size_t, C.c_double,
                                 C.c_double, C.c_int, C.POINTER(C.c_size_t), C.c_size_t, C.POINTER(C.c_size_t),
                                 C.c_void_p, C.c_size_t]

    def _int_fns(self, tag):
        ct = INT_CTYPES[tag]
        q = getattr(self.L, f"ref_quantize_{tag}")
        q.restype, q.argtypes = C.c_int, [ct, ct, C.c_double, C.POINTER(ct), C.POINTER(C.c_int)]
        r = getattr(self.L, f"ref_recover_{tag}")
        r.restype, r.argtypes = C.c_int, [ct, C.c_int, C.c_double, C.POINTER(ct)]
        i = getattr(self.L, f"ref_interp_{tag}")
        i.restype, i.argtypes = ct, [C.c_int, ct, ct, ct, ct]
        return ct, q, r, i

    def quantize_int(self, tag, data, pred, eb):
        """LinearQuantizer<T>::quantize_and_overwrite: (code, overwritten value, escaped)."""
        ct, q, _, _ = self._int_fns(tag)
        out, un = ct(0), C.c_int(0)
        code = q(ct(int(data)), ct(int(pred)), float(eb), C.byref(out), C.byref(un))
        return code, out.value, un.value

    def recover_int(self, tag, pred, code, eb):
        """LinearQuantizer<T>::recover_pred: (fits, value)."""
        ct, _, r, _ = self._int_fns(tag)
        out = ct(0)
        ok = r(ct(int(pred)), int(code), float(eb), C.byref(out))
        return ok, out.value if ok else 0

    def interp_int(self, tag, kind, a, b, c, d):
        ct, _, _, i = self._int_fns(tag)
        return i(kind, ct(int(a)), ct(int(b)), ct(int(c)), ct(int(d)))

    def sample(self, data, sbs, abseb, rate=0.005, profiling=-1):
        """profiling_block + sampleBlocks of the reference (Sample.hpp): (profiling branch, profiled
        starts as a (count, 4) array, the sampled blocks concatenated)."""
        data = np.ascontiguousarray(data)
        dtype = dtype_code(data.dtype)
        dims = (C.c_size_t * 4)(*data.shape)
        cap = data.size // 8 + 16
        st = np.zeros((cap, 4), dtype=np.uint64)
        blocks = np.zeros(data.size // 4 + (sbs + 1) ** data.ndim, dtype=data.dtype)
        nst = C.c_size_t(0)
        sp = C.POINTER(C.c_size_t)
        r = self.L.ref_sample(dtype, data.ndim, dims, data.ctypes.data, sbs, abseb, rate, profiling,
                              st.ctypes.data_as(sp), cap, C.byref(nst), blocks.ctypes.data, blocks.size)
        assert r >= 0 and nst.value <= cap, r
        nb = r >> 1
        return r & 1, st[: nst.value].copy(), blocks[: nb * (sbs + 1) ** data.ndim].copy()

    def rle_encode(self, codes):
        codes = np.ascontiguousarray(codes, dtype=np.int16)
        out = np.zeros(3 * codes.size + 64 * 1024 + 16 * 4096, dtype=np.uint8)
        ln = self.L.ref_rlefse_encode(_ptr(codes, _i16p), codes.size, _ptr(out, _u8p), out.size)
        assert ln >= 0
        return out[:ln].tobytes()

    def rle_decode(self, blob, n):
        b = np.frombuffer(blob, dtype=np.uint8).copy()
        out = np.zeros(n, dtype=np.int16)
        self.L.ref_rlefse_decode(_ptr(b, _u8p), b.size, n, _ptr(out, _i16p))
        return out

    def fse_compress(self, src):
        src = np.ascontiguousarray(src, dtype=np.uint8)
        out = np.zeros(65536 + 1024, dtype=np.uint8)
        r = self.L.ref_fse_compress(_ptr(out, _u8p), out.size, _ptr(src, _u8p), src.size)
        if self.L.ref_fse_iserror(r):
            return -1, b""
        return r, out[: r if r > 1 else 0].tobytes()

    def fse_normalize(self, count, table_log, max_sym):
        cnt = np.ascontiguousarray(count, dtype=np.uint32)
        norm = np.zeros(256, dtype=np.int16)
        r = self.L.ref_fse_normalize(norm.ctypes.data_as(C.POINTER(C.c_short)), table_log,
                                     cnt.ctypes.data_as(C.POINTER(C.c_uint)), int(cnt.sum()), max_sym)
        return r, norm[: max_sym + 1]


    def quantize_f32(self, ori, pred, eb):
        """LinearQuantizer<float>::quantize_and_overwrite after set_eb: (code, value left, escaped)."""
        o, u = C.c_float(0), C.c_int(0)
        q = self.L.ref_quantize_f32(float(ori), float(pred), float(eb), 1, C.byref(o), C.byref(u))
        return q, o.value, u.value

    def quantize_f64(self, ori, pred, eb):
        o, u = C.c_double(0), C.c_int(0)
        q = self.L.ref_quantize_f64(float(ori), float(pred), float(eb), 1, C.byref(o), C.byref(u))
        return q, o.value, u.value

    def recover_f32(self, pred, code, eb):
        return self.L.ref_recover_f32(float(pred), int(code), float(eb))

    def recover_f64(self, pred, code, eb):
        return self.L.ref_recover_f64(float(pred), int(code), float(eb))

    def interp_f32(self, kind, a, b, c, d):
        return self.L.ref_interp_f32(kind, float(a), float(b), float(c), float(d))

    def interp_f64(self, kind, a, b, c, d):
        return self.L.ref_interp_f64(kind, float(a), float(b), float(c), float(d))


def oracle():
    return Oracle()


def ref_or_none():
    try:
        return Ref()
    except (FileNotFoundError, OSError):
        return None
