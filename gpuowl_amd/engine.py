"""Python mirror of the engine's host interface (the reference's Gpu / FFTConfig,
src/Gpu.h:213-300, src/FFTConfig.h:69-87) over the C-ABI in include/owl.h."""
import ctypes as C

from ._lib import GpuInfo, LLResult, OwlError, PrpResult, check, load

DATA, AUX, CHECK, BASE = 0, 1, 2, 3


def n_words(E):
    return (E - 1) // 32 + 1


def best_fit(E, spec=""):
    """FFTConfig::bestFit: (spec, N, maxExp) of the FFT chosen for exponent E."""
    lib = load()
    buf = C.create_string_buffer(64)
    N, mx = C.c_uint32(), C.c_uint32()
    check(lib.owl_fft_best_fit(E, (spec or "").encode(), buf, 64, C.byref(N), C.byref(mx)))
    return buf.value.decode(), N.value, mx.value


def words_to_int(words):
    return sum(int(w) << (32 * i) for i, w in enumerate(words))


def int_to_words(x, E):
    return [(x >> (32 * i)) & 0xFFFFFFFF for i in range(n_words(E))]


class Gpu:
    """One exponent on the GPU: Gpu(E, spec="", long_carry=False, line_block=0)."""

    def __init__(self, E, spec="", long_carry=False, line_block=0):
        self.lib = load()
        self.h = C.c_void_p()
        check(self.lib.owl_gpu_create(E, (spec or "").encode(), int(long_carry), line_block, C.byref(self.h)))
        info = GpuInfo()
        check(self.lib.owl_gpu_info(self.h, C.byref(info)))
        self.E, self.N, self.W, self.H, self.R = info.E, info.N, info.W, info.H, info.R
        self.nb, self.row_threads, self.col_threads = info.nb, info.rowThreads, info.colThreads
        self.spec, self.max_exp, self.long_carry = info.spec.decode(), info.maxExp, bool(info.longCarry)

    def close(self):
        if self.h:
            check(self.lib.owl_gpu_destroy(self.h))
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- state I/O
    def write_words(self, buf, words):
        arr = (C.c_uint32 * len(words))(*words)
        check(self.lib.owl_gpu_write_words(self.h, buf, arr, len(words)))

    def write_int(self, buf, x):
        self.write_words(buf, int_to_words(x, self.E))

    def read_words(self, buf=DATA):
        n = n_words(self.E)
        arr = (C.c_uint32 * n)()
        z = C.c_int()
        check(self.lib.owl_gpu_read_words(self.h, buf, arr, n, C.byref(z)))
        return [0] * n if z.value else list(arr)

    def read_int(self, buf=DATA):
        return words_to_int(self.read_words(buf))

    def write_digits(self, buf, digits):
        arr = (C.c_int32 * self.N)(*digits)
        check(self.lib.owl_gpu_write_digits(self.h, buf, arr, self.N))

    def read_digits(self, buf=DATA):
        arr = (C.c_int32 * self.N)()
        check(self.lib.owl_gpu_read_digits(self.h, buf, arr, self.N))
        return list(arr)

    def res64(self, buf=DATA):
        r = C.c_uint64()
        check(self.lib.owl_gpu_res64(self.h, buf, C.byref(r)))
        return r.value

    # ---- squaring path
    def square_loop(self, out, inp, frm, to, mul3_last=False, ll=False):
        check(self.lib.owl_gpu_square_loop(self.h, out, inp, frm, to, int(mul3_last), int(ll)))

    def mod_mul(self, io_a, in_b, mul3=False):
        check(self.lib.owl_gpu_mod_mul(self.h, io_a, in_b, int(mul3)))

    def do_check(self, block_size):
        ok = C.c_int()
        check(self.lib.owl_gpu_do_check(self.h, block_size, C.byref(ok)))
        return bool(ok.value)

    def is_equal(self, a, b):
        eq = C.c_int()
        check(self.lib.owl_gpu_is_equal(self.h, a, b, C.byref(eq)))
        return bool(eq.value)

    def copy(self, dst, src):
        check(self.lib.owl_gpu_copy(self.h, dst, src))

    def read_roe(self):
        n, mx, mean, sd = C.c_uint32(), C.c_double(), C.c_double(), C.c_double()
        check(self.lib.owl_gpu_read_roe(self.h, C.byref(n), C.byref(mx), C.byref(mean), C.byref(sd)))
        return {"n": n.value, "max": mx.value, "mean": mean.value, "sd": sd.value}

    # ---- workflows
    def prp(self, iters, block_size=1000, check_step=0, res_step=0):
        cap = iters // res_step if res_step else 0
        res = (C.c_uint64 * max(cap, 1))()
        out = PrpResult()
        check(self.lib.owl_gpu_prp(self.h, iters, block_size, check_step, res if res_step else None, res_step,
                                   cap, C.byref(out)))
        return out, list(res)[:cap]

    def ll(self, iters, res_step=0):
        cap = iters // res_step if res_step else 0
        res = (C.c_uint64 * max(cap, 1))()
        out = LLResult()
        check(self.lib.owl_gpu_ll(self.h, iters, res if res_step else None, res_step, cap, C.byref(out)))
        return out, list(res)[:cap]

    def time_prp(self, iters=1000, warmup=30, block_size=200):
        us = C.c_double()
        check(self.lib.owl_gpu_time_prp(self.h, iters, warmup, block_size, C.byref(us)))
        return us.value

    def time_squares(self, iters, warmup=30, use_graph=False):
        us = C.c_double()
        check(self.lib.owl_gpu_time_squares(self.h, iters, warmup, int(use_graph), C.byref(us)))
        return us.value

    # ---- kernel-level access (numerics tests)
    PASS_LEADIN, PASS_SQUARE, PASS_FUSED, PASS_CARRYOUT, PASS_SPECTRUM, PASS_MUL = range(6)

    def read_freq(self, idx=0):
        import numpy as np
        a = np.empty(self.N, dtype=np.float64)
        check(self.lib.owl_gpu_freq_io(self.h, idx, a.ctypes.data_as(C.POINTER(C.c_double)), self.N, 0))
        return a[0::2] + 1j * a[1::2]

    def write_freq(self, z, idx=0):
        import numpy as np
        a = np.empty(self.N, dtype=np.float64)
        a[0::2], a[1::2] = np.real(z), np.imag(z)
        check(self.lib.owl_gpu_freq_io(self.h, idx, a.ctypes.data_as(C.POINTER(C.c_double)), self.N, 1))

    def run_pass(self, which, arg=0):
        check(self.lib.owl_gpu_pass(self.h, which, arg))

    def time_kernels(self, iters):
        a, b = C.c_double(), C.c_double()
        check(self.lib.owl_gpu_time_kernels(self.h, iters, C.byref(a), C.byref(b)))
        return a.value, b.value


__all__ = ["Gpu", "best_fit", "OwlError", "DATA", "AUX", "CHECK", "BASE", "n_words", "words_to_int",
           "int_to_words"]
