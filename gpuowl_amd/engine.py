"""Python mirror of the engine's host interface (the reference's Gpu / FFTConfig,
src/Gpu.h:213-300, src/FFTConfig.h:69-87) over the C-ABI in include/owl.h."""
import array
import ctypes as C

from ._lib import GpuInfo, LLResult, OwlError, PrpResult, ResultRec, ServeInfo, ServeResult, TuneEntry, check, load

DATA, AUX, CHECK, BASE = 0, 1, 2, 3


def n_words(E):
    return (E - 1) // 32 + 1


def kernel_names():
    """Names of every compiled kernel (the runtime-plan ones and the shape-specialized ones)."""
    lib = load()
    need = C.c_size_t()
    lib.owl_hip_kernel_names(None, 0, C.byref(need))
    buf = C.create_string_buffer(need.value)
    check(lib.owl_hip_kernel_names(buf, need.value, C.byref(need)))
    return buf.value.decode().split(",")


def best_fit(E, spec=""):
    """FFTConfig::bestFit: (spec, N, maxExp) of the FFT this engine chooses for exponent E."""
    lib = load()
    buf = C.create_string_buffer(64)
    N, mx = C.c_uint32(), C.c_uint32()
    check(lib.owl_fft_best_fit(E, (spec or "").encode(), buf, 64, C.byref(N), C.byref(mx)))
    return buf.value.decode(), N.value, mx.value


def best_fit_ref(E, spec=""):
    """(spec, N, maxExp) of the FFT the reference chooses for exponent E (its limits, no tune file)."""
    lib = load()
    buf = C.create_string_buffer(64)
    N, mx = C.c_uint32(), C.c_uint32()
    check(lib.owl_fft_best_fit_ref(E, (spec or "").encode(), buf, 64, C.byref(N), C.byref(mx)))
    return buf.value.decode(), N.value, mx.value


def words_to_int(words):
    """The integer of little-endian 32-bit words (linear time, through bytes)."""
    return int.from_bytes(array.array("I", words).tobytes(), "little")


def int_to_words(x, E):
    """(E-1)/32+1 little-endian 32-bit words of 0 <= x < 2^(32*n) (linear time, through bytes)."""
    n = n_words(E)
    return array.array("I", x.to_bytes(4 * n, "little")).tolist()


class Gpu:
    """One exponent on the GPU: Gpu(E, spec="", long_carry=False, line_block=0, generic=False,
    col_one_region=False, col_one_pair=False)."""

    def __init__(self, E, spec="", long_carry=False, line_block=0, generic=False, col_one_region=False,
                 col_one_pair=False):
        self.lib = load()
        self.h = C.c_void_p()
        flags = (1 if long_carry else 0) | (2 if generic else 0) | (4 if col_one_region else 0) | \
            (8 if col_one_pair else 0)
        check(self.lib.owl_gpu_create(E, (spec or "").encode(), flags, line_block, C.byref(self.h)))
        info = GpuInfo()
        check(self.lib.owl_gpu_info(self.h, C.byref(info)))
        self.E, self.N, self.W, self.H, self.R = info.E, info.N, info.W, info.H, info.R
        self.nb, self.row_threads, self.col_threads = info.nb, info.rowThreads, info.colThreads
        self.spec, self.max_exp, self.long_carry = info.spec.decode(), info.maxExp, bool(info.longCarry)
        self.row_kernel, self.col_kernel = info.rowKernel.decode(), info.colKernel.decode()

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

    def read_digits_noalloc(self, buf=DATA):
        """read_digits staged through the library's static device array (no device allocation)."""
        arr = (C.c_int32 * self.N)()
        check(self.lib.owl_gpu_read_digits_noalloc(self.h, buf, arr, self.N))
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

    def measure_roe(self, iters=2000, warmup=50):
        """Round-off statistics of `iters` squarings (Gpu::measureROE) with the Gumbel Z."""
        n, mx, mean, sd, z = C.c_uint32(), C.c_double(), C.c_double(), C.c_double(), C.c_double()
        check(self.lib.owl_gpu_measure_roe(self.h, iters, warmup, C.byref(n), C.byref(mx), C.byref(mean), C.byref(sd),
                                           C.byref(z)))
        return {"n": n.value, "max": mx.value, "mean": mean.value, "sd": sd.value, "z": z.value}

    # ---- workflows
    def arm_fault(self, k):
        """The next prp() flips one data bit right after iteration k (failure-handling tests)."""
        check(self.lib.owl_gpu_arm_fault(self.h, k))

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

    def prp_proof(self, residue_dir, proof_path, power, block_size=1000, check_step=0):
        """The full PRP-3 test with a proof of the given power written (and verified) at proof_path."""
        out = PrpResult()
        check(self.lib.owl_gpu_prp_proof(self.h, str(residue_dir).encode(), power, block_size, check_step,
                                         str(proof_path).encode(), C.byref(out)))
        return out

    def proof_verify(self, proof_path):
        """(valid, is_prime) of a proof file."""
        ok, prime = C.c_int(), C.c_int()
        check(self.lib.owl_gpu_proof_verify(self.h, str(proof_path).encode(), C.byref(ok), C.byref(prime)))
        return bool(ok.value), bool(prime.value)

    def cert(self, cert_path, squarings, remove_file=False):
        """(sha3-hash as the results line reports it, res64) of `squarings` squarings of the start
        value in cert_path (the reference's M<E>.cert file)."""
        buf = C.create_string_buffer(80)
        r = C.c_uint64()
        check(self.lib.owl_gpu_cert(self.h, str(cert_path).encode(), squarings, int(remove_file), buf, 80,
                                    C.byref(r)))
        return buf.value.decode(), r.value

    def exp_mul(self, A, h, B, square_b=False):
        """A^h * B (B squared first if square_b) as words."""
        n = n_words(self.E)
        a, b, out = (C.c_uint32 * n)(*A), (C.c_uint32 * n)(*B), (C.c_uint32 * n)()
        check(self.lib.owl_gpu_exp_mul(self.h, a, h, b, int(square_b), out, n))
        return list(out)

    def prp_saved(self, directory, iters, block_size=1000, check_step=0, res_step=0, n_savefiles=20):
        """prp() resumed from / checkpointed to the reference-format savefiles in directory/<E>/."""
        cap = iters // res_step if res_step else 0
        res = (C.c_uint64 * max(cap, 1))()
        out = PrpResult()
        check(self.lib.owl_gpu_prp_saved(self.h, str(directory).encode(), n_savefiles, iters, block_size, check_step,
                                         res if res_step else None, res_step, cap, C.byref(out)))
        return out, list(res)[:cap]

    def ll_saved(self, directory, iters, save_step=10_000, res_step=0, n_savefiles=20):
        """ll() resumed from / saved to the reference-format savefiles in directory/ll-<E>/."""
        cap = iters // res_step if res_step else 0
        res = (C.c_uint64 * max(cap, 1))()
        out = LLResult()
        check(self.lib.owl_gpu_ll_saved(self.h, str(directory).encode(), n_savefiles, save_step, iters,
                                        res if res_step else None, res_step, cap, C.byref(out)))
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

    def time_copy(self, iters):
        us = C.c_double()
        check(self.lib.owl_gpu_time_copy(self.h, iters, C.byref(us)))
        return us.value

    def time_kernels(self, iters):
        a, b = C.c_double(), C.c_double()
        check(self.lib.owl_gpu_time_kernels(self.h, iters, C.byref(a), C.byref(b)))
        return a.value, b.value


class Batch:
    """Many exponents advanced together: Batch(exponents, spec="", long_carry=False, generic=False).
    Jobs of equal FFT geometry share every launch."""

    def __init__(self, exponents, spec="", long_carry=False, generic=False):
        self.lib = load()
        self.h = C.c_void_p()
        arr = (C.c_uint32 * len(exponents))(*exponents)
        flags = (1 if long_carry else 0) | (2 if generic else 0)
        check(self.lib.owl_batch_create(arr, len(exponents), (spec or "").encode(), flags, C.byref(self.h)))
        n, ng, tw = C.c_uint32(), C.c_uint32(), C.c_uint64()
        check(self.lib.owl_batch_info(self.h, C.byref(n), C.byref(ng), C.byref(tw)))
        self.n, self.n_groups, self.total_words = n.value, ng.value, tw.value
        self.exponents = list(exponents)

    def job_info(self, i):
        info = GpuInfo()
        check(self.lib.owl_batch_job_info(self.h, i, C.byref(info)))
        return {"E": info.E, "N": info.N, "spec": info.spec.decode(), "row": info.rowKernel.decode(),
                "col": info.colKernel.decode(), "R": info.R, "long_carry": bool(info.longCarry)}

    def arm_fault(self, job, k):
        check(self.lib.owl_batch_arm_fault(self.h, job, k))

    def close(self):
        if self.h:
            check(self.lib.owl_batch_destroy(self.h))
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

    def prp(self, iters, block_size=1000, check_step=0, res_step=0):
        cap = iters // res_step if res_step else 0
        res = (C.c_uint64 * max(1, self.n * cap))()
        out = (PrpResult * self.n)()
        check(self.lib.owl_batch_prp(self.h, iters, block_size, check_step, res if res_step else None, res_step, cap,
                                     out))
        rows = [list(res)[i * cap:(i + 1) * cap] for i in range(self.n)]
        return list(out), rows

    def ll(self, iters, res_step=0):
        cap = iters // res_step if res_step else 0
        res = (C.c_uint64 * max(1, self.n * cap))()
        out = (LLResult * self.n)()
        check(self.lib.owl_batch_ll(self.h, iters, res if res_step else None, res_step, cap, out))
        rows = [list(res)[i * cap:(i + 1) * cap] for i in range(self.n)]
        return list(out), rows

    def time_squares(self, iters, warmup=10, use_graph=False):
        us = C.c_double()
        check(self.lib.owl_batch_time_squares(self.h, iters, warmup, int(use_graph), C.byref(us)))
        return us.value


__all__ = ["Gpu", "Batch", "best_fit", "OwlError", "DATA", "AUX", "CHECK", "BASE", "n_words", "words_to_int",
           "int_to_words"]


class Server:
    """Request loop with continuous batching (include/owl.h, owl_server_*): submit PRP/LL
    requests at any time; each step() admits what fits, runs one round of block_size
    iterations on every active job and returns the results finished in it."""

    def __init__(self, mem_budget=0, max_jobs=256, block_size=1000, check_step=50_000, spec="", long_carry=False,
                 generic=False):
        self.lib = load()
        self.h = C.c_void_p()
        flags = (1 if long_carry else 0) | (2 if generic else 0)
        check(self.lib.owl_server_create(mem_budget, max_jobs, block_size, check_step, (spec or "").encode(), flags,
                                         C.byref(self.h)))
        self.block_size = block_size

    def submit(self, E, kind="prp", iters=None):
        ll = kind == "ll"
        if iters is None:
            iters = E - 2 if ll else E
        rid = C.c_uint32()
        check(self.lib.owl_server_submit(self.h, E, 1 if ll else 0, iters, C.byref(rid)))
        return rid.value

    def arm_fault(self, rid, k):
        """Failure-handling tests: flip a data bit of waiting PRP request rid after the round
        that passes iteration k."""
        check(self.lib.owl_server_arm_fault(self.h, rid, k))

    def step(self):
        n = C.c_uint32()
        check(self.lib.owl_server_step(self.h, C.byref(n)))
        return self.take()

    def take(self):
        out = []
        buf = (ServeResult * 64)()
        while True:
            n = C.c_uint32()
            check(self.lib.owl_server_take(self.h, buf, 64, C.byref(n)))
            for r in buf[:n.value]:
                d = {f: getattr(r, f) for f, _ in ServeResult._fields_}
                d["res2048"] = d["res2048"].decode()
                out.append(d)
            if n.value < 64:
                return out

    def info(self):
        i = ServeInfo()
        check(self.lib.owl_server_info(self.h, C.byref(i)))
        return {f: getattr(i, f) for f, _ in ServeInfo._fields_}

    def now(self):
        t = C.c_double()
        check(self.lib.owl_server_now(self.h, C.byref(t)))
        return t.value

    def drain(self, max_rounds=1_000_000):
        """Step until nothing is waiting or active; returns every result."""
        out = []
        for _ in range(max_rounds):
            i = self.info()
            if not i["waiting"] and not i["active"]:
                break
            out += self.step()
        return out

    def close(self):
        if self.h:
            check(self.lib.owl_server_destroy(self.h))
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


def tune(E, spec="", iters=200):
    """Time every engine configuration for exponent E at bestFit(E, spec); fastest first.
    Each entry: width, line_block, flags, us, W, H, R, long_carry, row, col."""
    lib = load()
    buf = (TuneEntry * 128)()
    n = C.c_uint32()
    check(lib.owl_tune_run(E, (spec or "").encode(), iters, buf, 128, C.byref(n)))
    return [{"width": e.width, "line_block": e.lineBlock, "flags": e.flags, "us": e.usPerIter, "W": e.W, "H": e.H,
             "R": e.R, "long_carry": bool(e.longCarry), "row": e.rowKernel.decode(), "col": e.colKernel.decode()}
            for e in buf[:n.value]]


def tune_batch(exponents, spec="", iters=20):
    """Time batches of these exponents under every W x H split per FFT length and record the fastest
    for owl_batch_create. Each entry: width, line_block, flags, us (per batch iteration), W, H, R,
    long_carry, row, col."""
    lib = load()
    arr = (C.c_uint32 * len(exponents))(*exponents)
    buf = (TuneEntry * 256)()
    n = C.c_uint32()
    check(lib.owl_tune_batch(arr, len(exponents), (spec or "").encode(), iters, buf, 256, C.byref(n)))
    return [{"width": e.width, "line_block": e.lineBlock, "flags": e.flags, "us": e.usPerIter, "W": e.W, "H": e.H,
             "R": e.R, "long_carry": bool(e.longCarry), "row": e.rowKernel.decode(), "col": e.colKernel.decode()}
            for e in buf[:n.value]]


def tune_batch_clear():
    check(load().owl_tune_batch_clear())


def tune_set(N, width, line_block, flags, us=0.0):
    check(load().owl_tune_set(N, width, line_block, flags, us))


def tune_get(N):
    """The recorded choice for FFT length N, or None."""
    w, r, f, us, found = C.c_uint32(), C.c_uint32(), C.c_int(), C.c_double(), C.c_int()
    check(load().owl_tune_get(N, C.byref(w), C.byref(r), C.byref(f), C.byref(us), C.byref(found)))
    return {"width": w.value, "line_block": r.value, "flags": f.value, "us": us.value} if found.value else None


def tune_clear():
    check(load().owl_tune_clear())


def tune_load(path):
    check(load().owl_tune_load(str(path).encode()))


def tune_save(path):
    check(load().owl_tune_save(str(path).encode()))


# ---------------------------------------------------------------- savefiles (reference formats)
def crc32(data):
    """CRC-32/IEEE of a bytes-like object, as the savefiles carry it."""
    buf = bytes(data)
    out = C.c_uint32()
    check(load().owl_crc32(C.c_char_p(buf), len(buf), C.byref(out)))
    return out.value


def savefile_write_prp(path, E, k, block_size, res64, n_errors, elapsed, check_words):
    w = (C.c_uint32 * len(check_words))(*check_words)
    check(load().owl_savefile_write_prp(str(path).encode(), E, k, block_size, res64, n_errors, elapsed, w,
                                        len(check_words)))


def savefile_read_prp(path):
    E, k, b, n = C.c_uint32(), C.c_uint32(), C.c_uint32(), C.c_uint32()
    r, t = C.c_uint64(), C.c_double()
    lib = load()
    p = str(path).encode()
    check(lib.owl_savefile_read_prp(p, C.byref(E), C.byref(k), C.byref(b), C.byref(r), C.byref(n), C.byref(t), None, 0))
    w = (C.c_uint32 * n_words(E.value))()
    check(lib.owl_savefile_read_prp(p, C.byref(E), C.byref(k), C.byref(b), C.byref(r), C.byref(n), C.byref(t), w,
                                    len(w)))
    return {"E": E.value, "k": k.value, "block_size": b.value, "res64": r.value, "n_errors": n.value,
            "elapsed": t.value, "check": list(w)}


def savefile_write_ll(path, E, k, elapsed, data_words):
    w = (C.c_uint32 * len(data_words))(*data_words)
    check(load().owl_savefile_write_ll(str(path).encode(), E, k, elapsed, w, len(data_words)))


def savefile_read_ll(path):
    E, k, t = C.c_uint32(), C.c_uint32(), C.c_double()
    lib = load()
    p = str(path).encode()
    check(lib.owl_savefile_read_ll(p, C.byref(E), C.byref(k), C.byref(t), None, 0))
    w = (C.c_uint32 * n_words(E.value))()
    check(lib.owl_savefile_read_ll(p, C.byref(E), C.byref(k), C.byref(t), w, len(w)))
    return {"E": E.value, "k": k.value, "elapsed": t.value, "data": list(w)}


# ---------------------------------------------------------------- the reference's tune.txt
def fft_tune_load(path):
    """Load a tune.txt ("cost spec [# maxExp]" lines): best_fit(E) then takes its first FFT covering E."""
    check(load().owl_fft_tune_load(str(path).encode()))


def fft_tune_save(path):
    check(load().owl_fft_tune_save(str(path).encode()))


def fft_tune_clear():
    check(load().owl_fft_tune_clear())


def fft_tune_add(cost, spec):
    """TuneEntry::update: returns whether the list changed."""
    upd = C.c_int()
    check(load().owl_fft_tune_add(cost, spec.encode(), C.byref(upd)))
    return bool(upd.value)


def fft_tune_list():
    """[(cost, spec, maxExp)] in order."""
    out, lib, i = [], load(), 0
    while True:
        cost, mx, found = C.c_double(), C.c_uint32(), C.c_int()
        spec = C.create_string_buffer(64)
        check(lib.owl_fft_tune_entry(i, C.byref(cost), spec, 64, C.byref(mx), C.byref(found)))
        if not found.value:
            return out
        out.append((cost.value, spec.value.decode(), mx.value))
        i += 1


# ---------------------------------------------------------------- worktodo / results (reference formats)
TASK_KINDS = {0: None, 1: "prp", 2: "ll", 3: "cert"}


def worktodo_parse(line):
    """One worktodo line -> {"kind": "prp"|"ll"|"cert", "exponent", "aid", "squarings"}, or None."""
    kind, E, sq = C.c_int(), C.c_uint32(), C.c_uint32()
    aid = C.create_string_buffer(40)
    check(load().owl_worktodo_parse(line.encode(), C.byref(kind), C.byref(E), aid, 40, C.byref(sq)))
    k = TASK_KINDS[kind.value]
    return None if k is None else {"kind": k, "exponent": E.value, "aid": aid.value.decode(), "squarings": sq.value}


def worktodo_best(lines):
    """Index of the task the reference would take first (smallest CERT, else smallest exponent), or -1."""
    arr = (C.c_char_p * max(1, len(lines)))(*[s.encode() for s in lines])
    idx = C.c_int32()
    check(load().owl_worktodo_best(arr, len(lines), C.byref(idx)))
    return idx.value


def prp_type1(words, E):
    """(res64, res2048) of x * 9^-1 mod 2^E - 1 (the PRP-3 type-1 residue at k = E)."""
    w = (C.c_uint32 * len(words))(*words)
    r = C.c_uint64()
    buf = C.create_string_buffer(520)
    check(load().owl_prp_type1(w, len(words), E, C.byref(r), buf, 520))
    return r.value, buf.value.decode()


def result_json(kind, E, is_prime=False, res64=0, fft_length=0, res2048="", n_errors=0, aid="", user="", uid="",
                program_name=None, program_version=None, proof_power=0, proof_md5="", sha3_hash="", squarings=0):
    """One results line in the reference's JSON layout; kind "prp", "ll" or "cert" (a PRP line
    carries the proof object when proof_power > 0; a CERT line its sha3_hash and squarings)."""
    enc = lambda s: s.encode() if s is not None else None  # noqa: E731
    r = ResultRec(kind={"prp": 1, "ll": 2, "cert": 3}[kind], exponent=E, isPrime=int(bool(is_prime)), res64=res64,
                  res2048=enc(res2048), fftLength=fft_length, nErrors=n_errors, proofPower=proof_power,
                  proofMd5=enc(proof_md5), sha3Hash=enc(sha3_hash), squarings=squarings, aid=enc(aid),
                  user=enc(user), uid=enc(uid), programName=enc(program_name), programVersion=enc(program_version))
    buf = C.create_string_buffer(4096)
    check(load().owl_result_json_ex(C.byref(r), buf, 4096))
    return buf.value.decode()


def proof_md5(path):
    """MD5 of a file as 32 lowercase hex digits (the proof object's md5 in a PRP results line)."""
    buf = C.create_string_buffer(40)
    check(load().owl_proof_md5(str(path).encode(), buf, 40))
    return buf.value.decode()


# ---------------------------------------------------------------- proofs
def proof_points(E, power):
    """The iterations whose residues a PRP run keeps for a proof of this power."""
    lib = load()
    n = C.c_uint32()
    check(lib.owl_proof_points(E, power, None, 0, C.byref(n)))
    buf = (C.c_uint32 * n.value)()
    check(lib.owl_proof_points(E, power, buf, n.value, C.byref(n)))
    return list(buf)


def root1(n, k):
    """The twiddle table entry exp(-2 pi i k / n) as (cos, sin) doubles (TrigBufCache's root1)."""
    c, s = C.c_double(), C.c_double()
    check(load().owl_trig_root1(n, k, C.byref(c), C.byref(s)))
    return c.value, s.value


def sha3_256(data):
    buf = bytes(data)
    out = C.create_string_buffer(32)
    check(load().owl_sha3_256(C.c_char_p(buf), len(buf), out))
    return out.raw
