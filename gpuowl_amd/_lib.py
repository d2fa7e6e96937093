"""ctypes binding of libgpuowl_amd.so (include/owl.h, include/owl_hip.h).

The library is built in-tree by gpuowl_amd/build.py. There is no fallback:
if it is missing, importing the engine raises."""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgpuowl_amd.so")


class OwlError(RuntimeError):
    pass


class GpuInfo(C.Structure):
    _fields_ = [("E", C.c_uint32), ("N", C.c_uint32), ("W", C.c_uint32), ("H", C.c_uint32), ("R", C.c_uint32),
                ("nb", C.c_uint32), ("rowThreads", C.c_uint32), ("colThreads", C.c_uint32),
                ("maxExp", C.c_uint32), ("longCarry", C.c_int), ("spec", C.c_char * 64),
                ("rowKernel", C.c_char * 32), ("colKernel", C.c_char * 32)]


class PrpResult(C.Structure):
    _fields_ = [("res64", C.c_uint64), ("k", C.c_uint32), ("nErrors", C.c_uint32), ("nChecks", C.c_uint32),
                ("isPrime", C.c_int), ("maxRoe", C.c_double), ("secsPerIt", C.c_double)]


class LLResult(C.Structure):
    _fields_ = [("res64", C.c_uint64), ("k", C.c_uint32), ("isPrime", C.c_int), ("secsPerIt", C.c_double)]


class ServeResult(C.Structure):
    _fields_ = [("id", C.c_uint32), ("E", C.c_uint32), ("iters", C.c_uint32), ("ll", C.c_int),
                ("res64", C.c_uint64), ("isPrime", C.c_int), ("nErrors", C.c_uint32), ("nChecks", C.c_uint32),
                ("maxRoe", C.c_double), ("tSubmit", C.c_double), ("tAdmit", C.c_double), ("tDone", C.c_double),
                ("status", C.c_int), ("res64Type1", C.c_uint64), ("res2048", C.c_char * 516)]


class ServeInfo(C.Structure):
    _fields_ = [("waiting", C.c_uint32), ("active", C.c_uint32), ("groups", C.c_uint32), ("done", C.c_uint32),
                ("rounds", C.c_uint32), ("bytesInUse", C.c_uint64), ("budget", C.c_uint64),
                ("iterations", C.c_uint64), ("wordIterations", C.c_uint64)]


class ResultRec(C.Structure):
    """owl_result_t (include/owl.h)."""
    _fields_ = [("kind", C.c_int), ("exponent", C.c_uint32), ("isPrime", C.c_int), ("res64", C.c_uint64),
                ("res2048", C.c_char_p), ("fftLength", C.c_uint32), ("nErrors", C.c_uint32),
                ("proofPower", C.c_uint32), ("proofMd5", C.c_char_p), ("sha3Hash", C.c_char_p),
                ("squarings", C.c_uint32), ("aid", C.c_char_p), ("user", C.c_char_p), ("uid", C.c_char_p),
                ("programName", C.c_char_p), ("programVersion", C.c_char_p)]


class TuneEntry(C.Structure):
    _fields_ = [("width", C.c_uint32), ("lineBlock", C.c_uint32), ("flags", C.c_int), ("usPerIter", C.c_double),
                ("W", C.c_uint32), ("H", C.c_uint32), ("R", C.c_uint32), ("longCarry", C.c_int),
                ("rowKernel", C.c_char * 32), ("colKernel", C.c_char * 32)]


u32, i32, u64, dbl, vp, cp = C.c_uint32, C.c_int, C.c_uint64, C.c_double, C.c_void_p, C.c_char_p
P = C.POINTER

# name -> argtypes; every function returns int status
SIGNATURES = {
    "owl_hip_device_count": [P(i32)],
    "owl_hip_set_device": [i32],
    "owl_hip_device_info": [i32, cp, C.c_size_t, P(C.c_size_t), P(C.c_size_t), P(i32), P(i32)],
    "owl_hip_alloc": [P(vp), C.c_size_t],
    "owl_hip_free": [vp],
    "owl_hip_queue_create": [P(vp)],
    "owl_hip_queue_destroy": [vp],
    "owl_hip_finish": [vp],
    "owl_hip_write": [vp, vp, vp, C.c_size_t, i32],
    "owl_hip_read": [vp, vp, vp, C.c_size_t, i32],
    "owl_hip_copy": [vp, vp, vp, C.c_size_t],
    "owl_hip_fill32": [vp, vp, u32, C.c_size_t],
    "owl_hip_kernel_get": [cp, P(vp)],
    "owl_hip_kernel_names": [cp, C.c_size_t, P(C.c_size_t)],
    "owl_hip_kernel_max_threads": [vp, P(u32)],
    "owl_hip_launch": [vp, vp, u32, u32, u32, u32, P(vp)],
    "owl_hip_event_create": [P(vp)],
    "owl_hip_event_record": [vp, vp],
    "owl_hip_event_elapsed_ms": [vp, vp, P(C.c_float)],
    "owl_hip_event_destroy": [vp],
    "owl_hip_graph_begin": [vp],
    "owl_hip_graph_end": [vp, P(vp)],
    "owl_hip_graph_launch": [vp, vp],
    "owl_hip_graph_destroy": [vp],
    "owl_fft_best_fit": [u32, cp, cp, C.c_size_t, P(u32), P(u32)],
    "owl_fft_best_fit_ref": [u32, cp, cp, C.c_size_t, P(u32), P(u32)],
    "owl_fft_tune_load": [cp],
    "owl_fft_tune_save": [cp],
    "owl_fft_tune_clear": [],
    "owl_fft_tune_add": [dbl, cp, P(i32)],
    "owl_fft_tune_entry": [u32, P(dbl), cp, C.c_size_t, P(u32), P(i32)],
    "owl_gpu_create": [u32, cp, i32, u32, P(vp)],
    "owl_gpu_destroy": [vp],
    "owl_gpu_info": [vp, P(GpuInfo)],
    "owl_gpu_write_words": [vp, i32, P(u32), u32],
    "owl_gpu_read_words": [vp, i32, P(u32), u32, P(i32)],
    "owl_gpu_write_digits": [vp, i32, P(i32), u32],
    "owl_gpu_read_digits": [vp, i32, P(i32), u32],
    "owl_gpu_read_digits_noalloc": [vp, i32, P(i32), u32],
    "owl_gpu_res64": [vp, i32, P(u64)],
    "owl_gpu_square_loop": [vp, i32, i32, u32, u32, i32, i32],
    "owl_gpu_mod_mul": [vp, i32, i32, i32],
    "owl_gpu_do_check": [vp, u32, P(i32)],
    "owl_gpu_is_equal": [vp, i32, i32, P(i32)],
    "owl_gpu_copy": [vp, i32, i32],
    "owl_gpu_read_roe": [vp, P(u32), P(dbl), P(dbl), P(dbl)],
    "owl_gpu_measure_roe": [vp, u32, u32, P(u32), P(dbl), P(dbl), P(dbl), P(dbl)],
    "owl_gpu_prp": [vp, u32, u32, u32, P(u64), u32, u32, P(PrpResult)],
    "owl_gpu_arm_fault": [vp, u32],
    "owl_gpu_ll": [vp, u32, P(u64), u32, u32, P(LLResult)],
    "owl_gpu_prp_saved": [vp, cp, u32, u32, u32, u32, P(u64), u32, u32, P(PrpResult)],
    "owl_gpu_ll_saved": [vp, cp, u32, u32, u32, P(u64), u32, u32, P(LLResult)],
    "owl_savefile_write_prp": [cp, u32, u32, u32, u64, u32, dbl, P(u32), u32],
    "owl_savefile_read_prp": [cp, P(u32), P(u32), P(u32), P(u64), P(u32), P(dbl), P(u32), u32],
    "owl_savefile_write_ll": [cp, u32, u32, dbl, P(u32), u32],
    "owl_savefile_read_ll": [cp, P(u32), P(u32), P(dbl), P(u32), u32],
    "owl_crc32": [vp, C.c_size_t, P(u32)],
    "owl_trig_root1": [u32, u64, P(dbl), P(dbl)],
    "owl_proof_points": [u32, u32, P(u32), u32, P(u32)],
    "owl_gpu_prp_proof": [vp, cp, u32, u32, u32, cp, P(PrpResult)],
    "owl_gpu_proof_verify": [vp, cp, P(i32), P(i32)],
    "owl_gpu_exp_mul": [vp, P(u32), u64, P(u32), i32, P(u32), u32],
    "owl_sha3_256": [vp, C.c_size_t, cp],
    "owl_worktodo_parse": [cp, P(i32), P(u32), cp, C.c_size_t, P(u32)],
    "owl_worktodo_best": [P(cp), u32, P(i32)],
    "owl_prp_type1": [P(u32), u32, u32, P(u64), cp, C.c_size_t],
    "owl_result_json": [i32, u32, i32, u64, cp, u32, u32, cp, cp, cp, cp, cp, cp, C.c_size_t],
    "owl_result_json_ex": [P(ResultRec), cp, C.c_size_t],
    "owl_proof_md5": [cp, cp, C.c_size_t],
    "owl_gpu_cert": [vp, cp, u32, i32, cp, C.c_size_t, P(u64)],
    "owl_gpu_time_prp": [vp, u32, u32, u32, P(dbl)],
    "owl_gpu_time_squares": [vp, u32, u32, i32, P(dbl)],
    "owl_gpu_time_kernels": [vp, u32, P(dbl), P(dbl)],
    "owl_gpu_freq_io": [vp, i32, P(dbl), u32, i32],
    "owl_gpu_time_copy": [vp, u32, P(dbl)],
    "owl_gpu_pass": [vp, i32, i32],
    "owl_batch_create": [P(u32), u32, cp, i32, P(vp)],
    "owl_batch_destroy": [vp],
    "owl_batch_info": [vp, P(u32), P(u32), P(u64)],
    "owl_batch_job_info": [vp, u32, P(GpuInfo)],
    "owl_batch_prp": [vp, u32, u32, u32, P(u64), u32, u32, P(PrpResult)],
    "owl_batch_arm_fault": [vp, u32, u32],
    "owl_batch_ll": [vp, u32, P(u64), u32, u32, P(LLResult)],
    "owl_batch_time_squares": [vp, u32, u32, i32, P(dbl)],
    "owl_tune_run": [u32, cp, u32, P(TuneEntry), u32, P(u32)],
    "owl_tune_set": [u32, u32, u32, i32, dbl],
    "owl_tune_batch": [P(u32), u32, cp, u32, P(TuneEntry), u32, P(u32)],
    "owl_tune_batch_clear": [],
    "owl_tune_get": [u32, P(u32), P(u32), P(i32), P(dbl), P(i32)],
    "owl_tune_clear": [],
    "owl_tune_load": [cp],
    "owl_tune_save": [cp],
    "owl_server_create": [u64, u32, u32, u32, cp, i32, P(vp)],
    "owl_server_destroy": [vp],
    "owl_server_submit": [vp, u32, i32, u32, P(u32)],
    "owl_server_arm_fault": [vp, u32, u32],
    "owl_server_step": [vp, P(u32)],
    "owl_server_take": [vp, P(ServeResult), u32, P(u32)],
    "owl_server_info": [vp, P(ServeInfo)],
    "owl_server_now": [vp, P(dbl)],
}

_lib = None


def load():
    """Load the library (once). Raises OwlError if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise OwlError(f"{LIB_PATH} is missing: build it with `python gpuowl_amd/build.py` "
                       "(the HIP path has no fallback)")
    lib = C.CDLL(LIB_PATH)
    for name, args in SIGNATURES.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = C.c_int
    lib.owl_last_error.argtypes = []
    lib.owl_last_error.restype = C.c_char_p
    _lib = lib
    return lib


def check(rc):
    if rc != 0:
        msg = _lib.owl_last_error().decode() if _lib is not None else "library not loaded"
        raise OwlError(msg or f"error {rc}")
