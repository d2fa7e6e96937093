"""MI355X-native engine for PRPLL's hot path: one PRP (x -> x^2 mod 2^E - 1) or
LL (x -> x^2 - 2) iteration as an FP64 IBDWT, in hand-written HIP for gfx950.

The host engine is C++ behind the C-ABI of include/owl.h; this package only
loads it (ctypes) and mirrors the reference's Gpu / FFTConfig interface."""
from ._lib import LIB_PATH, OwlError, load
from .engine import (AUX, BASE, CHECK, DATA, Batch, Gpu, Server, best_fit, best_fit_ref, crc32, fft_tune_add, fft_tune_clear, fft_tune_list, fft_tune_load,
                     fft_tune_save, kernel_names, int_to_words, n_words,
                     proof_md5, proof_points, prp_type1, result_json, root1, sha3_256, savefile_read_ll, savefile_read_prp, savefile_write_ll,
                     savefile_write_prp, worktodo_best, worktodo_parse, tune, tune_batch, tune_batch_clear, tune_clear,
                     tune_get, tune_load, tune_save, tune_set, words_to_int)

__all__ = ["Gpu", "Batch", "Server", "kernel_names", "tune", "tune_batch", "tune_batch_clear", "tune_get", "tune_set", "tune_clear", "tune_load", "tune_save", "best_fit", "best_fit_ref", "OwlError", "load", "LIB_PATH", "DATA", "AUX", "CHECK", "BASE", "n_words",
           "words_to_int", "int_to_words", "crc32", "savefile_read_prp", "savefile_write_prp", "savefile_read_ll",
           "savefile_write_ll", "fft_tune_load", "fft_tune_save", "fft_tune_clear", "fft_tune_add", "fft_tune_list", "worktodo_parse",
           "worktodo_best", "prp_type1", "result_json", "proof_points", "sha3_256"]
