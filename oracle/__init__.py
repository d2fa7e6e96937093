"""ORACLE -- test infrastructure only. Imported by tests/, __graft_entry__.smoke()
and bench.py's cpu_baseline leg as the checker; the product path never uses it.

ctypes wrapper of oracle/_build/liboracle.so (gmp_loop.c: exact GMP PRP/LL loop;
ibdwt.c: plain-C restatement of the reference's IBDWT word layout, packing,
readout and carry) and of oracle/_ref/refcheck (the reference's own
src/state.cpp compiled in this tree, when present)."""
import ctypes as C
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "liboracle.so")
GMP_CLI = os.path.join(HERE, "_build", "gmp_loop")
REFCHECK = os.path.join(HERE, "_ref", "refcheck")
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} missing: run `make -C oracle`")
        L = C.CDLL(LIB)
        u32, i32, u64 = C.c_uint32, C.c_int32, C.c_uint64
        P = C.POINTER
        L.oracle_run.argtypes = [u32, C.c_int, u32, u32, P(u64), u32, P(u32)]
        L.oracle_run.restype = C.c_int
        L.oracle_mulmod.argtypes = [u32, P(u32), P(u32), C.c_int, P(u32)]
        L.oracle_square_n.argtypes = [u32, P(u32), u32, P(u32)]
        L.ibdwt_extra.argtypes = [u32, u32, u32]
        L.ibdwt_extra.restype = u32
        L.ibdwt_is_big.argtypes = [u32, u32, u32]
        L.ibdwt_bitlen.argtypes = [u32, u32, u32]
        L.ibdwt_bitlen.restype = u32
        L.ibdwt_compact.argtypes = [P(i32), u32, u32, P(u32)]
        L.ibdwt_expand.argtypes = [P(u32), u32, u32, P(i32)]
        L.ibdwt_residue_from_ends.argtypes = [P(i32), u32, u32]
        L.ibdwt_residue_from_ends.restype = u64
        L.ibdwt_square.argtypes = [P(i32), u32, u32, C.c_int, C.c_int]
        L.ibdwt_square.restype = C.c_double
        L.ibdwt_run.argtypes = [u32, u32, C.c_int, u32, u32, P(u64), u32]
        L.ibdwt_run.restype = C.c_double
        _lib = L
    return _lib


def n_words(E):
    return (E - 1) // 32 + 1


def _words(ws):
    return (C.c_uint32 * len(ws))(*ws)


def gmp_res64(E, iters, step, ll=False):
    cap = iters // step
    out = (C.c_uint64 * max(cap, 1))()
    n = lib().oracle_run(E, int(ll), iters, step, out, cap, None)
    return list(out)[:n]


def gmp_residue(E, iters, ll=False):
    """Full residue words after iters iterations from 3 (PRP) or 4 (LL)."""
    w = (C.c_uint32 * n_words(E))()
    dummy = (C.c_uint64 * 1)()
    lib().oracle_run(E, int(ll), iters, 0, dummy, 0, w)
    return list(w)


def gmp_mulmod(E, a, b, mul3=False):
    out = (C.c_uint32 * n_words(E))()
    lib().oracle_mulmod(E, _words(a), _words(b), int(mul3), out)
    return list(out)


def gmp_square_n(E, a, n):
    out = (C.c_uint32 * n_words(E))()
    lib().oracle_square_n(E, _words(a), n, out)
    return list(out)


def bitlen(N, E, k):
    return lib().ibdwt_bitlen(N, E, k)


def extra(N, E, k):
    return lib().ibdwt_extra(N, E, k)


def compact(digits, N, E):
    out = (C.c_uint32 * n_words(E))()
    rc = lib().ibdwt_compact((C.c_int32 * N)(*digits), N, E, out)
    if rc:
        raise ValueError("digit out of range")
    return list(out)


def expand(words, N, E):
    d = (C.c_int32 * N)()
    lib().ibdwt_expand(_words(words), N, E, d)
    return list(d)


def residue_from_ends(ends64, N, E):
    return lib().ibdwt_residue_from_ends((C.c_int32 * 64)(*ends64), N, E)


def ibdwt_square(digits, N, E, ll=False, mul3=False):
    d = (C.c_int32 * N)(*digits)
    roe = lib().ibdwt_square(d, N, E, int(ll), int(mul3))
    return list(d), roe


def ibdwt_run(E, N, iters, step, ll=False):
    cap = iters // step
    out = (C.c_uint64 * max(cap, 1))()
    roe = lib().ibdwt_run(E, N, int(ll), iters, step, out, cap)
    return list(out)[:cap], roe


def ref_available():
    return os.path.exists(REFCHECK)


def ref_expand(words, N, E):
    """expandBits of the reference's src/state.cpp (oracle/_ref/refcheck)."""
    inp = "\n".join(f"{w:x}" for w in words) + "\n"
    r = subprocess.run([REFCHECK, "expand", str(E), str(N)], input=inp, capture_output=True, text=True, check=True)
    return [int(x) for x in r.stdout.split()]


def ref_compact(digits, N, E):
    """compactBits of the reference's src/state.cpp (oracle/_ref/refcheck)."""
    inp = "\n".join(str(d) for d in digits) + "\n"
    r = subprocess.run([REFCHECK, "compact", str(E), str(N)], input=inp, capture_output=True, text=True, check=True)
    return [int(x, 16) for x in r.stdout.split()]


def gmp_cli():
    return GMP_CLI
