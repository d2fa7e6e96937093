"""PRP proofs restated from the reference for tests: the file layout (Proof::save / load,
src/Proof.cpp:61-83, header src/Proof.h:43), the hash chain (proof::hashWords,
src/Proof.cpp:22-30: SHA3-256 over (E-1)/8+1 little-endian bytes, the chain prefixed by the
previous 32-byte digest, h = its first 64 bits little-endian), verification (Proof::verify,
src/Proof.cpp:84-126), the proof points (ProofSet::ProofSet, src/Proof.cpp:130-160) and
construction (ProofSet::computeProof, src/Proof.cpp:262-297).

Written here independently of the engine's C++ (Proof.cpp, Gpu::makeProof / verifyProof), so
that a proof the engine writes is accepted by this reading of the reference's algorithm and
not only by the engine's own verifier. Arithmetic is mod 2^E - 1; `Arith` does it with Python
integers (small E) and `GmpArith` with the GMP oracle (E near 10^6)."""
import array
import hashlib
import re

HEADER = re.compile(rb"PRP PROOF\nVERSION=2\nHASHSIZE=64\nPOWER=(\d+)\nNUMBER=M(\d+)\n")


def nbytes(E):
    return (E - 1) // 8 + 1


def to_bytes(x, E):
    return x.to_bytes(nbytes(E), "little")


def h64(digest):
    return int.from_bytes(digest[:8], "little")


class Arith:
    """x mod 2^E - 1 with Python integers (reduction by folding, so no long division)."""

    def __init__(self, E):
        self.E, self.M = E, (1 << E) - 1

    def red(self, x):
        while x >> self.E:
            x = (x & self.M) + (x >> self.E)
        return 0 if x == self.M else x

    def mul(self, a, b):
        return self.red(a * b)

    def pow(self, a, h):
        r = 1
        for bit in bin(h)[2:] if h else "":
            r = self.mul(r, r)
            if bit == "1":
                r = self.mul(r, a)
        return self.red(r)

    def square_n(self, a, n):
        for _ in range(n):
            a = self.mul(a, a)
        return a


class GmpArith(Arith):
    """The same with the GMP oracle's mulmod and repeated squaring."""

    def __init__(self, E, oracle):
        super().__init__(E)
        self.o = oracle

    def _w(self, x):
        return array.array("I", x.to_bytes(4 * self.o.n_words(self.E), "little")).tolist()

    def _i(self, w):
        return self.red(int.from_bytes(array.array("I", w).tobytes(), "little"))

    def mul(self, a, b):
        return self._i(self.o.gmp_mulmod(self.E, self._w(a), self._w(b)))

    def square_n(self, a, n):
        return self._i(self.o.gmp_square_n(self.E, self._w(a), n)) if n else a

    def pow(self, a, h):
        if not h:
            return 1
        r, run = a, 0
        for bit in bin(h)[3:]:
            run += 1
            if bit == "1":
                r = self.mul(self.square_n(r, run), a)
                run = 0
        return self.square_n(r, run)


def parse(raw):
    m = HEADER.match(raw)
    if not m:
        raise ValueError("invalid proof header")
    power, E = int(m.group(1)), int(m.group(2))
    n, off = nbytes(E), m.end()
    if len(raw) != off + (power + 1) * n:
        raise ValueError("proof size")
    parts = [raw[off + i * n: off + (i + 1) * n] for i in range(power + 1)]
    return E, power, parts[0], parts[1:]


def verify(raw, arith=None):
    """(valid, is_prime) of a proof file's bytes, as Proof::verify decides it."""
    E, power, b_bytes, middles = parse(raw)
    ar = arith or Arith(E)
    digest = hashlib.sha3_256(b_bytes).digest()
    A, B = 3, ar.red(int.from_bytes(b_bytes, "little"))
    span = E
    for mb in middles:
        digest = hashlib.sha3_256(digest + mb).digest()
        h = h64(digest)
        Mi = ar.red(int.from_bytes(mb, "little"))
        if span % 2:
            B = ar.mul(B, B)
        B = ar.mul(ar.pow(Mi, h), B)
        A = ar.mul(ar.pow(A, h), Mi)
        span = (span + 1) // 2
    return ar.square_n(A, span) == B, int.from_bytes(b_bytes, "little") == 9


def points(E, power):
    """The 2^power iterations whose residues a proof uses, ascending (E last)."""
    pts, span = [0], (E + 1) // 2
    for _ in range(power):
        pts += [x + span for x in pts]
        span = (span + 1) // 2
    pts[0] = E
    return sorted(pts)


def make(E, power):
    """The proof file's bytes for 2^E - 1, from x_k = 3^(2^k) computed with Python integers."""
    ar = Arith(E)
    pts = points(E, power)
    want, res, x = set(pts), {}, 3
    for k in range(1, E + 1):
        x = ar.mul(x, x)
        if k in want:
            res[k] = x
    B = res[E]
    digest = hashlib.sha3_256(to_bytes(B, E)).digest()
    middles, hashes = [], []
    for p in range(power):
        s = 1 << (power - p - 1)
        stack = []
        for i in range(1 << p):
            stack.append(res[pts[s * (2 * i + 1) - 1]])
            k = 0
            while i & (1 << k):
                top = stack.pop()
                stack[-1] = ar.mul(ar.pow(stack[-1], hashes[p - 1 - k]), top)
                k += 1
        assert len(stack) == 1
        middles.append(stack[0])
        digest = hashlib.sha3_256(digest + to_bytes(stack[0], E)).digest()
        hashes.append(h64(digest))
    head = b"PRP PROOF\nVERSION=2\nHASHSIZE=64\nPOWER=%d\nNUMBER=M%d\n" % (power, E)
    return head + to_bytes(B, E) + b"".join(to_bytes(m, E) for m in middles)
