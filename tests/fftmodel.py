"""numpy FP64 model of each GPU pass, in the engine's HBM layout (test infrastructure).

Layout: element (line l, slot s) at (l//R)*W*R + s*R + l%R. Complex index j = l + H*c,
word k = 2j + rep. Frequency slots: 2q -> column frequency q, 2q+1 -> W-q (q >= 1);
slot 0 -> 0, slot 1 -> W/2. Forward transform: row FFT over c, twiddle exp(-2 pi i l a/ND),
column FFT over l; the pairing couples K = a + W b with ND - K."""
import numpy as np


class Model:
    def __init__(self, E, W, H, R):
        self.E, self.W, self.H, self.R = E, W, H, R
        self.ND = W * H
        self.N = 2 * self.ND
        N = self.N
        step = N - E % N
        k = np.arange(N, dtype=np.int64)
        self.extra = (step * k) % N
        self.bits = E // N + ((self.extra + step) < N).astype(np.int64)
        self.w = np.exp2(self.extra.astype(np.longdouble) / N).astype(np.float64)
        self.iw = np.exp2(-self.extra.astype(np.longdouble) / N).astype(np.float64)
        s = np.zeros(W, dtype=np.int64)
        q = np.arange(1, W // 2)
        s[0], s[W // 2] = 0, 1
        s[q], s[W - q] = 2 * q, 2 * q + 1
        self.slot = s                                  # frequency a -> slot
        l = np.arange(H)[:, None]
        self.lay = (l // R) * W * R + np.arange(W)[None, :] * R + (l % R)  # [l, col] -> index
        self.tw = np.exp(-2j * np.pi * ((np.arange(H)[:, None] * np.arange(W)[None, :]) % self.ND) / self.ND)

    # ---- layout helpers
    def to_layout(self, A):      # A[l, a] (frequency columns) -> HBM array
        out = np.empty(self.ND, dtype=complex)
        out[self.lay[:, self.slot]] = A
        return out

    def from_layout(self, buf):  # HBM array -> A[l, a]
        return buf[self.lay[:, self.slot]]

    def digits_to_rows(self, d):  # sequential digits -> weighted complex Z[l, c]
        x = d * self.w
        z = x[0::2] + 1j * x[1::2]          # j = l + H c
        return z.reshape(self.W, self.H).T

    # ---- passes
    def lead_in(self, d):
        return self.to_layout(np.fft.fft(self.digits_to_rows(d), axis=1))

    def col_square(self, buf, mul_spec=None):
        A = self.from_layout(buf)
        Z = np.fft.fft(A * self.tw, axis=0)                 # Z[b, a] = Z_{a + W b}
        ZK = Z.reshape(-1)                                   # index b*W + a == K
        K = np.arange(self.ND)
        P = (self.ND - K) % self.ND
        wK = np.exp(-2j * np.pi * K / self.ND)
        if mul_spec is None:
            Zp = ZK * ZK - (1 + wK) * (ZK - np.conj(ZK[P])) ** 2 / 4
        else:
            Y = mul_spec
            Zp = ZK * Y - (1 + wK) * (ZK - np.conj(ZK[P])) * (Y - np.conj(Y[P])) / 4
        B = np.fft.ifft(Zp.reshape(self.H, self.W), axis=0) * self.H
        return self.to_layout(B * np.conj(self.tw))

    def spectrum(self, buf):
        """COL_FWD: the column transform, flattened by K = a + W b (for mul_spec)."""
        return np.fft.fft(self.from_layout(buf) * self.tw, axis=0).reshape(-1)

    def spectrum_layout(self, buf):
        Z = np.fft.fft(self.from_layout(buf) * self.tw, axis=0)  # [b, a]
        return self.to_layout(Z)

    def row_inverse(self, buf):
        """-> real words v_k = (weighted product)/w_k, unrounded, sequential order."""
        A = self.from_layout(buf)
        z = np.fft.ifft(A, axis=1) * self.W / self.ND       # [l, c]
        zj = z.T.reshape(-1)
        v = np.empty(self.N)
        v[0::2], v[1::2] = zj.real, zj.imag
        return v * self.iw

    def _lowbits(self, u, b):
        m = np.int64(1) << b
        lo = u & (m - 1)
        return np.where(lo >= (m >> 1), lo - m, lo)

    def carry_blocks(self, v, mul3=False, ll=False):
        """Round, then carry along each block of R lines (per column). Returns digits
        (sequential) and the carry out of each block [lb, c]."""
        u = np.rint(v).astype(np.int64)
        roe = float(np.max(np.abs(v - np.rint(v))))
        if mul3:
            u *= 3
        if ll:
            u[0] -= 2
        W, H, R = self.W, self.H, self.R
        nb = H // R
        # word k = 2(l + H c) + rep; arrange as [c, lb, r, rep]
        U = u.reshape(W, nb, R, 2)
        Bts = self.bits.reshape(W, nb, R, 2)
        D = np.empty_like(U)
        s = np.zeros((W, nb), dtype=np.int64)
        for r in range(R):
            for rep in range(2):
                t = U[:, :, r, rep] + s
                b = Bts[:, :, r, rep]
                d = self._lowbits(t, b)
                s = (t - d) >> b
                D[:, :, r, rep] = d
        return D, s, roe

    def _pred_carry(self, s):
        """carry arriving at block (lb, c): from (lb-1, c), or for lb = 0 from (nb-1, c-1) (wrapping)."""
        cin = np.empty_like(s)
        cin[:, 1:] = s[:, :-1]
        cin[:, 0] = np.roll(s[:, -1], 1)
        return cin

    def fused_digits(self, v, ll=False):
        """The fused row pass's digits: block carry chain, then the predecessor's block
        carry rippled through this block's 2R words; what is left after the block's last
        word stays in that word."""
        D, s, roe = self.carry_blocks(v, ll=ll)
        c = self._pred_carry(s)
        Bts = self.bits.reshape(D.shape)
        R = self.R
        for r in range(R):
            for rep in range(2):
                last = (r == R - 1 and rep == 1)
                t = D[:, :, r, rep] + c
                if last:
                    D[:, :, r, rep] = t
                else:
                    b = Bts[:, :, r, rep]
                    d = self._lowbits(t, b)
                    c = (t - d) >> b
                    D[:, :, r, rep] = d
        return D.reshape(-1), roe

    def full_digits(self, v, mul3=False, ll=False):
        """Lead-out: fully normalized balanced digits."""
        D, s, roe = self.carry_blocks(v, mul3=mul3, ll=ll)
        d = D.reshape(-1).copy()
        carry = 0
        # sequential ripple of the block carries in word order (k = 0 .. N-1, wrapping)
        flat_s = s  # [c, lb] -> carry into the block after it in word order
        bpb = 2 * self.R
        order_s = flat_s.reshape(-1)  # blocks in word order: c * nb + lb
        add = np.zeros(len(order_s), dtype=np.int64)
        add[1:] = order_s[:-1]
        add[0] = order_s[-1]
        pending = add
        for _ in range(64):
            if not pending.any():
                break
            nxt = np.zeros_like(pending)
            idx = np.nonzero(pending)[0]
            for bi in idx:
                c = pending[bi]
                k = bi * bpb
                n = 0
                while c and n < self.N:
                    kk = (k + n) % self.N
                    t = int(d[kk]) + int(c)
                    b = int(self.bits[kk])
                    lo = t & ((1 << b) - 1)
                    if lo >= (1 << (b - 1)):
                        lo -= 1 << b
                    c = (t - lo) >> b
                    d[kk] = lo
                    n += 1
            pending = nxt
        return d, roe

    def random_digits(self, rng):
        b = self.bits
        return rng.integers(-(1 << (b - 1)), 1 << (b - 1), dtype=np.int64)
