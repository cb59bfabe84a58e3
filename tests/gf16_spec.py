"""Independent numpy restatement of the MAYO maths, used to pin the oracle.

Written in the per-equation matrix form of the MAYO specification (one v x v
matrix P1_a per equation a, explicit matrix products over GF(16), E^l by
repeated multiplication by z) rather than the byte-lane loops of the
reference and the oracle, so an index, transpose or dropped-term mistake in
either shows up as a mismatch.
"""
from __future__ import annotations

import numpy as np


def gf16_mul_closed(a: int, b: int) -> int:
    """Carry-less product reduced modulo x^4 + x + 1."""
    p = 0
    for i in range(4):
        if (b >> i) & 1:
            p ^= a << i
    for deg in (6, 5, 4):
        if (p >> deg) & 1:
            p ^= 0b10011 << (deg - 4)
    return p & 0xF


MUL = np.array([[gf16_mul_closed(a, b) for b in range(16)] for a in range(16)], np.uint8)
INV = np.zeros(16, np.uint8)
for _a in range(1, 16):
    INV[_a] = [b for b in range(16) if MUL[_a, b] == 1][0]


def unpack_nibbles(buf: bytes | np.ndarray, count: int) -> np.ndarray:
    b = np.frombuffer(bytes(buf), np.uint8) if not isinstance(buf, np.ndarray) else buf
    out = np.empty(2 * b.size, np.uint8)
    out[0::2] = b & 0xF
    out[1::2] = b >> 4
    return out[:count]


def pack_nibbles(vals: np.ndarray) -> bytes:
    v = np.asarray(vals, np.uint8)
    if v.size % 2:
        v = np.concatenate([v, [0]])
    return (v[0::2] | (v[1::2] << 4)).astype(np.uint8).tobytes()


def matmul(A: np.ndarray, B: np.ndarray) -> np.ndarray:
    """GF(16) matrix product (XOR-accumulated table products)."""
    A = np.asarray(A, np.uint8)
    B = np.asarray(B, np.uint8)
    prod = MUL[A[:, :, None], B[None, :, :]]  # i x k x j
    return np.bitwise_xor.reduce(prod, axis=1)


def matvec(A: np.ndarray, x: np.ndarray) -> np.ndarray:
    return matmul(A, np.asarray(x, np.uint8)[:, None])[:, 0]


def upper(X: np.ndarray) -> np.ndarray:
    """Upper(X): off-diagonal entries folded into the upper triangle."""
    U = np.triu(X ^ X.T)
    np.fill_diagonal(U, np.diag(X))
    return U


class Layout:
    """Byte layout of the expanded keys as matrices per equation."""

    def __init__(self, p):
        self.p = p

    def mvecs(self, buf: bytes, count: int) -> np.ndarray:
        """count m-vectors -> (count, m) nibbles."""
        mb = self.p.m_bytes
        raw = np.frombuffer(bytes(buf), np.uint8)[: count * mb].reshape(count, mb)
        out = np.empty((count, self.p.m), np.uint8)
        out[:, 0::2] = raw & 0xF
        out[:, 1::2] = raw >> 4
        return out

    def upper_matrices(self, buf: bytes, dim: int) -> np.ndarray:
        """Upper-triangular dim x dim matrix of m-vectors -> (m, dim, dim)."""
        cnt = dim * (dim + 1) // 2
        vecs = self.mvecs(buf, cnt)
        out = np.zeros((self.p.m, dim, dim), np.uint8)
        e = 0
        for r in range(dim):
            for c in range(r, dim):
                out[:, r, c] = vecs[e]
                e += 1
        return out

    def rect_matrices(self, buf: bytes, rows: int, cols: int) -> np.ndarray:
        vecs = self.mvecs(buf, rows * cols).reshape(rows, cols, self.p.m)
        return np.transpose(vecs, (2, 0, 1)).copy()

    def O(self, buf: bytes) -> np.ndarray:
        p = self.p
        return unpack_nibbles(buf, p.v * p.o).reshape(p.v, p.o)


def E_times(p, y: np.ndarray) -> np.ndarray:
    """z * y(z) mod f(z), f = z^m + sum_{i<4} tail_i z^i, by long division."""
    prod = np.concatenate([[0], y]).astype(np.uint8)  # degree <= m
    lead = prod[p.m]
    prod = prod[: p.m].copy()
    for i, c in enumerate(p.f_tail):
        prod[i] ^= MUL[lead, c]
    return prod


def E_pow(p, y: np.ndarray, ell: int) -> np.ndarray:
    for _ in range(ell):
        y = E_times(p, y)
    return y


def pair_schedule(k: int):
    """(ell, i, j) in the order of the MAYO specification's double loop."""
    ell = 0
    for i in range(k):
        for j in range(k - 1, i - 1, -1):
            yield ell, i, j
            ell += 1


def quad_vectors(P: np.ndarray, S: list) -> dict:
    """u_ij[a] = S_i^T P_a S_j + S_j^T P_a S_i (i<j), S_i^T P_a S_i (i==j)."""
    m = P.shape[0]
    out = {}
    k = len(S)
    for i in range(k):
        for j in range(i, k):
            u = np.zeros(m, np.uint8)
            for a in range(m):
                t = matvec(P[a], S[j])
                u[a] = np.bitwise_xor.reduce(MUL[S[i], t])
                if i != j:
                    t2 = matvec(P[a], S[i])
                    u[a] ^= np.bitwise_xor.reduce(MUL[S[j], t2])
            out[(i, j)] = u
    return out


def public_map(p, P1m, P2m, P3m, s: np.ndarray) -> np.ndarray:
    """y = sum_l E^l u_l over the full n x n forms (s: k x n nibbles)."""
    v, o, n = p.v, p.o, p.n
    P = np.zeros((p.m, n, n), np.uint8)
    P[:, :v, :v] = P1m
    P[:, :v, v:] = P2m
    P[:, v:, v:] = P3m
    U = quad_vectors(P, [s[i] for i in range(p.k)])
    y = np.zeros(p.m, np.uint8)
    for ell, i, j in pair_schedule(p.k):
        y ^= E_pow(p, U[(i, j)], ell)
    return y
