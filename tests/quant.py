"""numpy restatements of the operand formats, for exact expected values.

* BF16 split: hi = RN_bf16(x32), lo = RN_bf16(x32 - hi), x32 = RN_f32(x).
* OCP FP8 E4M3: round to nearest even, saturating at 448, subnormal step 2^-9;
  one power-of-two scale 2^e per (occupied i, 64-row Q block), e the smallest
  integer with 448 * 2^e >= max|x32| over the block.
* INT8 digit slices: per column (i, a) of b one exponent f: E = frexp exponent
  of m = max_Q |x| (m < 2^E), f = E - 1 if m <= (1 - 2^-7) 2^E else E (0 for a
  zero column); Y = RN(x 2^(30 - f)) and its signed byte digits
  Y = d1 2^24 + d2 2^16 + d3 2^8 + d4 peeled from the bottom;
  the contraction keeps the digit products of levels p + q = 2..6, summed
  exactly, and combines them in FP64 as the kernel does.
The GPU quantisers (ferric_amd/csrc/pack.hip) must agree bit for bit, which the
GPU tests check through the energies these values produce.
"""
from __future__ import annotations

import numpy as np


def bf16_bits(x) -> np.ndarray:
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def bf16_value(bits) -> np.ndarray:
    return (np.asarray(bits).astype(np.uint32) << 16).view(np.float32)


def bf16_split(b) -> tuple[np.ndarray, np.ndarray]:
    """(hi, lo) as float64 arrays of the BF16 values."""
    x32 = np.asarray(b).astype(np.float32)
    hi = bf16_value(bf16_bits(x32))
    lo = bf16_value(bf16_bits(x32 - hi))
    return hi.astype(np.float64), lo.astype(np.float64)


def e4m3_encode(x) -> np.ndarray:
    x = np.asarray(x, dtype=np.float32)
    sign = np.signbit(x).astype(np.uint8) << 7
    a = np.abs(x).astype(np.float64)
    sub = a < 2.0 ** -6
    k = np.rint(a * 512.0)
    _, e2 = np.frexp(np.where(sub, 1.0, a))
    ex = e2 - 1
    q = np.rint(np.ldexp(np.where(sub, 1.0, a), 3 - ex))
    carry = q >= 16
    q = np.where(carry, 8.0, q)
    ex = np.where(carry, ex + 1, ex)
    code = np.minimum(((ex + 7) << 3) | (q.astype(np.int64) - 8), 0x7E)
    code = np.where(a >= 448.0, 0x7E, np.where(sub, k.astype(np.int64), code))
    return (code.astype(np.uint8) | sign).astype(np.uint8)


def e4m3_decode(codes) -> np.ndarray:
    c = np.asarray(codes).astype(np.int64)
    s = np.where(c & 0x80, -1.0, 1.0)
    e = (c >> 3) & 0xF
    m = c & 7
    return s * np.where(e == 0, m * 2.0 ** -9, (8 + m) * np.exp2(e - 10.0))


def fp8_scale_exponents(amax) -> np.ndarray:
    amax = np.asarray(amax, dtype=np.float64)
    pos = amax > 0
    safe = np.where(pos, amax, 448.0)
    e = np.ceil(np.log2(safe / 448.0)).astype(np.int64)
    e = np.where(448.0 * np.exp2(e - 1.0) >= safe, e - 1, e)
    e = np.where(448.0 * np.exp2(e * 1.0) < safe, e + 1, e)
    return np.clip(np.where(pos, e, 0), -127, 127)


def fp8_dequantized(b, nocc, nvir) -> np.ndarray:
    """The FP64 values the FP8 path contracts: decode(q) * 2^e, shape of b."""
    x32 = np.asarray(b).astype(np.float32)
    naux = x32.shape[0]
    out = np.empty(x32.shape, dtype=np.float64)
    for i in range(nocc):
        cols = slice(i * nvir, (i + 1) * nvir)
        for q0 in range(0, naux, 64):
            blk = x32[q0:q0 + 64, cols]
            e = int(fp8_scale_exponents(np.abs(blk).max()))
            scaled = np.ldexp(blk.astype(np.float64), -e).astype(np.float32)
            out[q0:q0 + 64, cols] = np.ldexp(e4m3_decode(e4m3_encode(scaled)), e)
    return out


def int8x4_digits(b):
    """(e, d): column exponents (nov,) and digits (4, naux, nov) int64 of b."""
    x = np.asarray(b, dtype=np.float64)
    m = np.abs(x).max(axis=0) if x.shape[0] else np.zeros(x.shape[1])
    _, e = np.frexp(m)
    e = np.where(m <= np.ldexp(1.0 - 2.0 ** -7, e), e - 1, e)
    e = np.where(m > 0, e, 0).astype(np.int64)
    y = np.rint(np.ldexp(x, (30 - e)[None, :])).astype(np.int64)
    d = np.empty((4,) + x.shape, dtype=np.int64)
    for k in (3, 2, 1):
        d[k] = ((y & 0xFF) ^ 0x80) - 0x80
        y = (y - d[k]) >> 8
    d[0] = y
    return e, d


def int8x4_g(b):
    """The (nov, nov) matrix g = (ia|jb) the INT8X4 kernel forms, bit for bit:
    level sums S_L (exact int64), t = S6 2^-32 + S5 2^-24 + S4 2^-16 + S3 2^-8 + S2
    summed in that order in FP64, then g = t 2^(e_ia + e_jb - 12)."""
    return int8x4_g_rect(b, b)


def int8x4_g_rect(ba, bb):
    """int8x4_g for two factors (the opposite-spin rectangle): (nov_a, nov_b)
    g[ia, JB] from the digits of each factor under its own row exponents."""
    ea, da = int8x4_digits(ba)
    eb, db = int8x4_digits(bb)
    s = [np.zeros((da.shape[2], db.shape[2]), dtype=np.int64) for _ in range(5)]
    # Each digit product sum is an integer of magnitude <= 128^2 naux < 2^53,
    # and so is every partial sum of it: an FP64 BLAS product forms it exactly
    # in any association (and much faster than an int64 matmul).
    fa = da.astype(np.float64)
    fb = db.astype(np.float64)
    assert 128 ** 2 * max(da.shape[1], 1) < 2 ** 53
    for p in range(4):
        for q in range(4):
            if p + q < 5:
                s[p + q] += (fa[p].T @ fb[q]).astype(np.int64)
    t = s[4].astype(np.float64) * 2.0 ** -32
    t = t + s[3].astype(np.float64) * 2.0 ** -24
    t = t + s[2].astype(np.float64) * 2.0 ** -16
    t = t + s[1].astype(np.float64) * 2.0 ** -8
    t = t + s[0].astype(np.float64)
    return np.ldexp(t, ea[:, None] + eb[None, :] - 12)


def int8x4_g_bound_rect(ba, bb):
    """int8x4_g_bound for two factors: per-element bound on
    |int8x4_g_rect(ba, bb) - ba^T bb|."""
    x = np.asarray(ba, dtype=np.float64)
    y = np.asarray(bb, dtype=np.float64)
    naux = x.shape[0]
    ex, _ = int8x4_digits(x)
    ey, _ = int8x4_digits(y)
    ux, uy = np.ldexp(1.0, ex - 31), np.ldexp(1.0, ey - 31)
    sx, sy = np.abs(x).sum(axis=0), np.abs(y).sum(axis=0)
    unit = np.ldexp(1.0, ex[:, None] + ey[None, :] - 60)
    bnd = np.outer(ux, sy) + np.outer(sx, uy) + naux * np.outer(ux, uy)
    bnd += naux * (2.0 * 128 ** 2 * 2 ** 8 + 128 ** 2) * unit
    bnd += 6 * 2.0 ** -53 * (np.abs(x).T @ np.abs(y) + bnd)
    return bnd


def int8x4_g_bound(b):
    """Per-element bound (nov, nov) on |int8x4_g(b) - b^T b|: with x^ = Y 2^(f - 30),
    |x - x^| <= u = 2^(f - 31), so sum_Q |x^ y^ - x y| <= u_x S_y + u_y S_x + naux u_x u_y
    (S = sum_Q |x|); the dropped levels 7 and 8 add at most
    naux (2 * 128^2 * 2^8 + 128^2) 2^(f_x + f_y - 60); the FP64 combination rounds
    at most 6 times at 2^-53 of a partial no larger than sum_Q |x^ y^|."""
    x = np.asarray(b, dtype=np.float64)
    naux = x.shape[0]
    e, _ = int8x4_digits(x)
    u = np.ldexp(1.0, e - 31)
    s = np.abs(x).sum(axis=0)
    unit = np.ldexp(1.0, e[:, None] + e[None, :] - 60)
    bnd = np.outer(u, s) + np.outer(s, u) + naux * np.outer(u, u)
    bnd += naux * (2.0 * 128 ** 2 * 2 ** 8 + 128 ** 2) * unit
    bnd += 6 * 2.0 ** -53 * (np.abs(x).T @ np.abs(x) + bnd)
    return bnd


def int8x4_energy_bound(b, eps, nocc, nvir, first_occ=0, nocc_total=None):
    """(os, ss) bounds on |E(int8x4_g(b)) - E(b^T b)|, propagating int8x4_g_bound
    through the closed-shell pair sums (rimp2.rs:984-1018): per unique pair
    i <= j with f = 1 or 2, |d(g^2/D)| <= (2|g| d + d^2)/|D| and
    |d(g_ab (g_ab - g_ba)/D)| <= (|g_ab| (d_ab + d_ba) + |g_ab - g_ba| d_ab
    + d_ab (d_ab + d_ba))/|D|."""
    x = np.asarray(b, dtype=np.float64)
    eps = np.asarray(eps, dtype=np.float64)
    if nocc_total is None:
        nocc_total = first_occ + nocc
    g = x.T @ x
    dl = int8x4_g_bound(x)
    eo = eps[first_occ:first_occ + nocc]
    ev = eps[nocc_total:nocc_total + nvir]
    bos = bss = 0.0
    for i in range(nocc):
        for j in range(i, nocc):
            f = 1.0 if i == j else 2.0
            G = g[i * nvir:(i + 1) * nvir, j * nvir:(j + 1) * nvir]
            d = dl[i * nvir:(i + 1) * nvir, j * nvir:(j + 1) * nvir]
            D = np.abs(eo[i] + eo[j] - ev[:, None] - ev[None, :])
            bos += f * np.sum((2 * np.abs(G) * d + d * d) / D)
            bss += f * np.sum((np.abs(G) * (d + d.T) + np.abs(G - G.T) * d + d * (d + d.T)) / D)
    return bos, bss
