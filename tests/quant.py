"""numpy restatements of the operand formats, for exact expected values.

* BF16 split: hi = RN_bf16(x32), lo = RN_bf16(x32 - hi), x32 = RN_f32(x).
* OCP FP8 E4M3: round to nearest even, saturating at 448, subnormal step 2^-9;
  one power-of-two scale 2^e per (occupied i, 64-row Q block), e the smallest
  integer with 448 * 2^e >= max|x32| over the block.
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
