"""The reference's own unit-test input generators, restated bit for bit.

* ``lcg_case``       crates/ferric-mp2/src/rimp2.rs:2891-2914 (seed 0x243f6a8885a308d3,
                     values 0.05*(U-0.5) filled in row-major order, eps ladder
                     -0.9-0.07*i / 0.15+0.011*i)
* ``resident_case``  crates/ferric-mp2/tests/gpu_rimp2_resident.rs:92-120 (values
                     0.1*(U-0.5), occupied eps -1.5+1.2U, virtual 0.2+1.8U)
* ``xorshift_case``  crates/ferric-mp2/src/zvector.rs:1135-1146 (seed 0x2545f4914f6cdd1d,
                     values U-0.5; t2 first, then B)

All three map a 64-bit state to (state >> 11) / 2^53, exactly as the Rust code.
"""
from __future__ import annotations

import numpy as np

_M64 = (1 << 64) - 1


def _lcg_stream(seed: int):
    s = seed & _M64
    while True:
        s = (s * 6364136223846793005 + 1442695040888963407) & _M64
        yield (s >> 11) / float(1 << 53)


def lcg_case(naux: int, nocc: int, nvir: int, first_occ: int):
    """(b_ov, eps, nocc, nvir, first_occ, nocc_total) of rimp2.rs:2891-2914."""
    width = nocc * nvir
    nocc_total = first_occ + nocc
    g = _lcg_stream(0x243F6A8885A308D3)
    b = np.array([(next(g) - 0.5) * 0.05 for _ in range(naux * width)]).reshape(naux, width)
    eps = np.array([-0.9 - 0.07 * i if i < nocc_total else 0.15 + 0.011 * i
                    for i in range(nocc_total + nvir)])
    return b, eps, nocc, nvir, first_occ, nocc_total


def resident_case(naux: int, nocc: int, nvir: int, first_occ: int, seed: int):
    """gpu_rimp2_resident.rs:92-120."""
    g = _lcg_stream(seed)
    b = np.array([0.1 * (next(g) - 0.5) for _ in range(naux * nocc * nvir)]).reshape(
        naux, nocc * nvir)
    nocc_total = first_occ + nocc
    eps = [(-1.5 + 1.2 * next(g)) for _ in range(nocc_total)]
    eps += [(0.2 + 1.8 * next(g)) for _ in range(nvir)]
    return b, np.array(eps), nocc, nvir, first_occ, nocc_total


def xorshift_stream(seed: int):
    s = seed & _M64
    while True:
        s ^= (s << 13) & _M64
        s ^= s >> 7
        s ^= (s << 17) & _M64
        yield (s >> 11) / float(1 << 53) - 0.5


def xorshift_case(nocc: int = 4, nvir: int = 7, naux: int = 9):
    """zvector.rs:1135-1146: t2 (nov x nov) then b_ov (naux x nov)."""
    nov = nocc * nvir
    g = xorshift_stream(0x2545F4914F6CDD1D)
    t2 = np.array([next(g) for _ in range(nov * nov)]).reshape(nov, nov)
    b = np.array([next(g) for _ in range(naux * nov)]).reshape(naux, nov)
    return t2, b, nocc, nvir, naux


# Shapes the reference's own tests use.
LCG_CASES = [(60, 5, 12, 0), (120, 8, 40, 2), (40, 3, 1, 0)]
RESIDENT_CASES = [(150, 9, 31, 2, 5), (120, 7, 23, 1, 3), (300, 5, 17, 0, 9),
                  (200, 6, 20, 0, 21), (8, 2, 3, 0, 1)]  # gpu_rimp2_resident.rs:286-448
