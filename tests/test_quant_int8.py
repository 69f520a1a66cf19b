"""Pins of the INT8 digit-slice format restated in tests/quant.py (CPU only).

The GPU tests compare the INT8X4 kernel with the oracle energy of
quant.int8x4_g; these checks fix that restatement from the mathematics:
digits in their ranges and reassembling Y exactly, Y within half a unit of
x 2^(30 - e), the level expansion equal to the exact integer product Y_ia Y_jb
up to the dropped levels 7 and 8 (computed separately and added back), and g
within the format's error bound of the FP64 product.
"""
import numpy as np

from tests import quant


def _b(seed, naux=150, nov=40, boost=True):
    rng = np.random.default_rng(seed)
    b = rng.standard_normal((naux, nov)) * np.sqrt(0.09 / naux)
    if boost:
        b[rng.choice(naux, max(1, naux // 100), replace=False)] *= 20.0
    b[:, 3] = 0.0                         # a zero column
    b[:, 5] *= 1e-30                      # a tiny one
    return b.astype(np.float32).astype(np.float64)


def test_digits_ranges_and_reassembly():
    b = _b(1)
    e, d = quant.int8x4_digits(b)
    assert np.all(np.abs(d[0]) <= 127)
    assert np.all((d[1:] >= -128) & (d[1:] <= 127))
    y = ((d[0] * 256 + d[1]) * 256 + d[2]) * 256 + d[3]
    exact = np.ldexp(b, (30 - e)[None, :])
    assert np.all(np.abs(y - exact) <= 0.5)           # RN to an integer
    assert np.all(np.abs(y) <= 2 ** 31 - 2 ** 24)
    m = np.abs(b).max(axis=0)
    nz = m > 0
    # the row's largest value fills 2^29..2^31 of Y: at least 30 bits used
    assert np.all(m[nz] <= np.ldexp(1.0 - 2.0 ** -7, e[nz] + 1))
    assert np.all(m[nz] > np.ldexp(0.5 - 2.0 ** -8, e[nz] + 1) / 2)
    assert e[3] == 0 and np.all(d[:, :, 3] == 0)


def test_levels_reproduce_the_integer_product():
    """Y_ia Y_jb = sum_{p,q} d_p d'_q 2^(8 (6 - p - q)) (zero-based digits): the
    level weights int8x4_g applies, in exact integers."""
    b = _b(2, naux=90, nov=24)
    _, d = quant.int8x4_digits(b)
    y = ((d[0] * 256 + d[1]) * 256 + d[2]) * 256 + d[3]
    exact = y.T @ y
    total = np.zeros_like(exact)
    for p in range(4):
        for q in range(4):
            total += (d[p].T @ d[q]) * (1 << (8 * (6 - p - q)))
    assert np.array_equal(total, exact)


def test_g_against_fp64_product_within_the_format_bound():
    b = _b(3)
    g = quant.int8x4_g(b)
    ref = b.T @ b
    e, _ = quant.int8x4_digits(b)
    # |dg| <= rounding of Y (half a unit 2^(e - 30) per element, both sides,
    # against |Y| < 2^31) plus the dropped levels (|d_p d_q| <= 2^14, weights
    # 2^8 and 1): bounded by naux 2^(e_a + e_b - 60) (2^31 + 2^31 + 3 * 2^22).
    naux = b.shape[0]
    unit = np.ldexp(1.0, e[:, None] + e[None, :] - 60)
    bnd = naux * unit * (2.0 ** 32 + 3 * 2.0 ** 22) + 1e-300
    assert np.all(np.abs(g - ref) <= bnd)
    # and is far tighter than FP32 in practice: relative to |g| it is ~1e-9
    big = np.abs(ref) > 1e-3 * np.abs(ref).max()
    assert np.max(np.abs(g - ref)[big] / np.abs(ref)[big]) < 1e-6


def test_a_dropped_level_is_caught():
    """Keeping only levels 2..5 (a plausible off-by-one) moves g well outside
    what the kept levels give: the restatement's level count is pinned."""
    b = _b(4, naux=120, nov=30)
    e, d = quant.int8x4_digits(b)
    full = quant.int8x4_g(b)
    s6 = sum(d[p].T @ d[q] for p in range(4) for q in range(4) if p + q == 4)
    short = full - np.ldexp(s6.astype(np.float64) * 2.0 ** -32, e[:, None] + e[None, :] - 12)
    ref = b.T @ b
    assert np.max(np.abs(short - ref)) > 10 * np.max(np.abs(full - ref))


def test_energy_bound_holds_and_is_not_loose_by_orders(oracle_mod):
    """On a small problem the emulated format's energy error lies inside the
    propagated bound, and the bound is within a few orders of the error (a
    bound that came out as 0 or as |E| would pass nothing useful)."""
    rng = np.random.default_rng(9)
    nocc, nvir, naux = 4, 30, 160
    b = (rng.standard_normal((naux, nocc * nvir)) * np.sqrt(0.09 / naux))
    b[rng.choice(naux, 2, replace=False)] *= 20.0
    eps = np.concatenate([np.sort(rng.uniform(-1.3, -0.3, nocc)),
                          np.sort(np.exp(rng.uniform(np.log(0.1), np.log(6.0), nvir)))])
    ref = oracle_mod.spin_components_from_g(b.T @ b, eps, nocc, nvir)
    got = oracle_mod.spin_components_from_g(quant.int8x4_g(b), eps, nocc, nvir)
    bos, bss = quant.int8x4_energy_bound(b, eps, nocc, nvir)
    assert abs(got[0] - ref[0]) <= bos and abs(got[1] - ref[1]) <= bss
    assert 0 < bos < 1e-6 * abs(ref[0]) and 0 < bss < 1e-6 * abs(ref[1])
