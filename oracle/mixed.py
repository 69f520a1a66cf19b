"""The reference's mixed-precision G-block arithmetic and its error model, restated.

TEST INFRASTRUCTURE ONLY (the checker): imported by tests/ and never by the
product path.

* ``gemm_f32_f64acc_host`` restates the host twin of the device mixed GEMM
  (crates/ferric-core/src/gpu/mixed_host.rs:54-86): both operands rounded to
  f32, the contraction split into k-panels of ``k_panel`` rows, each panel an
  f32 product (numpy's float32 matmul, an sgemm like the reference's OpenBLAS
  one) started from zero, each panel's result added into an f64 accumulator.
  ``MIXED_K_PANEL_DEFAULT`` = 64 (crates/ferric-core/src/gpu/mixed.rs:70).
* ``gamma``, ``mixed_error_factor`` (mixed_host.rs:26-49) and the per-element
  factors of crates/ferric-mp2/tests/common/rimp2_eps.rs: ``eps_storage``,
  ``eps_device``, ``eps_accumulation`` (each multiplies
  S_ab = sum_P |B_P,ia||B_P,jb|) and ``eta_abs`` (the f32 underflow term).
* ``f32_across_panels_host``: the mutant the reference's flush test uses
  (crates/ferric-mp2/tests/gpu_rimp2_mixed.rs:411-447), the panels summed in f32
  instead of f64.
"""
from __future__ import annotations

import math

import numpy as np

U32 = 2.0 ** -24   # unit roundoff of binary32 (not machine epsilon)
U64 = 2.0 ** -53
MIXED_K_PANEL_DEFAULT = 64


def gamma(n: int, u: float) -> float:
    """gamma_n(u) = n u / (1 - n u)."""
    nu = n * u
    return nu / (1.0 - nu)


def mixed_error_factor(k: int, k_panel: int) -> float:
    """2 u32 + gamma_b(u32) + gamma_ceil(k/b)(u64), b clamped to [1, k]."""
    k = max(k, 1)
    b = min(max(k_panel, 1), k)
    return 2.0 * U32 + gamma(b, U32) + gamma(-(-k // b), U64)


def eps_storage(naux: int) -> float:
    """The f64 product of the f32-rounded B against the exact f64 product."""
    return 2.0 * U32 + U32 * U32 + gamma(naux, U64) * (1.0 + (1.0 + U32) * (1.0 + U32))


def eps_device(naux: int, b: int = MIXED_K_PANEL_DEFAULT) -> float:
    """The mixed device kernel against the exact f64 product (storage,
    f32 panels of b products, f64 sum of the panels, the f64 reference's own
    gamma_naux(u64))."""
    naux = max(naux, 1)
    b = min(max(b, 1), naux)
    panels = -(-naux // b)
    return ((1.0 + U32) * (1.0 + U32) * (1.0 + gamma(b, U32)) * (1.0 + gamma(panels, U64))
            - 1.0 + gamma(naux, U64))


def eps_accumulation(naux: int, b: int = MIXED_K_PANEL_DEFAULT) -> float:
    """The accumulation part alone: the device against the f64 product of the
    f32-ROUNDED B."""
    naux = max(naux, 1)
    b = min(max(b, 1), naux)
    panels = -(-naux // b)
    return (((1.0 + gamma(b, U32)) * (1.0 + gamma(panels, U64)) - 1.0 + gamma(naux, U64))
            * (1.0 + U32) * (1.0 + U32))


def eta_abs(naux: int, max_abs_b: float) -> float:
    """The f32 underflow term: naux 2^-149 (2 + 2 max|B|)."""
    return naux * math.ldexp(1.0, -149) * (2.0 + 2.0 * max_abs_b)


def gemm_f32_f64acc_host(left, right, k_panel: int = MIXED_K_PANEL_DEFAULT) -> np.ndarray:
    """out = left . right by the mixed algorithm (mixed_host.rs:54-86)."""
    left = np.asarray(left, dtype=np.float64)
    right = np.asarray(right, dtype=np.float64)
    m, k = left.shape
    k2, n = right.shape
    if k2 != k:
        raise ValueError("inner dimensions differ")
    out = np.zeros((m, n))
    if k == 0 or m == 0 or n == 0:
        return out
    a32 = left.astype(np.float32)
    b32 = right.astype(np.float32)
    kb = max(k_panel, 1)
    for k0 in range(0, k, kb):
        k1 = min(k0 + kb, k)
        c32 = a32[:, k0:k1] @ b32[k0:k1, :]
        out += c32.astype(np.float64)
    return out


def f32_across_panels_host(left, right, k_panel: int = MIXED_K_PANEL_DEFAULT) -> np.ndarray:
    """The mutant: the same f32 panels, but summed in f32 across panels."""
    a32 = np.asarray(left, dtype=np.float64).astype(np.float32)
    b32 = np.asarray(right, dtype=np.float64).astype(np.float32)
    k = a32.shape[1]
    acc = np.zeros((a32.shape[0], b32.shape[1]), dtype=np.float32)
    for k0 in range(0, k, max(k_panel, 1)):
        k1 = min(k0 + k_panel, k)
        acc = (acc + a32[:, k0:k1] @ b32[k0:k1, :]).astype(np.float32)
    return acc.astype(np.float64)


def flush_construction(k_panel: int = MIXED_K_PANEL_DEFAULT) -> np.ndarray:
    """B_ov (3 k_panel, 2) of gpu_rimp2_mixed.rs:418-429: column 0 all ones,
    column 1 holding +2^20, 2^-9 and -2^20 at the first row of three panels.
    Column 0 . column 1 is 2^-9 under the f64 flush and 0 under the mutant."""
    naux = 3 * k_panel
    x = np.zeros(naux)
    x[0] = 2.0 ** 20
    x[k_panel] = 2.0 ** -9
    x[2 * k_panel] = -(2.0 ** 20)
    return np.stack([np.ones(naux), x], axis=1)
