"""Unrestricted (open-shell) RI-MP2 blocks and the signed factor on the device.

Mirrors the reference's open-shell energy surface:

* :class:`SpinChannel` -- one spin's ``(b, eps, nocc, nvir, first_occ,
  nocc_total)`` (crates/ferric-mp2/src/u_rimp2.rs:669-679);
* :func:`try_u_same_spin_on_device` / :func:`try_u_opposite_spin_on_device`
  (crates/ferric-mp2/src/rimp2_gpu.rs:884-902) -- the energy or ``None`` for
  the caller's CPU path, after the same checks as the closed-shell dispatcher;
* :func:`u_same_spin` / :func:`u_opposite_spin` -- the device energies
  (u_rimp2.rs:693-760), raising on failure;
* :func:`spin_components_from_signed_b_ov` -- the closed-shell components of
  ``K = pos^T pos - neg^T neg`` (rimp2.rs:1559-1607).

``b`` may be a numpy array (host entry, staged by the library) or a
contiguous CUDA tensor (device entry on the current torch stream).
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import GpuError, LayoutError, Precision, check, lib
from .rimp2 import SpinComponents, _is_torch_cuda, settings_precision


@dataclass
class SpinChannel:
    b: object             # (naux, nocc*nvir) numpy array or CUDA tensor
    eps: object           # full per-spin orbital energies
    nocc: int
    nvir: int
    first_occ: int = 0
    nocc_total: int | None = None

    def __post_init__(self):
        if self.nocc_total is None:
            self.nocc_total = self.first_occ + self.nocc


def _b_shape(b, nocc, nvir):
    if _is_torch_cuda(b):
        import torch

        if b.dtype not in (torch.float64, torch.float32):
            raise LayoutError(f"b dtype {b.dtype} is not float64/float32")
        if b.dim() != 2 or not b.is_contiguous():
            raise LayoutError("b must be a contiguous 2-D (naux, nocc*nvir) tensor")
        naux, width = b.shape
    else:
        b = np.ascontiguousarray(np.asarray(b, dtype=np.float64))
        if b.ndim != 2:
            raise LayoutError("b must be 2-D (naux, nocc*nvir)")
        naux, width = b.shape
    if width != nocc * nvir:
        raise LayoutError(f"b is {naux}x{width} but nocc*nvir = {nocc * nvir}")
    return b, naux


class _Chan:
    """A SpinChannel marshalled for the C ABI (keeps its arrays alive)."""

    def __init__(self, ch: SpinChannel):
        self.b, self.naux = _b_shape(ch.b, ch.nocc, ch.nvir)
        self.dev = _is_torch_cuda(self.b)
        self.eps = np.ascontiguousarray(np.asarray(ch.eps, dtype=np.float64))
        ptr = self.b.data_ptr() if self.dev else self.b.ctypes.data
        self.c = _lib.SpinChannelC(ctypes.c_void_p(ptr), ch.nocc, ch.nvir,
                                   self.eps.ctypes.data_as(_lib._c_dp), self.eps.size,
                                   ch.first_occ, ch.nocc_total)

    def dtype(self):
        import torch

        return _lib.DType.F64 if self.b.dtype == torch.float64 else _lib.DType.F32

    def stream(self):
        import torch

        return ctypes.c_void_p(torch.cuda.current_stream(self.b.device).cuda_stream)


def u_same_spin(ch: SpinChannel, precision: Precision = Precision.FP64) -> float:
    """Same-spin energy 1/4 sum_{i<=j} f_ij sum_ab (g_ab - g_ba)^2 / D of one
    channel (u_rimp2.rs:693-729)."""
    c = _Chan(ch)
    out = ctypes.c_double()
    L = lib()
    if c.dev:
        st = L.fh_u_same_spin_dev(ctypes.byref(c.c), int(c.dtype()), c.naux, int(precision),
                                  c.stream(), ctypes.byref(out))
    else:
        st = L.fh_u_same_spin(ctypes.byref(c.c), c.naux, int(precision), ctypes.byref(out))
    check(st, "fh_u_same_spin")
    return out.value


def u_opposite_spin(ch_a: SpinChannel, ch_b: SpinChannel,
                    precision: Precision = Precision.FP64) -> float:
    """Opposite-spin energy sum_{i,J,a,B} (ia|JB)^2 / D (u_rimp2.rs:731-760);
    both channels fitted in one aux basis. FP64, FP32 or INT8X4."""
    a, b = _Chan(ch_a), _Chan(ch_b)
    if a.naux != b.naux:
        raise LayoutError(f"alpha has {a.naux} aux rows, beta {b.naux}: one aux basis is required")
    if a.dev != b.dev:
        raise LayoutError("both channels must be host arrays or both device tensors")
    out = ctypes.c_double()
    L = lib()
    if a.dev:
        if a.b.dtype != b.b.dtype:
            raise LayoutError("both channels must have one dtype")
        st = L.fh_u_opposite_spin_dev(ctypes.byref(a.c), ctypes.byref(b.c), int(a.dtype()),
                                      a.naux, int(precision), a.stream(), ctypes.byref(out))
    else:
        st = L.fh_u_opposite_spin(ctypes.byref(a.c), ctypes.byref(b.c), a.naux, int(precision),
                                  ctypes.byref(out))
    check(st, "fh_u_opposite_spin")
    return out.value


def _mixed_precision(mixed_ok: bool) -> Precision:
    return settings_precision() if mixed_ok else Precision.FP64


def try_u_same_spin_on_device(ch: SpinChannel, mixed_ok: bool = False):
    """rimp2_gpu.rs:884-889: the energy, or None for the CPU path (FERRIC_GPU=off
    or a device failure). A reduced-precision range refusal retries in FP64 on
    the device, as the reference's fallback ladder does."""
    if os.environ.get("FERRIC_GPU", "").lower() == "off":
        return None
    prec = _mixed_precision(mixed_ok)
    try:
        try:
            return u_same_spin(ch, prec)
        except _lib.RangeError:
            return u_same_spin(ch, Precision.FP64)
    except GpuError:
        return None


def try_u_opposite_spin_on_device(ch_a: SpinChannel, ch_b: SpinChannel, mixed_ok: bool = False):
    """rimp2_gpu.rs:893-902: the energy, or None for the CPU path. The
    opposite-spin block runs in FP64 (FP32 or INT8X4 when mixed precision is
    allowed and selected)."""
    if os.environ.get("FERRIC_GPU", "").lower() == "off":
        return None
    prec = _mixed_precision(mixed_ok)
    if prec not in (Precision.FP64, Precision.FP32, Precision.INT8X4):
        prec = Precision.FP64
    try:
        try:
            return u_opposite_spin(ch_a, ch_b, prec)
        except _lib.RangeError:
            return u_opposite_spin(ch_a, ch_b, Precision.FP64)
    except GpuError:
        return None


def u_mp2_components(ch_a: SpinChannel, ch_b: SpinChannel,
                     precision: Precision = Precision.FP64):
    """(E_aa, E_bb, E_ab): the three blocks of the UMP2 correlation energy."""
    return (u_same_spin(ch_a, precision), u_same_spin(ch_b, precision),
            u_opposite_spin(ch_a, ch_b, precision))


def spin_components_from_signed_b_ov(pos, neg, eps, nocc: int, nvir: int, first_occ: int = 0,
                                     nocc_total: int | None = None,
                                     precision: Precision = Precision.FP64) -> SpinComponents:
    """Closed-shell spin components of K = pos^T pos - neg^T neg
    (rimp2.rs:1559-1607); pos and neg share the column layout (i*nvir + a) and
    may differ in row count. FP64 or FP32."""
    if nocc_total is None:
        nocc_total = first_occ + nocc
    bp, naux_p = _b_shape(pos, nocc, nvir)
    bn, naux_n = _b_shape(neg, nocc, nvir)
    e = np.ascontiguousarray(np.asarray(eps, dtype=np.float64))
    out = _lib.SpinComponentsC()
    L = lib()
    if _is_torch_cuda(bp):
        import torch

        if not _is_torch_cuda(bn) or bn.dtype != bp.dtype:
            raise LayoutError("pos and neg must be device tensors of one dtype")
        dt = _lib.DType.F64 if bp.dtype == torch.float64 else _lib.DType.F32
        stream = torch.cuda.current_stream(bp.device).cuda_stream
        st = L.fh_rimp2_energy_signed_dev(
            ctypes.c_void_p(bp.data_ptr()), naux_p, ctypes.c_void_p(bn.data_ptr()), naux_n,
            int(dt), nocc, nvir, e.ctypes.data_as(_lib._c_dp), e.size, first_occ, nocc_total,
            int(precision), ctypes.c_void_p(stream), ctypes.byref(out))
    else:
        st = L.fh_rimp2_energy_signed(
            bp.ctypes.data_as(_lib._c_dp), naux_p, bn.ctypes.data_as(_lib._c_dp), naux_n, nocc,
            nvir, e.ctypes.data_as(_lib._c_dp), e.size, first_occ, nocc_total, int(precision),
            ctypes.byref(out))
    check(st, "fh_rimp2_energy_signed")
    return SpinComponents(out.e_os, out.e_ss, out.e_total)
