"""Host mirror of ferric's RI-MP2 pair-energy surface over the HIP library.

Names, argument meaning and error behaviour follow crates/ferric-mp2/src/rimp2.rs:

* ``spin_components_from_b_ov``          rimp2.rs:1311-1320 (FP64 device)
* ``spin_components_from_b_ov_kappa``    rimp2.rs:1330-1341 (may use reduced precision)
* ``spin_components_from_b_ov_kappa_f64`` rimp2.rs:1343-1357 (FP64 device)
* ``try_spin_components_on_device``      rimp2_gpu.rs:105-170 (None means "run the CPU path")
* ``active_occ``                          rimp2.rs:99-106
* ``scs_mp2`` / ``ScsMp2Config``          crates/ferric-mp2/src/scs.rs:27-49, :93-114

``b_ov`` is row-major (naux, nocc*nvir) over ACTIVE occupied orbitals; ``eps`` is
the full MO energy vector (occupied i at eps[first_occ + i], virtual a at
eps[nocc_total + a]). ``b_ov`` may be a host numpy array (FP64; uploaded for the
call, as the reference's Stage does) or a device torch tensor (FP64 or FP32;
used in place). Every call runs on the GPU; failures raise the GpuError
hierarchy of :mod:`ferric_amd._lib` -- this package has no CPU path.
"""
from __future__ import annotations

import ctypes
import math
import os
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import GpuError, LayoutError, Precision, check, lib

__all__ = [
    "SpinComponents", "Precision", "active_occ", "spin_components_from_b_ov",
    "spin_components_from_b_ov_kappa", "spin_components_from_b_ov_kappa_f64",
    "try_spin_components_on_device", "spin_components", "ScsMp2Config", "scs_mp2",
    "settings_precision", "energy_adjoint", "mixed_energy_operator", "mixed_ok",
    "spin_components_fitted", "mp2_density",
]


@dataclass(frozen=True)
class SpinComponents:
    """Spin-component resolved MP2 correlation energy (rimp2.rs:134-141)."""

    e_os: float
    e_ss: float
    e_total: float


def active_occ(nocc_total: int, frozen_core: int) -> int:
    """Active occupied count; an error when the freeze covers every occupied
    orbital (``frozen_core == 0`` with zero occupied is allowed), rimp2.rs:99-106."""
    if frozen_core != 0 and frozen_core >= nocc_total:
        raise LayoutError(
            f"frozen_core = {frozen_core} freezes all {nocc_total} occupied orbitals — "
            "nothing left to correlate")
    return nocc_total - frozen_core


_PREC_NAMES = {
    "f64": Precision.FP64, "fp64": Precision.FP64,
    "f32": Precision.FP32, "fp32": Precision.FP32, "mixed": Precision.MIXED,
    "bf16x3": Precision.BF16X3, "bf16": Precision.BF16X3,
    "fp8": Precision.FP8_E4M3, "fp8_e4m3": Precision.FP8_E4M3,
    "int8x4": Precision.INT8X4, "int8": Precision.INT8X4,
}


def settings_precision() -> Precision:
    """Reduced precision allowed for the energy kernel by the process settings.

    ``FERRIC_GPU_PRECISION`` (f64 | mixed | fp32 | bf16x3 | fp8 | int8x4), the reference's
    precision knob (crates/ferric-core/src/gpu/precision.rs:283); default f64.
    ``mixed`` is the reference's own mixed arithmetic (FP32 over 64-row Q panels
    summed in FP64, Precision.MIXED); the other reduced modes are this library's.
    A malformed value is an error, as with the reference's ConfigVar.
    """
    v = os.environ.get("FERRIC_GPU_PRECISION", "f64").strip().lower()
    if v not in _PREC_NAMES:
        raise LayoutError(f"FERRIC_GPU_PRECISION={v!r} is not one of {sorted(_PREC_NAMES)}")
    return _PREC_NAMES[v]


def _kappa_c(kappa) -> float:
    return float("nan") if kappa is None else float(kappa)


def _is_torch_cuda(x) -> bool:
    return type(x).__module__.startswith("torch") and getattr(x, "is_cuda", False)


def spin_components(b_ov, eps, nocc: int, nvir: int, first_occ: int = 0,
                    nocc_total: int | None = None, kappa=None,
                    precision: Precision = Precision.FP64, row_partials: bool = False):
    """The device energy. Returns SpinComponents (and the per-i partials when
    ``row_partials``). Raises GpuError subclasses on failure."""
    if nocc_total is None:
        nocc_total = first_occ + nocc
    precision = Precision(precision)
    e = np.ascontiguousarray(np.asarray(eps, dtype=np.float64))
    out = _lib.SpinComponentsC()
    L = lib()
    if _is_torch_cuda(b_ov):
        import torch

        if b_ov.dtype == torch.float64:
            dt = _lib.DType.F64
        elif b_ov.dtype == torch.float32:
            dt = _lib.DType.F32
        else:
            raise LayoutError(f"b_ov dtype {b_ov.dtype} is not float64/float32")
        if b_ov.dim() != 2 or not b_ov.is_contiguous():
            raise LayoutError("b_ov must be a contiguous 2-D (naux, nocc*nvir) tensor")
        naux, width = b_ov.shape
        if width != nocc * nvir:
            raise LayoutError(f"b_ov is {naux}x{width} but nocc*nvir = {nocc * nvir}")
        rows = np.zeros(2 * max(nocc, 1))
        stream = torch.cuda.current_stream(b_ov.device).cuda_stream
        st = L.fh_rimp2_energy_dev(
            ctypes.c_void_p(b_ov.data_ptr()), int(dt), naux, nocc, nvir,
            e.ctypes.data_as(_lib._c_dp), e.size, first_occ, nocc_total, _kappa_c(kappa),
            int(precision), ctypes.c_void_p(stream), ctypes.byref(out),
            rows.ctypes.data_as(_lib._c_dp))
        check(st, "fh_rimp2_energy_dev")
        sc = SpinComponents(out.e_os, out.e_ss, out.e_total)
        if row_partials:
            return sc, rows[:2 * nocc].reshape(nocc, 2)
        return sc
    b = np.ascontiguousarray(np.asarray(b_ov, dtype=np.float64))
    if b.ndim != 2:
        raise LayoutError("b_ov must be 2-D (naux, nocc*nvir)")
    naux, width = b.shape
    if width != nocc * nvir:
        raise LayoutError(f"b_ov is {naux}x{width} but nocc*nvir = {nocc * nvir}")
    if row_partials:
        import torch

        bt = torch.from_numpy(b).cuda()
        return spin_components(bt, e, nocc, nvir, first_occ, nocc_total, kappa, precision, True)
    st = L.fh_rimp2_energy(b.ctypes.data_as(_lib._c_dp), naux, nocc, nvir,
                           e.ctypes.data_as(_lib._c_dp), e.size, first_occ, nocc_total,
                           _kappa_c(kappa), int(precision), ctypes.byref(out))
    check(st, "fh_rimp2_energy")
    return SpinComponents(out.e_os, out.e_ss, out.e_total)


def energy_adjoint(b_ov, eps, nocc: int, nvir: int, first_occ: int = 0,
                   nocc_total: int | None = None, c_os: float = 1.0, c_ss: float = 1.0,
                   precision: Precision = Precision.FP64):
    """E = c_os E_os + c_ss E_ss and its adjoint on the device.

    Returns ``(SpinComponents, E, gamma, deps)``: ``gamma`` = dE/dB in the layout
    of ``b_ov`` (naux, nocc*nvir) -- 4 x_ov of the reference's gradient
    intermediates (crates/ferric-mp2/src/rimp2.rs:1832-1840, zvector.rs:651-697)
    -- and ``deps`` = dE/deps over the full orbital-energy vector (2 diag P on the
    active occupied and virtual blocks, 0 for frozen core). A device tensor
    ``b_ov`` gives a device FP64 ``gamma``; a host array gives a numpy one.
    Precision FP64, BF16X3, or INT8X4 (the energy and T~ from the INT8 digit-slice
    kernel, the gamma GEMMs in BF16x3).
    """
    if nocc_total is None:
        nocc_total = first_occ + nocc
    precision = Precision(precision)
    e = np.ascontiguousarray(np.asarray(eps, dtype=np.float64))
    deps = np.zeros(e.size)
    out = _lib.SpinComponentsC()
    L = lib()
    if _is_torch_cuda(b_ov):
        import torch

        if b_ov.dtype == torch.float64:
            dt = _lib.DType.F64
        elif b_ov.dtype == torch.float32:
            dt = _lib.DType.F32
        else:
            raise LayoutError(f"b_ov dtype {b_ov.dtype} is not float64/float32")
        if b_ov.dim() != 2 or not b_ov.is_contiguous():
            raise LayoutError("b_ov must be a contiguous 2-D (naux, nocc*nvir) tensor")
        naux, width = b_ov.shape
        if width != nocc * nvir:
            raise LayoutError(f"b_ov is {naux}x{width} but nocc*nvir = {nocc * nvir}")
        gamma = torch.empty((naux, width), dtype=torch.float64, device=b_ov.device)
        stream = torch.cuda.current_stream(b_ov.device).cuda_stream
        st = L.fh_rimp2_energy_adjoint_dev(
            ctypes.c_void_p(b_ov.data_ptr()), int(dt), naux, nocc, nvir,
            e.ctypes.data_as(_lib._c_dp), e.size, first_occ, nocc_total, float(c_os),
            float(c_ss), int(precision), ctypes.c_void_p(stream), ctypes.byref(out),
            ctypes.c_void_p(gamma.data_ptr()), deps.ctypes.data_as(_lib._c_dp))
        check(st, "fh_rimp2_energy_adjoint_dev")
    else:
        b = np.ascontiguousarray(np.asarray(b_ov, dtype=np.float64))
        if b.ndim != 2:
            raise LayoutError("b_ov must be 2-D (naux, nocc*nvir)")
        naux, width = b.shape
        if width != nocc * nvir:
            raise LayoutError(f"b_ov is {naux}x{width} but nocc*nvir = {nocc * nvir}")
        gamma = np.zeros_like(b)
        st = L.fh_rimp2_energy_adjoint(
            b.ctypes.data_as(_lib._c_dp), naux, nocc, nvir, e.ctypes.data_as(_lib._c_dp),
            e.size, first_occ, nocc_total, float(c_os), float(c_ss), int(precision),
            ctypes.byref(out), gamma.ctypes.data_as(_lib._c_dp), deps.ctypes.data_as(_lib._c_dp))
        check(st, "fh_rimp2_energy_adjoint")
    sc = SpinComponents(out.e_os, out.e_ss, out.e_total)
    return sc, c_os * sc.e_os + c_ss * sc.e_ss, gamma, deps


def mp2_density(b_ov, eps, nocc: int, nvir: int, first_occ: int = 0,
                nocc_total: int | None = None, c_os: float = 1.0, c_ss: float = 1.0):
    """The MP2 one-particle densities on the device: ``(SpinComponents, P_oo,
    P_vv)`` with P_oo (nocc, nocc) and P_vv (nvir, nvir) over the active
    orbitals, build_mp2_density (crates/ferric-mp2/src/oo_rimp2.rs:1192-1300)
    for (c_os, c_ss) = (1, 1), its spin-scaled form (rimp2.rs:1917-2052)
    otherwise. ``b_ov``: host numpy array or contiguous CUDA tensor."""
    if nocc_total is None:
        nocc_total = first_occ + nocc
    e = np.ascontiguousarray(np.asarray(eps, dtype=np.float64))
    p_oo = np.zeros((nocc, nocc))
    p_vv = np.zeros((nvir, nvir))
    out = _lib.SpinComponentsC()
    L = lib()
    if _is_torch_cuda(b_ov):
        import torch

        if b_ov.dtype not in (torch.float64, torch.float32):
            raise LayoutError(f"b_ov dtype {b_ov.dtype} is not float64/float32")
        if b_ov.dim() != 2 or not b_ov.is_contiguous():
            raise LayoutError("b_ov must be a contiguous 2-D (naux, nocc*nvir) tensor")
        naux, width = b_ov.shape
        if width != nocc * nvir:
            raise LayoutError(f"b_ov is {naux}x{width} but nocc*nvir = {nocc * nvir}")
        dt = _lib.DType.F64 if b_ov.dtype == torch.float64 else _lib.DType.F32
        stream = torch.cuda.current_stream(b_ov.device).cuda_stream
        st = L.fh_rimp2_mp2_density_dev(
            ctypes.c_void_p(b_ov.data_ptr()), int(dt), naux, nocc, nvir,
            e.ctypes.data_as(_lib._c_dp), e.size, first_occ, nocc_total, float(c_os), float(c_ss),
            ctypes.c_void_p(stream), ctypes.byref(out), p_oo.ctypes.data_as(_lib._c_dp),
            p_vv.ctypes.data_as(_lib._c_dp))
        check(st, "fh_rimp2_mp2_density_dev")
    else:
        b = np.ascontiguousarray(np.asarray(b_ov, dtype=np.float64))
        if b.ndim != 2:
            raise LayoutError("b_ov must be 2-D (naux, nocc*nvir)")
        naux, width = b.shape
        if width != nocc * nvir:
            raise LayoutError(f"b_ov is {naux}x{width} but nocc*nvir = {nocc * nvir}")
        st = L.fh_rimp2_mp2_density(
            b.ctypes.data_as(_lib._c_dp), naux, nocc, nvir, e.ctypes.data_as(_lib._c_dp), e.size,
            first_occ, nocc_total, float(c_os), float(c_ss), ctypes.byref(out),
            p_oo.ctypes.data_as(_lib._c_dp), p_vv.ctypes.data_as(_lib._c_dp))
        check(st, "fh_rimp2_mp2_density")
    return SpinComponents(out.e_os, out.e_ss, out.e_total), p_oo, p_vv


def spin_components_from_g(g, eps, nocc: int, nvir: int, first_occ: int = 0,
                           nocc_total: int | None = None) -> SpinComponents:
    """rimp2.rs:984-1018 on the device: the components from an explicit dense
    (nocc*nvir) x (nocc*nvir) (ia|jb) matrix (the robust fit's non-Gram G).
    ``g``: numpy array (host entry) or contiguous float64 CUDA tensor (device
    entry on the current torch stream). FP64."""
    if nocc_total is None:
        nocc_total = first_occ + nocc
    e = np.ascontiguousarray(np.asarray(eps, dtype=np.float64))
    nov = nocc * nvir
    out = _lib.SpinComponentsC()
    L = lib()
    if _is_torch_cuda(g):
        import torch

        if g.dtype != torch.float64 or not g.is_contiguous() or tuple(g.shape) != (nov, nov):
            raise LayoutError(f"g must be a contiguous float64 ({nov}, {nov}) tensor")
        stream = torch.cuda.current_stream(g.device).cuda_stream
        st = L.fh_rimp2_energy_from_g_dev(ctypes.c_void_p(g.data_ptr()), nocc, nvir,
                                          e.ctypes.data_as(_lib._c_dp), e.size, first_occ,
                                          nocc_total, ctypes.c_void_p(stream), ctypes.byref(out))
    else:
        gn = np.ascontiguousarray(np.asarray(g, dtype=np.float64))
        if gn.shape != (nov, nov):
            raise LayoutError(f"g is {gn.shape}, expected ({nov}, {nov})")
        st = L.fh_rimp2_energy_from_g(gn.ctypes.data_as(_lib._c_dp), nocc, nvir,
                                      e.ctypes.data_as(_lib._c_dp), e.size, first_occ,
                                      nocc_total, ctypes.byref(out))
    check(st, "fh_rimp2_energy_from_g")
    return SpinComponents(out.e_os, out.e_ss, out.e_total)


# Operator kinds of ferric_integrals::operator::OperatorKind that reach the
# RI-MP2 energy (crates/ferric-integrals/src/operator.rs:5-40), by name.
MIXED_ENERGY_OPERATORS = frozenset({"coulomb", "erfc_coulomb", "terfc"})
OPERATOR_KINDS = MIXED_ENERGY_OPERATORS | frozenset(
    {"erf_coulomb", "terf", "yukawa", "slater_geminal", "cgtg", "gaussian", "delta"})


def mixed_energy_operator(kind: str = "coulomb", composite: bool = False) -> bool:
    """rimp2.rs:1293-1300: the mixed kernel's error map was measured for the
    primitive Coulomb, erfc and terfc operators only; composite (fitted)
    operators and every other kind keep the device in FP64."""
    k = str(kind).strip().lower()
    if k not in OPERATOR_KINDS:
        raise LayoutError(f"unknown operator kind {kind!r}; one of {sorted(OPERATOR_KINDS)}")
    return (not composite) and k in MIXED_ENERGY_OPERATORS


def mixed_ok(kappa=None, fit_metric: str = "coulomb", op_kind: str = "coulomb",
             composite: bool = False) -> bool:
    """The caller-side rule of ri_mp2_spin_components_fitted (rimp2.rs:1277):
    reduced precision only without kappa regularisation, with the Coulomb fit
    metric (omega-RI has no measured error map) and a mixed-energy operator."""
    fm = str(fit_metric).strip().lower()
    if fm not in ("coulomb", "erfc"):
        raise LayoutError(f"unknown fit metric {fit_metric!r}; expected 'coulomb' or 'erfc'")
    return kappa is None and fm == "coulomb" and mixed_energy_operator(op_kind, composite)


def spin_components_fitted(b_ov, eps, nocc, nvir, first_occ, nocc_total, kappa=None,
                           fit_metric: str = "coulomb", op_kind: str = "coulomb",
                           composite: bool = False) -> SpinComponents:
    """The pair-energy call of ri_mp2_spin_components_fitted (rimp2.rs:1266-1281):
    kappa validated (finite and > 0), ``mixed_ok`` from the operator and fit
    metric, then dispatch_spin_components."""
    if kappa is not None and not (math.isfinite(kappa) and kappa > 0.0):
        raise LayoutError(f"RiMp2Config.kappa must be finite and > 0 (got {kappa}); "
                          "use None for plain MP2")
    ok = mixed_ok(kappa, fit_metric, op_kind, composite)
    return _dispatch(b_ov, eps, nocc, nvir, first_occ, nocc_total, kappa, ok)


def _dispatch(b_ov, eps, nocc, nvir, first_occ, nocc_total, kappa, mixed_ok) -> SpinComponents:
    """dispatch_spin_components (rimp2.rs:1360-1375) minus the CPU branch: the
    reduced-precision kernel only when ``mixed_ok`` and the settings allow it."""
    prec = settings_precision() if mixed_ok else Precision.FP64
    try:
        return spin_components(b_ov, eps, nocc, nvir, first_occ, nocc_total, kappa, prec)
    except _lib.RangeError:
        if prec == Precision.FP64:
            raise
        # Reduced precision refused (operand range): FP64 on the device, as
        # the reference's mixed -> f64 fallback (rimp2_gpu.rs:149-161).
        return spin_components(b_ov, eps, nocc, nvir, first_occ, nocc_total, kappa,
                               Precision.FP64)


def spin_components_from_b_ov(b_ov, eps, nocc, nvir, first_occ, nocc_total) -> SpinComponents:
    """rimp2.rs:1311-1320: FP64 device, plain MP2."""
    return _dispatch(b_ov, eps, nocc, nvir, first_occ, nocc_total, None, False)


def spin_components_from_b_ov_kappa(b_ov, eps, nocc, nvir, first_occ, nocc_total,
                                    kappa=None) -> SpinComponents:
    """rimp2.rs:1330-1341: reduced precision allowed by the settings."""
    return _dispatch(b_ov, eps, nocc, nvir, first_occ, nocc_total, kappa, True)


def spin_components_from_b_ov_kappa_f64(b_ov, eps, nocc, nvir, first_occ, nocc_total,
                                        kappa=None) -> SpinComponents:
    """rimp2.rs:1343-1357: the device stays FP64."""
    return _dispatch(b_ov, eps, nocc, nvir, first_occ, nocc_total, kappa, False)


def try_spin_components_on_device(b_ov, eps, nocc, nvir, first_occ, nocc_total, kappa=None,
                                  mixed_ok: bool = False):
    """rimp2_gpu.rs:105-170: SpinComponents, or None when the caller should run
    its own CPU path (every GpuError, including a missing device)."""
    if os.environ.get("FERRIC_GPU", "auto").strip().lower() == "off":
        return None
    try:
        return _dispatch(b_ov, eps, nocc, nvir, first_occ, nocc_total, kappa, mixed_ok)
    except GpuError:
        return None


@dataclass(frozen=True)
class ScsMp2Config:
    """SCS coefficients (crates/ferric-mp2/src/scs.rs:27-49; Grimme's 6/5, 1/3)."""

    c_os: float = 6.0 / 5.0
    c_ss: float = 1.0 / 3.0


def scs_mp2(sc: SpinComponents, cfg: ScsMp2Config = ScsMp2Config()) -> float:
    """E = c_os E_os + c_ss E_ss, applied after the kernel (scs.rs:105-106)."""
    if not (math.isfinite(cfg.c_os) and math.isfinite(cfg.c_ss)):
        raise LayoutError("SCS coefficients must be finite")
    return cfg.c_os * sc.e_os + cfg.c_ss * sc.e_ss
