"""ctypes binding of the C ABI in include/ferric_hip.h (libferric_hip.so).

Loading fails loudly when the library has not been built: there is no
fallback implementation anywhere in this package.
"""
from __future__ import annotations

import ctypes
import enum
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libferric_hip.so")

_c_sz = ctypes.c_size_t
_c_dp = ctypes.POINTER(ctypes.c_double)
_c_vp = ctypes.c_void_p
_c_int = ctypes.c_int
_c_double = ctypes.c_double


class Status(enum.IntEnum):
    OK = 0
    LAYOUT = 1
    POOL_FULL = 2
    DEVICE = 3
    HIP = 4
    KERNEL = 5
    RANGE = 6
    UNSUPPORTED = 7


class Precision(enum.IntEnum):
    """Arithmetic of the (ia|jb) contraction (denominators and sums stay FP64)."""

    FP64 = 0
    FP32 = 1
    BF16X3 = 2
    FP8_E4M3 = 3
    INT8X4 = 4
    MIXED = 5    # FP32 over 64-row Q panels, FP64 across panels (the reference's Mixed)


class DType(enum.IntEnum):
    F64 = 0
    F32 = 1


class SpinComponentsC(ctypes.Structure):
    _fields_ = [("e_os", ctypes.c_double), ("e_ss", ctypes.c_double),
                ("e_total", ctypes.c_double)]


class StatsC(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in (
        "calls", "offloaded", "fp64_calls", "fp32_calls", "bf16x3_calls", "fp8_calls",
        "adjoint_calls", "batched_calls", "h2d_bytes", "d2h_bytes", "kernel_launches",
        "fail_layout", "fail_pool_full", "fail_hip", "fail_range", "fail_unsupported",
        "pool_peak_bytes", "int8x4_calls", "mixed_calls")]


class TimingC(ctypes.Structure):
    _fields_ = [("pack_ms", ctypes.c_double), ("contract_ms", ctypes.c_double),
                ("reduce_ms", ctypes.c_double), ("total_ms", ctypes.c_double),
                ("items", ctypes.c_uint64), ("tile", ctypes.c_uint32),
                ("reserved", ctypes.c_uint32)]


class DomainC(ctypes.Structure):
    _fields_ = [("b_offset", ctypes.c_uint64), ("eps_offset", ctypes.c_uint64),
                ("nocc", ctypes.c_uint32), ("nvir", ctypes.c_uint32),
                ("naux", ctypes.c_uint32), ("reserved", ctypes.c_uint32)]


class PnoDiagnosticsC(ctypes.Structure):
    _fields_ = [("kept_pnos", ctypes.c_uint64), ("max_discarded_weight", ctypes.c_double)]


class SpinChannelC(ctypes.Structure):
    _fields_ = [("b", ctypes.c_void_p), ("nocc", ctypes.c_size_t), ("nvir", ctypes.c_size_t),
                ("eps", ctypes.POINTER(ctypes.c_double)), ("n_eps", ctypes.c_size_t),
                ("first_occ", ctypes.c_size_t), ("nocc_total", ctypes.c_size_t)]


# GpuError taxonomy (crates/ferric-core/src/gpu/device.rs:16-36).
class GpuError(RuntimeError):
    status = Status.HIP


class LayoutError(GpuError):
    status = Status.LAYOUT


class PoolFullError(GpuError):
    status = Status.POOL_FULL


class DeviceError(GpuError):
    status = Status.DEVICE


class HipError(GpuError):
    status = Status.HIP


class KernelError(GpuError):
    status = Status.KERNEL


class RangeError(GpuError):
    status = Status.RANGE


class UnsupportedError(GpuError):
    status = Status.UNSUPPORTED


_ERRORS = {c.status: c for c in (LayoutError, PoolFullError, DeviceError, HipError, KernelError,
                                 RangeError, UnsupportedError)}

# Every symbol the header declares, with (restype, argtypes).
SIGNATURES = {
    "fh_abi_version": (_c_int, []),
    "fh_init": (_c_int, [_c_int]),
    "fh_last_error": (ctypes.c_char_p, []),
    "fh_set_pool_limit": (_c_int, [_c_sz]),
    "fh_get_pool_limit": (_c_int, [ctypes.POINTER(_c_sz)]),
    "fh_release_workspace": (_c_int, []),
    "fh_get_stats": (_c_int, [ctypes.POINTER(StatsC)]),
    "fh_reset_stats": (_c_int, []),
    "fh_tune_set": (_c_int, [_c_int, _c_sz, _c_sz, _c_int]),
    "fh_tune_clear": (_c_int, []),
    "fh_set_timing": (_c_int, [_c_int]),
    "fh_last_timing": (_c_int, [ctypes.POINTER(TimingC)]),
    "fh_rimp2_energy": (_c_int, [_c_dp, _c_sz, _c_sz, _c_sz, _c_dp, _c_sz, _c_sz, _c_sz,
                                 _c_double, _c_int, ctypes.POINTER(SpinComponentsC)]),
    "fh_rimp2_energy_dev": (_c_int, [_c_vp, _c_int, _c_sz, _c_sz, _c_sz, _c_dp, _c_sz, _c_sz,
                                     _c_sz, _c_double, _c_int, _c_vp,
                                     ctypes.POINTER(SpinComponentsC), _c_dp]),
    "fh_rimp2_energy_dev_noalloc": (_c_int, [_c_vp, _c_int, _c_sz, _c_sz, _c_sz, _c_dp, _c_sz,
                                             _c_sz, _c_sz, _c_double, _c_int, _c_vp,
                                             ctypes.POINTER(SpinComponentsC), _c_dp]),
    "fh_rimp2_mp2_density": (_c_int, [_c_dp, _c_sz, _c_sz, _c_sz, _c_dp, _c_sz, _c_sz, _c_sz,
                                      _c_double, _c_double, ctypes.POINTER(SpinComponentsC),
                                      _c_dp, _c_dp]),
    "fh_rimp2_mp2_density_dev": (_c_int, [_c_vp, _c_int, _c_sz, _c_sz, _c_sz, _c_dp, _c_sz,
                                          _c_sz, _c_sz, _c_double, _c_double, _c_vp,
                                          ctypes.POINTER(SpinComponentsC), _c_dp, _c_dp]),
    "fh_rimp2_energy_adjoint": (_c_int, [_c_dp, _c_sz, _c_sz, _c_sz, _c_dp, _c_sz, _c_sz,
                                         _c_sz, _c_double, _c_double, _c_int,
                                         ctypes.POINTER(SpinComponentsC), _c_dp, _c_dp]),
    "fh_rimp2_energy_adjoint_dev": (_c_int, [_c_vp, _c_int, _c_sz, _c_sz, _c_sz, _c_dp, _c_sz,
                                             _c_sz, _c_sz, _c_double, _c_double, _c_int, _c_vp,
                                             ctypes.POINTER(SpinComponentsC), _c_vp, _c_dp]),
    "fh_rimp2_batched_dev": (_c_int, [ctypes.POINTER(DomainC), _c_sz, _c_vp, _c_sz, _c_vp,
                                      _c_sz, _c_int, _c_int, _c_vp, _c_dp, _c_dp]),
    "fh_rimp2_pair_energies": (_c_int, [_c_dp, _c_sz, _c_sz, _c_sz, _c_dp, _c_sz, _c_sz, _c_sz,
                                        _c_int, _c_dp, _c_dp]),
    "fh_rimp2_pair_energies_dev": (_c_int, [_c_vp, _c_int, _c_sz, _c_sz, _c_sz, _c_dp, _c_sz,
                                            _c_sz, _c_sz, _c_int, _c_vp, _c_dp, _c_dp]),
    "fh_dlpno_pno_pair_energies": (_c_int, [_c_dp, _c_sz, _c_sz, _c_sz, _c_dp, _c_sz, _c_sz,
                                            _c_sz, _c_vp, _c_sz, ctypes.c_double, _c_dp, _c_dp,
                                            ctypes.POINTER(PnoDiagnosticsC)]),
    "fh_dlpno_pno_pair_energies_dev": (_c_int, [_c_vp, _c_sz, _c_sz, _c_sz, _c_dp, _c_sz, _c_sz,
                                                _c_sz, _c_vp, _c_sz, ctypes.c_double, _c_vp,
                                                _c_dp, _c_dp, ctypes.POINTER(PnoDiagnosticsC)]),
    "fh_domain_fit_pair_blocks": (_c_int, [_c_dp, _c_dp, _c_sz, _c_sz, _c_sz, _c_dp, _c_dp,
                                           ctypes.c_double, _c_vp, _c_sz, _c_dp,
                                           ctypes.POINTER(_c_sz)]),
    "fh_domain_fit_pair_blocks_dev": (_c_int, [_c_vp, _c_vp, _c_sz, _c_sz, _c_sz, _c_dp, _c_dp,
                                               ctypes.c_double, _c_vp, _c_sz, _c_vp, _c_vp,
                                               ctypes.POINTER(_c_sz)]),
    "fh_rimp2_energy_from_g": (_c_int, [_c_dp, _c_sz, _c_sz, _c_dp, _c_sz, _c_sz, _c_sz,
                                        ctypes.POINTER(SpinComponentsC)]),
    "fh_rimp2_energy_from_g_dev": (_c_int, [_c_vp, _c_sz, _c_sz, _c_dp, _c_sz, _c_sz, _c_sz,
                                            _c_vp, ctypes.POINTER(SpinComponentsC)]),
    "fh_pair_energies_from_blocks_dev": (_c_int, [_c_vp, _c_vp, _c_sz, _c_sz, _c_sz, _c_dp,
                                                  _c_sz, _c_sz, _c_sz, _c_vp, _c_dp, _c_dp]),
    "fh_pair_energies_from_g_dev": (_c_int, [_c_vp, _c_sz, _c_sz, _c_dp, _c_sz, _c_sz, _c_sz,
                                             _c_vp, _c_dp, _c_dp]),
    "fh_dlpno_pno_pair_energies_from_g_dev": (_c_int, [_c_vp, _c_sz, _c_sz, _c_dp, _c_sz, _c_sz,
                                                       _c_sz, _c_vp, _c_sz, ctypes.c_double,
                                                       _c_vp, _c_dp, _c_dp,
                                                       ctypes.POINTER(PnoDiagnosticsC)]),
    "fh_dress_b_ov_dev": (_c_int, [_c_vp, _c_sz, _c_sz, _c_vp, _c_sz, _c_vp, _c_sz, _c_vp, _c_sz,
                                   _c_vp, _c_vp]),
    "fh_u_same_spin": (_c_int, [ctypes.POINTER(SpinChannelC), _c_sz, _c_int, _c_dp]),
    "fh_u_same_spin_dev": (_c_int, [ctypes.POINTER(SpinChannelC), _c_int, _c_sz, _c_int, _c_vp,
                                    _c_dp]),
    "fh_u_opposite_spin": (_c_int, [ctypes.POINTER(SpinChannelC), ctypes.POINTER(SpinChannelC),
                                    _c_sz, _c_int, _c_dp]),
    "fh_u_opposite_spin_dev": (_c_int, [ctypes.POINTER(SpinChannelC),
                                        ctypes.POINTER(SpinChannelC), _c_int, _c_sz, _c_int,
                                        _c_vp, _c_dp]),
    "fh_rimp2_energy_signed": (_c_int, [_c_dp, _c_sz, _c_dp, _c_sz, _c_sz, _c_sz, _c_dp, _c_sz,
                                        _c_sz, _c_sz, _c_int, ctypes.POINTER(SpinComponentsC)]),
    "fh_rimp2_energy_signed_dev": (_c_int, [_c_vp, _c_sz, _c_vp, _c_sz, _c_int, _c_sz, _c_sz,
                                            _c_dp, _c_sz, _c_sz, _c_sz, _c_int, _c_vp,
                                            ctypes.POINTER(SpinComponentsC)]),
}

_lib = None
ABI_VERSION = 11  # FERRIC_HIP_ABI_VERSION of include/ferric_hip.h


class LibraryNotBuilt(ImportError):
    pass


def lib() -> ctypes.CDLL:
    """The loaded library (raises LibraryNotBuilt if it is absent)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise LibraryNotBuilt(
                f"{LIB_PATH} is missing: build it with `python -m ferric_amd.build` "
                "(there is no fallback implementation)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        if L.fh_abi_version() != ABI_VERSION:
            raise LibraryNotBuilt(
                f"{LIB_PATH} has ABI version {L.fh_abi_version()}, this binding expects "
                f"{ABI_VERSION}: rebuild it with `python -m ferric_amd.build`")
        _lib = L
        from . import autotune

        autotune.install(autotune.load_table())
    return _lib


def last_error() -> str:
    msg = lib().fh_last_error()
    return msg.decode() if msg else ""


def check(status: int, what: str = "") -> None:
    if status == Status.OK:
        return
    cls = _ERRORS.get(Status(status), GpuError)
    raise cls(f"{what}: {Status(status).name}: {last_error()}")


def set_timing(on: bool) -> None:
    check(lib().fh_set_timing(1 if on else 0), "fh_set_timing")


def last_timing() -> dict:
    t = TimingC()
    check(lib().fh_last_timing(ctypes.byref(t)), "fh_last_timing")
    return {n: getattr(t, n) for n, _ in TimingC._fields_}


def stats() -> dict:
    s = StatsC()
    check(lib().fh_get_stats(ctypes.byref(s)), "fh_get_stats")
    return {n: getattr(s, n) for n, _ in StatsC._fields_}
