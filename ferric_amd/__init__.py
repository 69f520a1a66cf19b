"""ferric_amd: MI355X-native RI-MP2 pair-energy path (gfx950 HIP kernels).

The compute lives in ``lib/libferric_hip.so`` (C ABI: ``include/ferric_hip.h``);
this package is its host-side mirror of ferric's Python/Rust surface.
"""
from ._lib import (DeviceError, GpuError, HipError, KernelError, LayoutError,  # noqa: F401
                   LibraryNotBuilt, PoolFullError, Precision, RangeError, UnsupportedError,
                   lib, stats)
from .rimp2 import (ScsMp2Config, SpinComponents, active_occ, scs_mp2,  # noqa: F401
                    spin_components, spin_components_from_b_ov,
                    spin_components_from_b_ov_kappa, spin_components_from_b_ov_kappa_f64,
                    try_spin_components_on_device)

__version__ = "0.1.0"
