"""fh_rimp2_energy_dev_noalloc against fh_rimp2_energy_dev (run with -m gpu).

The no-allocation entry runs the same kernels on a workspace carved from a
fixed device array, so its energies and per-row partials equal the pool entry's
bit for bit: checked on two calls in a row, the second on a smaller input (the
array keeps the first call's contents), in FP64 and in each reduced precision,
and a workspace beyond the array is FH_POOL_FULL without touching the device.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from ferric_amd import Precision, _lib, synthetic  # noqa: E402

SCRATCH_BYTES = 128 * 1024 * 1024


def _call(name, t, eps, nocc, nvir, prec, kappa=float("nan")):
    L = _lib.lib()
    out = _lib.SpinComponentsC()
    rows = np.zeros(2 * nocc)
    e = np.ascontiguousarray(eps, dtype=np.float64)
    dt = _lib.DType.F64 if t.dtype == torch.float64 else _lib.DType.F32
    stream = torch.cuda.current_stream().cuda_stream
    st = getattr(L, name)(ctypes.c_void_p(t.data_ptr()), int(dt), t.shape[0], nocc, nvir,
                          e.ctypes.data_as(_lib._c_dp), e.size, 0, nocc, kappa, int(prec),
                          ctypes.c_void_p(stream), ctypes.byref(out),
                          rows.ctypes.data_as(_lib._c_dp))
    return st, (out.e_os, out.e_ss, out.e_total), rows


def _both(p, prec, kappa=float("nan")):
    t = torch.from_numpy(np.ascontiguousarray(p.b_ov)).cuda()
    print("no-alloc", p.name, prec.name, flush=True)
    st_a, e_a, r_a = _call("fh_rimp2_energy_dev_noalloc", t, p.eps, p.nocc, p.nvir, prec, kappa)
    assert st_a == 0, _lib.last_error()
    st_p, e_p, r_p = _call("fh_rimp2_energy_dev", t, p.eps, p.nocc, p.nvir, prec, kappa)
    assert st_p == 0, _lib.last_error()
    assert e_a == e_p, (e_a, e_p)
    assert np.array_equal(r_a, r_p)
    return e_a


@pytest.mark.parametrize("prec", [Precision.FP64, Precision.MIXED, Precision.BF16X3,
                                  Precision.FP8_E4M3, Precision.INT8X4])
def test_noalloc_equals_pool_entry_large_then_small(prec):
    big = synthetic.decane_scale()
    small = synthetic.water_shapes()
    _both(big, prec)
    _both(small, prec)
    _both(small, prec, kappa=1.3)


def test_noalloc_refuses_a_workspace_beyond_the_array():
    # packed FP64 operand alone: 20 * 1008 * 1024 * 8 B = 165 MB > 128 MiB
    nocc, nvir, naux = 20, 1000, 1024
    t = torch.zeros((naux, nocc * nvir), dtype=torch.float32, device="cuda")
    eps = np.concatenate([np.linspace(-1.0, -0.5, nocc), np.linspace(0.5, 2.0, nvir)])
    before = _lib.stats()["fail_pool_full"]
    st, _, _ = _call("fh_rimp2_energy_dev_noalloc", t, eps, nocc, nvir, Precision.FP64)
    assert st == _lib.Status.POOL_FULL
    assert "fixed" in _lib.last_error() and str(SCRATCH_BYTES) in _lib.last_error()
    assert _lib.stats()["fail_pool_full"] == before + 1
    # the pool entry still runs the same shape, and the no-alloc entry the next
    # small one
    st, _, _ = _call("fh_rimp2_energy_dev", t, eps, nocc, nvir, Precision.FP64)
    assert st == 0
    _both(synthetic.water_shapes(), Precision.FP64)
