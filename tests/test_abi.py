"""The C ABI library loads and exports every symbol include/ferric_hip.h declares.

These run without a GPU: only symbol lookup and the argument checks that the
library performs before touching a device (the reference's shape checks come
first, crates/ferric-mp2/src/rimp2_gpu.rs:211-237).
"""
import ctypes
import math
import os
import re

import numpy as np
import pytest

from ferric_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "ferric_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fh_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_binding_table():
    assert declared_functions() == sorted(_lib.SIGNATURES)


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    for name in declared_functions():
        assert hasattr(L, name), name
    assert L.fh_abi_version() == 11


def test_enum_values_match_header():
    text = open(HEADER).read()
    for st in _lib.Status:
        assert re.search(rf"FH_{st.name}\s*=\s*{int(st)}\b", text), st
    for p in _lib.Precision:
        assert re.search(rf"FH_{p.name}\s*=\s*{int(p)}\b", text), p


def test_struct_layouts():
    assert ctypes.sizeof(_lib.SpinComponentsC) == 24
    assert ctypes.sizeof(_lib.StatsC) == 19 * 8
    assert ctypes.sizeof(_lib.DomainC) == 32
    assert ctypes.sizeof(_lib.TimingC) == 48
    assert ctypes.sizeof(_lib.PnoDiagnosticsC) == 16


def test_pno_entry_checks_before_the_device():
    """t_cut_pno < 0 (local_pno.rs:112-118) and malformed pairs are FH_LAYOUT
    before any device work, so they are testable without a GPU."""
    b = np.zeros((4, 6))
    eps = np.array([-1.0, -0.8, 0.5, 0.6, 0.7])
    L = _lib.lib()

    def call(pairs, t):
        pl = np.ascontiguousarray(np.asarray(pairs, dtype=np.int32).reshape(-1, 2))
        os_, ss = np.zeros(len(pl)), np.zeros(len(pl))
        return L.fh_dlpno_pno_pair_energies(
            b.ctypes.data_as(_lib._c_dp), 4, 2, 3, eps.ctypes.data_as(_lib._c_dp), 5, 0, 2,
            ctypes.c_void_p(pl.ctypes.data), len(pl), t, os_.ctypes.data_as(_lib._c_dp),
            ss.ctypes.data_as(_lib._c_dp), None)

    assert call([(0, 1)], -1e-3) == _lib.Status.LAYOUT
    assert call([(1, 0)], 0.0) == _lib.Status.LAYOUT
    assert call([(0, 2)], 0.0) == _lib.Status.LAYOUT
    assert call([(-1, 1)], 0.0) == _lib.Status.LAYOUT


def test_batched_descriptor_checks_before_the_device():
    """Malformed domain descriptors are FH_LAYOUT before any device work: an
    empty dimension, a block past the buffer, eps past its buffer, and a
    domain beyond the kernel's range (nocc nvir >= 2^22)."""
    from ferric_amd.batched import descriptors

    L = _lib.lib()
    eo, es = np.zeros(1), np.zeros(1)
    fake = ctypes.c_void_p(0x1000)   # never dereferenced: the checks come first

    def call(nocc, nvir, naux, b_len, eps_len, b_off=0, e_off=0):
        d = descriptors([nocc], [nvir], [naux], [b_off], [e_off])
        return L.fh_rimp2_batched_dev(d, 1, fake, b_len, fake, eps_len,
                                      int(_lib.Precision.FP64), 0, None,
                                      eo.ctypes.data_as(_lib._c_dp), es.ctypes.data_as(_lib._c_dp))

    assert call(0, 5, 8, 100, 100) == _lib.Status.LAYOUT
    assert call(2, 5, 8, 79, 100) == _lib.Status.LAYOUT
    assert call(2, 5, 8, 80, 6) == _lib.Status.LAYOUT
    assert call(2, 5, 8, 80, 7, e_off=1) == _lib.Status.LAYOUT
    assert call(4096, 1024, 1, 1 << 22, 5120) == _lib.Status.LAYOUT
    assert call(2, 5, 8, 80, 7, b_off=1) == _lib.Status.LAYOUT
    with pytest.raises(_lib.LayoutError):
        descriptors([2], [-5], [8], [0], [0])


def test_from_g_entries_check_before_the_device():
    eps = np.array([-1.0, -0.8, 0.5, 0.6, 0.7])
    g = np.zeros((6, 6))
    L = _lib.lib()
    assert L.fh_rimp2_energy_from_g(g.ctypes.data_as(_lib._c_dp), 2, 3,
                                    eps.ctypes.data_as(_lib._c_dp), 5, 0, 2,
                                    None) == _lib.Status.LAYOUT
    out = _lib.SpinComponentsC()
    assert L.fh_rimp2_energy_from_g(g.ctypes.data_as(_lib._c_dp), 2, 3,
                                    eps.ctypes.data_as(_lib._c_dp), 4, 0, 2,
                                    ctypes.byref(out)) == _lib.Status.LAYOUT  # eps too short
    os_, ss = np.zeros(1), np.zeros(1)
    for pair in ((1, 0), (0, 2)):
        pl = np.array(pair, dtype=np.int32)
        assert L.fh_pair_energies_from_blocks_dev(
            ctypes.c_void_p(0), ctypes.c_void_p(pl.ctypes.data), 1, 2, 3,
            eps.ctypes.data_as(_lib._c_dp), 5, 0, 2, None, os_.ctypes.data_as(_lib._c_dp),
            ss.ctypes.data_as(_lib._c_dp)) == _lib.Status.LAYOUT


def test_domain_fit_entry_checks_before_the_device():
    a, v = np.zeros((4, 6)), np.eye(4)
    xyz, cen = np.zeros((4, 3)), np.zeros((2, 3))
    L = _lib.lib()

    def call(pairs, radius):
        pl = np.ascontiguousarray(np.asarray(pairs, dtype=np.int32).reshape(-1, 2))
        out = np.zeros((len(pl), 3, 3))
        return L.fh_domain_fit_pair_blocks(
            a.ctypes.data_as(_lib._c_dp), v.ctypes.data_as(_lib._c_dp), 4, 2, 3,
            xyz.ctypes.data_as(_lib._c_dp), cen.ctypes.data_as(_lib._c_dp), radius,
            ctypes.c_void_p(pl.ctypes.data), len(pl), out.ctypes.data_as(_lib._c_dp), None)

    assert call([(0, 1)], -1.0) == _lib.Status.LAYOUT
    assert call([(0, 1)], float("nan")) == _lib.Status.LAYOUT
    assert call([(1, 0)], 1.0) == _lib.Status.LAYOUT
    assert call([(0, 2)], 1.0) == _lib.Status.LAYOUT


def _energy(b, naux, nocc, nvir, eps, first_occ, nocc_total, kappa=float("nan"), prec=0):
    out = _lib.SpinComponentsC()
    bp = b.ctypes.data_as(_lib._c_dp) if b is not None else None
    ep = eps.ctypes.data_as(_lib._c_dp)
    st = _lib.lib().fh_rimp2_energy(bp, naux, nocc, nvir, ep, eps.size, first_occ, nocc_total,
                                    kappa, prec, ctypes.byref(out))
    return st, out


def test_shape_checks_before_any_device_work():
    L = _lib.lib()
    eps = np.zeros(3)
    # eps too short for the index ranges -> Layout
    st, _ = _energy(np.zeros((4, 6)), 4, 2, 3, eps, 0, 2)
    assert st == _lib.Status.LAYOUT
    assert "orbital energies" in _lib.last_error()
    # nocc * nvir overflows -> Layout
    st, _ = _energy(np.zeros((1, 1)), 1, (1 << 64) - 1, 2, np.zeros(4), 0, 1)
    assert st == _lib.Status.LAYOUT
    # empty channel -> zeros, OK
    st, out = _energy(np.zeros((4, 0)), 4, 0, 3, eps, 0, 0)
    assert st == _lib.Status.OK and (out.e_os, out.e_ss, out.e_total) == (0.0, 0.0, 0.0)
    st, out = _energy(np.zeros((4, 0)), 4, 3, 0, eps, 0, 3)
    assert st == _lib.Status.OK
    # naux == 0 -> Layout (the CPU path owns this degenerate shape)
    st, _ = _energy(np.zeros((0, 6)), 0, 2, 3, np.zeros(5), 0, 2)
    assert st == _lib.Status.LAYOUT
    # kappa must be finite and > 0
    for k in (0.0, -1.0, math.inf):
        st, _ = _energy(np.zeros((4, 6)), 4, 2, 3, np.zeros(5), 0, 2, kappa=k)
        assert st == _lib.Status.LAYOUT, k
    # unknown precision
    st, _ = _energy(np.zeros((4, 6)), 4, 2, 3, np.zeros(5), 0, 2, prec=9)
    assert st == _lib.Status.LAYOUT
    s = _lib.stats()
    assert s["fail_layout"] >= 6
    assert L.fh_reset_stats() == 0
    assert _lib.stats()["calls"] == 0


def test_python_errors_are_typed():
    from ferric_amd import rimp2

    with pytest.raises(_lib.LayoutError):
        rimp2.spin_components(np.zeros((4, 6)), np.zeros(3), 2, 3)
    with pytest.raises(_lib.LayoutError):
        rimp2.spin_components(np.zeros((4, 7)), np.zeros(5), 2, 3)
    sc = rimp2.spin_components(np.zeros((4, 0)), np.zeros(3), 0, 3)
    assert sc.e_total == 0.0


def test_try_on_device_returns_none_on_failure(monkeypatch):
    from ferric_amd import rimp2

    assert rimp2.try_spin_components_on_device(np.zeros((4, 6)), np.zeros(3), 2, 3, 0, 2) is None
    monkeypatch.setenv("FERRIC_GPU", "off")
    assert rimp2.try_spin_components_on_device(np.zeros((4, 0)), np.zeros(3), 0, 3, 0, 0) is None


def test_spin_channel_layout():
    # fh_spin_channel: pointer, 2 x size_t, pointer, 3 x size_t
    assert ctypes.sizeof(_lib.SpinChannelC) == 7 * 8


def test_unrestricted_argument_checks_before_device():
    from ferric_amd import unrestricted as u

    eps = np.zeros(10)
    ch = _lib.SpinChannelC(None, 2, 3, eps.ctypes.data_as(_lib._c_dp), eps.size, 0, 2)
    out = ctypes.c_double()
    L = _lib.lib()
    # unsupported precision for the opposite-spin rectangle
    st = L.fh_u_opposite_spin(ctypes.byref(ch), ctypes.byref(ch), 4, 2, ctypes.byref(out))
    assert st == _lib.Status.UNSUPPORTED
    # eps too short (needs nocc_total + nvir = 5 here it is 10, so shrink)
    short = _lib.SpinChannelC(None, 2, 3, eps.ctypes.data_as(_lib._c_dp), 4, 0, 2)
    st = L.fh_u_same_spin(ctypes.byref(short), 4, 0, ctypes.byref(out))
    assert st == _lib.Status.LAYOUT
    # empty channel: zero energy without a device
    empty = _lib.SpinChannelC(None, 0, 3, eps.ctypes.data_as(_lib._c_dp), eps.size, 0, 0)
    st = L.fh_u_same_spin(ctypes.byref(empty), 4, 0, ctypes.byref(out))
    assert st == _lib.Status.OK and out.value == 0.0
    # the Python layer refuses channels fitted in different aux bases
    with pytest.raises(_lib.LayoutError):
        u.u_opposite_spin(u.SpinChannel(np.zeros((4, 6)), eps, 2, 3),
                          u.SpinChannel(np.zeros((5, 6)), eps, 2, 3))
    # signed: FP8 is refused before any device work
    sc = _lib.SpinComponentsC()
    b = np.zeros((4, 6))
    st = L.fh_rimp2_energy_signed(b.ctypes.data_as(_lib._c_dp), 4, b.ctypes.data_as(_lib._c_dp),
                                  4, 2, 3, eps.ctypes.data_as(_lib._c_dp), eps.size, 0, 2, 3,
                                  ctypes.byref(sc))
    assert st == _lib.Status.UNSUPPORTED
