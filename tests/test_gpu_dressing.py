"""The dressing / MO transform on the GPU (rocBLAS GEMMs), run with -m gpu.

Against the oracle's restatement of stream_dressed_mo_band (rimp2.rs:648-756)
at 1e-12 relative, over output bands, and end to end on real water: (Q|mu nu),
MO coefficients and V^-1/2 -> B on the device -> the pair kernel reproduces
the ORCA / PySCF RI-MP2 energy of the distorted water to 1e-10 Eh.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import ferric_amd  # noqa: E402
from ferric_amd import rimp2  # noqa: E402
from ferric_amd.dressing import stream_dressed_mo_band  # noqa: E402

FIX = os.path.join(os.path.dirname(__file__), "golden", "water_ccpvdz_real.npz")


def _inv_sqrt(v):
    w, u = np.linalg.eigh(v)
    return (u / np.sqrt(w)) @ u.T


def _cuda(x):
    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


def test_against_oracle_and_bands(oracle_mod):
    rng = np.random.default_rng(8)
    naux, nao, nl, nr = 333, 37, 6, 29
    a = rng.standard_normal((naux, nao, nao))
    a = a + a.transpose(0, 2, 1)
    m = rng.standard_normal((naux, naux)) / naux
    cl, cr = rng.standard_normal((nao, nl)), rng.standard_normal((nao, nr))
    print("dressing", naux, nao, nl, nr, flush=True)
    want = oracle_mod.dress_b_ov(a, m, cl, cr)
    got = stream_dressed_mo_band(_cuda(a), _cuda(m), _cuda(cl), _cuda(cr)).cpu().numpy()
    assert np.max(np.abs(got - want)) <= 1e-12 * np.max(np.abs(want))
    band = stream_dressed_mo_band(_cuda(a), _cuda(m), _cuda(cl), _cuda(cr),
                                  output_band=(100, 290)).cpu().numpy()
    assert np.max(np.abs(band - want[100:290])) <= 1e-12 * np.max(np.abs(want))


def test_real_water_end_to_end():
    w = np.load(FIX)
    c = w["dist_c"]
    b = stream_dressed_mo_band(_cuda(w["dist_j3"]), _cuda(_inv_sqrt(w["dist_j2"])),
                               _cuda(c[:, :5]), _cuda(c[:, 5:]))
    sc = rimp2.spin_components(b, w["dist_eps"], 5, 19)
    print(f"real water via device dressing: {sc.e_total:.13f}", flush=True)
    assert abs(sc.e_os - float(w["dist_e_os_ref"])) <= 1e-10
    assert abs(sc.e_ss - float(w["dist_e_ss_ref"])) <= 1e-10
    assert abs(sc.e_total - float(w["dist_e_corr_orca"])) <= 1e-10


def test_errors():
    a = torch.zeros((4, 3, 3), dtype=torch.float64, device="cuda")
    v = torch.zeros((4, 4), dtype=torch.float64, device="cuda")
    c = torch.zeros((3, 2), dtype=torch.float64, device="cuda")
    with pytest.raises(ferric_amd.LayoutError):
        stream_dressed_mo_band(a, v, c, c, output_band=(2, 6))
    with pytest.raises(ferric_amd.LayoutError):
        stream_dressed_mo_band(a.float(), v, c, c)
