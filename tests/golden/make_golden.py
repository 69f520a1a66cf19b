"""Writes tests/golden/rimp2_golden.npz: inputs and brute-force expected outputs.

The expected values come from this script's own numpy brute force over the full
(i, j) double loop (no unique-pair symmetry, no oracle code), the formulation of
the reference's naive-reference tests (crates/ferric-mp2/src/rimp2.rs:2856-2927
and :3106-3185), summed with math.fsum. The adjoint values come from the closed
forms Gamma = 4 x_ov, dE/deps = 2 diag(P) (rimp2.rs:1832-1840, :1987-2052;
zvector.rs:651-697) and are cross-checked here against central finite
differences of the brute-force energy before anything is written.

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from tests.golden.generators import (LCG_CASES, RESIDENT_CASES, lcg_case,  # noqa: E402
                                     resident_case)


def brute_energy(b, eps, nocc, nvir, first_occ, nocc_total, kappa=None):
    """(E_os, E_ss) over the full ordered (i, j) double loop."""
    eo = eps[first_occ:first_occ + nocc]
    ev = eps[nocc_total:nocc_total + nvir]
    k = (b.T @ b).reshape(nocc, nvir, nocc, nvir)
    gab = k.transpose(0, 2, 1, 3)          # [i,j,a,b] = (ia|jb)
    gba = k.transpose(0, 2, 3, 1)          # [i,j,a,b] = (ib|ja)
    d = (eo[:, None, None, None] + eo[None, :, None, None]
         - ev[None, None, :, None] - ev[None, None, None, :])
    damp = 1.0 if kappa is None else (1.0 - np.exp(kappa * d)) ** 2
    os_terms = damp * gab * gab / d
    ss_terms = damp * gab * (gab - gba) / d
    return math.fsum(os_terms.ravel()), math.fsum(ss_terms.ravel())


def closed_form_adjoint(b, eps, nocc, nvir, first_occ, nocc_total, c_os, c_ss):
    """Gamma = dE/dB and dE/deps of E = c_os E_os + c_ss E_ss."""
    eo = eps[first_occ:first_occ + nocc]
    ev = eps[nocc_total:nocc_total + nvir]
    k = (b.T @ b).reshape(nocc, nvir, nocc, nvir).transpose(0, 2, 1, 3)  # [i,j,a,b]
    d = (eo[:, None, None, None] + eo[None, :, None, None]
         - ev[None, None, :, None] - ev[None, None, None, :])
    t = k / d
    tt = (c_os + c_ss) * t - c_ss * t.transpose(0, 1, 3, 2)
    bq = b.reshape(-1, nocc, nvir)
    x = np.einsum("ijab,pjb->pia", tt, bq).reshape(b.shape[0], nocc * nvir)
    gamma = 4.0 * x
    deps = np.zeros_like(eps)
    w = t * tt
    deps[first_occ:first_occ + nocc] = -2.0 * w.sum(axis=(1, 2, 3))
    deps[nocc_total:nocc_total + nvir] = 2.0 * w.sum(axis=(0, 1, 3))
    return gamma, deps


def _fd_check(b, eps, nocc, nvir, first_occ, nocc_total, c_os, c_ss, gamma, deps, rng):
    h = 1e-5

    def e_of(bb, ee):
        os_, ss_ = brute_energy(bb, ee, nocc, nvir, first_occ, nocc_total)
        return c_os * os_ + c_ss * ss_

    for _ in range(16):
        p = rng.integers(b.shape[0])
        q = rng.integers(b.shape[1])
        bp = b.copy()
        bm = b.copy()
        bp[p, q] += h
        bm[p, q] -= h
        fd = (e_of(bp, eps) - e_of(bm, eps)) / (2 * h)
        assert abs(fd - gamma[p, q]) <= 1e-8 * max(1.0, abs(gamma[p, q])), (fd, gamma[p, q])
    for q in range(eps.size):
        ep = eps.copy()
        em = eps.copy()
        ep[q] += h
        em[q] -= h
        fd = (e_of(b, ep) - e_of(b, em)) / (2 * h)
        assert abs(fd - deps[q]) <= 1e-8 * max(1.0, abs(deps[q])), (q, fd, deps[q])


def water_like():
    from ferric_amd.synthetic import water_shapes

    p = water_shapes()
    return p.b_ov, p.eps, p.nocc, p.nvir, p.first_occ, p.nocc_total


def main():
    out = {}
    cases = [(f"lcg_{na}_{no}_{nv}_{fc}", lcg_case(na, no, nv, fc)) for na, no, nv, fc in LCG_CASES]
    cases += [(f"res_{na}_{no}_{nv}_{fc}_{sd}", resident_case(na, no, nv, fc, sd))
              for na, no, nv, fc, sd in RESIDENT_CASES]
    cases.append(("water_shapes", water_like()))
    rng = np.random.default_rng(7)
    names = []
    for name, (b, eps, nocc, nvir, fc, nt) in cases:
        names.append(name)
        out[f"{name}/b"] = b
        out[f"{name}/eps"] = eps
        out[f"{name}/dims"] = np.array([nocc, nvir, fc, nt, b.shape[0]])
        out[f"{name}/energy"] = np.array(brute_energy(b, eps, nocc, nvir, fc, nt))
        out[f"{name}/energy_kappa1.3"] = np.array(brute_energy(b, eps, nocc, nvir, fc, nt, 1.3))
        if nocc * nvir * b.shape[0] <= 12000:
            for tag, (c_os, c_ss) in (("mp2", (1.0, 1.0)), ("scs", (1.2, 1.0 / 3.0))):
                gamma, deps = closed_form_adjoint(b, eps, nocc, nvir, fc, nt, c_os, c_ss)
                _fd_check(b, eps, nocc, nvir, fc, nt, c_os, c_ss, gamma, deps, rng)
                out[f"{name}/gamma_{tag}"] = gamma
                out[f"{name}/deps_{tag}"] = deps
    out["names"] = np.array(names)
    path = os.path.join(HERE, "rimp2_golden.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, len(names), "cases")


if __name__ == "__main__":
    main()
