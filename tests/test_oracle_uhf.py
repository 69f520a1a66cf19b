"""Pins of the oracle's unrestricted blocks and signed factor (CPU only).

* closed-shell reduction (the reference's own u_rimp2_matches_closed_shell
  check, u_rimp2.rs:1732): with alpha == beta, E_aa = E_bb = E_ss / 2 and
  E_ab = E_os of the restricted kernel;
* an independent brute force in the spin-orbital basis: E_UMP2 =
  1/4 sum_ijab |<ij||ab>|^2 / D over all spin orbitals, with alpha and beta
  channels of different sizes and frozen cores, must equal E_aa + E_bb + E_ab;
* the signed factor: rows that appear in both pos and neg cancel exactly, and
  the result equals the dense-G path (rimp2.rs:984-1018) on K = P^T P - N^T N;
* teeth: a restatement with a dropped factor (fac = 1) or a flipped exchange
  sign fails the brute force.
"""
import numpy as np
import pytest

import oracle


def _channel(rng, naux, nocc, nvir, fc=0):
    b = rng.standard_normal((naux, nocc * nvir)) * np.sqrt(0.09 / naux)
    occ = np.sort(rng.uniform(-1.4, -0.4, fc + nocc))
    vir = np.sort(np.exp(rng.uniform(np.log(0.15), np.log(4.0), nvir)))
    return b, np.concatenate([occ, vir]), nocc, nvir, fc, fc + nocc


def _spin_orbital_energy(ch_a, ch_b):
    """1/4 sum |<ij||ab>|^2 / D over all spin orbitals, from the RI factors."""
    ba, ea, noa, nva, foa, nta = ch_a
    bb, eb, nob, nvb, fob, ntb = ch_b
    naux = ba.shape[0]
    # spin-orbital lists: (spin, index)
    occ = [(0, i) for i in range(noa)] + [(1, i) for i in range(nob)]
    vir = [(0, a) for a in range(nva)] + [(1, a) for a in range(nvb)]
    B = {0: ba.reshape(naux, noa, nva), 1: bb.reshape(naux, nob, nvb)}
    eo = {0: ea[foa:foa + noa], 1: eb[fob:fob + nob]}
    ev = {0: ea[nta:nta + nva], 1: eb[ntb:ntb + nvb]}

    def coul(si, i, sa, a, sj, j, sb, b):   # (ia|jb) in the spin-orbital basis
        if si != sa or sj != sb:
            return 0.0
        return float(B[si][:, i, a] @ B[sj][:, j, b])

    e = 0.0
    for (si, i) in occ:
        for (sj, j) in occ:
            for (sa, a) in vir:
                for (sb, b) in vir:
                    g = coul(si, i, sa, a, sj, j, sb, b) - coul(si, i, sb, b, sj, j, sa, a)
                    d = eo[si][i] + eo[sj][j] - ev[sa][a] - ev[sb][b]
                    e += g * g / d
    return 0.25 * e


def test_closed_shell_reduction():
    rng = np.random.default_rng(11)
    b, eps, no, nv, fc, nt = _channel(rng, 14, 4, 6, fc=1)
    eos, ess, _ = oracle.spin_components(b, eps, no, nv, fc, nt)
    e_aa = oracle.u_same_spin(b, eps, no, nv, fc, nt)
    e_ab = oracle.u_opposite_spin((b, eps, no, nv, fc, nt), (b, eps, no, nv, fc, nt))
    assert e_aa == pytest.approx(0.5 * ess, rel=1e-13, abs=1e-16)
    assert e_ab == pytest.approx(eos, rel=1e-13)


@pytest.mark.parametrize("seed", [0, 1])
def test_spin_orbital_brute_force(seed):
    rng = np.random.default_rng(100 + seed)
    naux = 11
    ch_a = _channel(rng, naux, 3, 5, fc=1)
    ch_b = _channel(rng, naux, 2, 6, fc=0)
    e_aa = oracle.u_same_spin(*ch_a)
    e_bb = oracle.u_same_spin(*ch_b)
    e_ab = oracle.u_opposite_spin(ch_a, ch_b)
    want = _spin_orbital_energy(ch_a, ch_b)
    assert e_aa < 0 and e_bb < 0 and e_ab < 0
    assert e_aa + e_bb + e_ab == pytest.approx(want, rel=1e-12)
    # the two orders of the opposite-spin pair give the same energy
    assert oracle.u_opposite_spin(ch_b, ch_a) == pytest.approx(e_ab, rel=1e-12)


def _numpy_same_spin(b, eps, no, nv, fc, nt, fac_off=2.0, sign=-1.0):
    B = b.reshape(b.shape[0], no, nv)
    eo, ev = eps[fc:fc + no], eps[nt:nt + nv]
    e = 0.0
    for i in range(no):
        for j in range(i, no):
            g = B[:, i, :].T @ B[:, j, :]
            k = g + sign * g.T
            d = eo[i] + eo[j] - ev[:, None] - ev[None, :]
            e += (1.0 if i == j else fac_off) * np.sum(k * k / d)
    return 0.25 * e


def test_same_spin_teeth():
    rng = np.random.default_rng(5)
    ch = _channel(rng, 12, 4, 5, fc=1)
    got = oracle.u_same_spin(*ch)
    assert got == pytest.approx(_numpy_same_spin(*ch), rel=1e-12)
    assert abs(_numpy_same_spin(*ch, fac_off=1.0) - got) > 1e-3 * abs(got)
    assert abs(_numpy_same_spin(*ch, sign=+1.0) - got) > 1e-3 * abs(got)


def test_signed_factor():
    rng = np.random.default_rng(7)
    no, nv = 3, 5
    b, eps, *_ = _channel(rng, 10, no, nv)
    y = rng.standard_normal((4, no * nv)) * 0.05
    plain = oracle.spin_components(b, eps, no, nv)
    # rows present in both factors cancel
    stacked = oracle.signed_spin_components(np.vstack([b, y]), y, eps, no, nv)
    assert stacked == pytest.approx(plain, rel=1e-12)
    # the dense-G path on K = P^T P - N^T N
    neg = rng.standard_normal((6, no * nv)) * 0.02
    g = b.T @ b - neg.T @ neg
    want = oracle.spin_components_from_g(g, eps, no, nv)
    got = oracle.signed_spin_components(b, neg, eps, no, nv)
    assert got == pytest.approx(want, rel=1e-12)
    # a wrong sign of the negative factor is caught
    wrong = oracle.spin_components_from_g(b.T @ b + neg.T @ neg, eps, no, nv)
    assert abs(wrong[2] - got[2]) > 1e-6 * abs(got[2])
