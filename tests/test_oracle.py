"""Pins the CPU oracle (oracle/rimp2_oracle.c) before anything is compared with it.

Each check is chosen so that a plausible restatement mistake fails at least one:
the reference's own test shapes and generators against an independent full
(i, j) brute force, the spin-orbital identity, the reference's mutation
(testdata/mutations/ri-mp2.patch), the kappa limits, thread-count determinism,
and for the adjoint: the scalar definition of x_ov (zvector.rs:1135-1175),
central finite differences, Euler homogeneity and the trace identity.
"""
import math

import numpy as np
import pytest

from tests.golden.generators import (LCG_CASES, RESIDENT_CASES, lcg_case, resident_case,
                                     xorshift_case)


def naive_full_width(b, eps, nocc, nvir, first_occ, nocc_total):
    """The full-width reference of rimp2.rs:2858-2884 (every j from 0), restated."""
    e_total = 0.0
    for i in range(nocc):
        g_i = b[:, i * nvir:(i + 1) * nvir].T @ b
        for j in range(i, nocc):
            fac = 1.0 if i == j else 2.0
            e_ij = eps[first_occ + i] + eps[first_occ + j]
            gab = g_i[:, j * nvir:(j + 1) * nvir]
            d = e_ij - eps[nocc_total:nocc_total + nvir][:, None] - \
                eps[nocc_total:nocc_total + nvir][None, :]
            e_total += fac * np.sum((gab * gab + gab * (gab - gab.T)) / d)
    return e_total


# -- energy ---------------------------------------------------------------------

@pytest.mark.parametrize("case", LCG_CASES)
def test_tail_gemm_matches_full_width_reference(oracle_mod, case):
    """rimp2.rs:2856-2927: rel < 1e-13 against the full-width formulation."""
    b, eps, nocc, nvir, fc, nt = lcg_case(*case)
    want = naive_full_width(b, eps, nocc, nvir, fc, nt)
    got = oracle_mod.spin_components(b, eps, nocc, nvir, fc, nt)
    assert abs(want - got[2]) / max(abs(want), 1e-30) < 1e-13


def test_oracle_matches_golden_brute_force(oracle_mod, golden):
    """rimp2.rs:3106-3185: unique-pair weighting vs the naive full (i,j) loop, rel < 1e-12."""
    for name, c in golden.items():
        got = oracle_mod.spin_components(c["b"], c["eps"], c["nocc"], c["nvir"], c["first_occ"],
                                         c["nocc_total"])
        os_, ss_ = c["energy"]
        for g, w in ((got[0], os_), (got[1], ss_), (got[2], os_ + ss_)):
            assert abs(g - w) <= 1e-12 * max(abs(w), abs(os_)), (name, g, w)
        gk = oracle_mod.spin_components(c["b"], c["eps"], c["nocc"], c["nvir"], c["first_occ"],
                                        c["nocc_total"], kappa=1.3)
        kos, kss = c["energy_kappa"]
        assert abs(gk[0] - kos) <= 1e-12 * abs(kos), name
        assert abs(gk[1] - kss) <= 1e-12 * abs(kos), name


def test_golden_inputs_regenerate_bit_for_bit(golden):
    for na, no, nv, fc in LCG_CASES:
        b, eps, *_ = lcg_case(na, no, nv, fc)
        c = golden[f"lcg_{na}_{no}_{nv}_{fc}"]
        assert np.array_equal(b, c["b"]) and np.array_equal(eps, c["eps"])
    for na, no, nv, fc, sd in RESIDENT_CASES:
        b, eps, *_ = resident_case(na, no, nv, fc, sd)
        c = golden[f"res_{na}_{no}_{nv}_{fc}_{sd}"]
        assert np.array_equal(b, c["b"]) and np.array_equal(eps, c["eps"])


def spin_orbital_energy(b, eo, ev):
    """scripts/mp2_variant_identities.py claim 1, restated: 1/4 sum |<ij||ab>|^2/D
    in the spin-orbital basis, split by spin of the occupied pair."""
    no, nv = eo.size, ev.size
    n = no + nv
    # B over all orbitals with zero occ-occ / vir-vir blocks is enough: only
    # (ia|jb) enters <ij||ab> = (ia|jb) - (ib|ja).
    k = (b.T @ b).reshape(no, nv, no, nv)
    spin = np.array([s for p in range(n) for s in (0, 1)])
    spat = np.array([p for p in range(n) for s in (0, 1)])
    occ = [q for q in range(2 * n) if spat[q] < no]
    vir = [q for q in range(2 * n) if spat[q] >= no]
    e_all = np.concatenate([eo, ev])
    so_os = so_ss = 0.0
    for I in occ:
        for J in occ:
            for A in vir:
                for B in vir:
                    i, j, a, bb = spat[I], spat[J], spat[A] - no, spat[B] - no
                    direct = k[i, a, j, bb] if (spin[I] == spin[A] and spin[J] == spin[B]) else 0.0
                    exch = k[i, bb, j, a] if (spin[I] == spin[B] and spin[J] == spin[A]) else 0.0
                    d = e_all[i] + e_all[j] - e_all[no + a] - e_all[no + bb]
                    t = 0.25 * (direct - exch) ** 2 / d
                    if spin[I] == spin[J]:
                        so_ss += t
                    else:
                        so_os += t
    return so_os, so_ss


def test_spin_orbital_identity(oracle_mod):
    rng = np.random.default_rng(20261105)
    no, nv, naux = 3, 4, 9
    b = rng.standard_normal((naux, no * nv)) * 0.3
    eo = np.sort(rng.uniform(-1.5, -0.5, no))
    ev = np.sort(rng.uniform(0.3, 2.0, nv))
    eps = np.concatenate([eo, ev])
    got = oracle_mod.spin_components(b, eps, no, nv)
    so_os, so_ss = spin_orbital_energy(b, eo, ev)
    assert got[0] == pytest.approx(so_os, rel=1e-12)
    assert got[1] == pytest.approx(so_ss, rel=1e-12)


def _pair_loop_numpy(b, eps, nocc, nvir, first_occ, nocc_total, mutation=None):
    """The rimp2.rs:1399-1544 loop in numpy, optionally with a planted mistake
    (the test of the tests: a mutation must fail the brute-force checks)."""
    os_ = ss_ = 0.0
    ev = eps[nocc_total:nocc_total + nvir]
    for i in range(nocc):
        g_i = b[:, i * nvir:(i + 1) * nvir].T @ b[:, i * nvir:]
        for j in range(i, nocc):
            fac = 1.0 if i == j else (1.0 if mutation == "fac1" else 2.0)
            col = (j - i) * nvir if mutation != "offset" else (j - i) * nvir // 2
            gab = g_i[:, col:col + nvir]
            gba = gab.T
            d = eps[first_occ + i] + eps[first_occ + j] - ev[:, None] - ev[None, :]
            os_ += fac * np.sum(gab * gab / d)
            if mutation == "ss_sign":   # testdata/mutations/ri-mp2.patch
                ss_ += fac * np.sum(gab * (gab + gba) / d)
            else:
                ss_ += fac * np.sum(gab * (gab - gba) / d)
    return os_, ss_


@pytest.mark.parametrize("mutation", ["ss_sign", "fac1", "offset"])
def test_mutations_are_caught(golden, mutation):
    c = golden["res_150_9_31_2_5"]
    args = (c["b"], c["eps"], c["nocc"], c["nvir"], c["first_occ"], c["nocc_total"])
    clean = _pair_loop_numpy(*args)
    mutated = _pair_loop_numpy(*args, mutation=mutation)
    want = c["energy"]
    assert abs(clean[0] - want[0]) < 1e-12 and abs(clean[1] - want[1]) < 1e-12
    worst = max(abs(mutated[0] - want[0]), abs(mutated[1] - want[1]))
    assert worst > 1e-6, f"mutation {mutation} not caught"


def test_every_pair_energy_is_nonpositive(oracle_mod, golden):
    """Claim 1b of the reference's identity script: e_ij <= 0 for every pair."""
    for name, c in golden.items():
        nocc, nvir = c["nocc"], c["nvir"]
        for i in range(nocc):
            g = oracle_mod.tail_block(c["b"], nocc, nvir, i)
            os_i, ss_i = oracle_mod.i_partial(g, i, nocc, nvir, c["eps"], c["first_occ"],
                                              c["nocc_total"])
            assert os_i + ss_i <= 0.0 and os_i < 0.0, name


def test_kappa_limits(oracle_mod, golden):
    """rimp2.rs:51-57: kappa -> inf recovers MP2, kappa -> 0 switches it off."""
    c = golden["res_120_7_23_1_3"]
    args = (c["b"], c["eps"], c["nocc"], c["nvir"], c["first_occ"], c["nocc_total"])
    plain = oracle_mod.spin_components(*args)
    big = oracle_mod.spin_components(*args, kappa=1e6)
    small = oracle_mod.spin_components(*args, kappa=1e-7)
    assert big[2] == pytest.approx(plain[2], rel=1e-14)
    assert abs(small[2]) < 1e-12 * abs(plain[2])
    mid = oracle_mod.spin_components(*args, kappa=1.0)
    assert plain[2] < mid[2] < 0.0


def test_thread_count_determinism(oracle_mod, golden):
    c = golden["res_150_9_31_2_5"]
    args = (c["b"], c["eps"], c["nocc"], c["nvir"], c["first_occ"], c["nocc_total"])
    ref = oracle_mod.spin_components(*args, nthreads=1)
    for t in (2, 5, 8):
        assert oracle_mod.spin_components(*args, nthreads=t) == ref


def test_blas_driver_matches_c_oracle(oracle_mod, golden):
    """The OpenBLAS-backed baseline driver computes the same energy."""
    c = golden["res_200_6_20_0_21"]
    args = (c["b"], c["eps"], c["nocc"], c["nvir"], c["first_occ"], c["nocc_total"])
    ref = oracle_mod.spin_components(*args)
    got = oracle_mod.spin_components_blas(*args, nthreads=3)
    assert got[0] == pytest.approx(ref[0], rel=1e-13)
    assert got[1] == pytest.approx(ref[1], rel=1e-13)


def test_dense_g_form_matches(oracle_mod, golden):
    """rimp2.rs:984-1018 against the B form."""
    c = golden["lcg_120_8_40_2"]
    b = c["b"]
    g = b.T @ b
    got = oracle_mod.spin_components_from_g(g, c["eps"], c["nocc"], c["nvir"], c["first_occ"],
                                            c["nocc_total"])
    want = c["energy"]
    assert got[0] == pytest.approx(want[0], rel=1e-12)
    assert got[1] == pytest.approx(want[1], rel=1e-12)


def test_active_occ(oracle_mod):
    assert oracle_mod.active_occ(5, 0) == 5
    assert oracle_mod.active_occ(5, 2) == 3
    assert oracle_mod.active_occ(0, 0) == 0
    with pytest.raises(oracle_mod.OracleError):
        oracle_mod.active_occ(5, 5)


def test_empty_shapes(oracle_mod):
    assert oracle_mod.spin_components(np.zeros((4, 0)), np.zeros(3), 0, 3) == (0.0, 0.0, 0.0)
    assert oracle_mod.spin_components(np.zeros((4, 0)), np.zeros(3), 3, 0) == (0.0, 0.0, 0.0)
    with pytest.raises(oracle_mod.OracleError):
        oracle_mod.spin_components(np.zeros((4, 6)), np.zeros(3), 2, 3)


def test_energy_bound_is_positive_and_linear(oracle_mod, golden):
    c = golden["res_120_7_23_1_3"]
    args = (c["b"], c["eps"], c["nocc"], c["nvir"], c["first_occ"], c["nocc_total"])
    b1 = oracle_mod.energy_bound(*args, eps_g=1e-8)
    b2 = oracle_mod.energy_bound(*args, eps_g=2e-8)
    assert b1[0] > 0 and b1[1] > 0
    assert b2[0] == pytest.approx(2 * b1[0], rel=1e-3)


def _sign_aligned_energy_shift(oracle_mod, c, eps_g):
    """E(G + delta) - E(G) for the per-element perturbation of the bound's
    model with every element pushed away from zero: delta = sign(g) eps_g S."""
    b, eps = c["b"], c["eps"]
    nocc, nvir, fc, nt = c["nocc"], c["nvir"], c["first_occ"], c["nocc_total"]
    g = b.T @ b
    s = np.abs(b).T @ np.abs(b)
    e0 = oracle_mod.spin_components_from_g(g, eps, nocc, nvir, fc, nt)
    e1 = oracle_mod.spin_components_from_g(g + np.sign(g) * eps_g * s, eps, nocc, nvir, fc, nt)
    return e1[0] - e0[0], e1[1] - e0[1]


@pytest.mark.parametrize("name", ["res_120_7_23_1_3", "res_150_9_31_2_5", "water_shapes"])
def test_energy_bound_is_attained_not_loose(oracle_mod, golden, name):
    """The bound of tests/common/rimp2_error_bound.rs (restated) is a worst case
    over per-element G perturbations |delta_ab| <= eps_g S_ab. Pushing every
    element away from zero by exactly eps_g S_ab raises every g_ab^2, so the
    opposite-spin shift equals the opposite-spin bound up to its summation term
    (a bound loose by a factor, a dropped f_ij = 2 or a wrong S fails); the
    same-spin shift stays inside its bound and reaches a sizeable part of it."""
    c = golden[name]
    args = (c["b"], c["eps"], c["nocc"], c["nvir"], c["first_occ"], c["nocc_total"])
    eps_g = 1e-6
    bos, bss = oracle_mod.energy_bound(*args, eps_g=eps_g)
    dos, dss = _sign_aligned_energy_shift(oracle_mod, c, eps_g)
    assert dos < 0
    assert 0.999 * bos <= abs(dos) <= bos, (dos, bos)
    assert 0.05 * bss <= abs(dss) <= bss, (dss, bss)


def test_energy_bound_rows_sum_to_the_bound(oracle_mod, golden):
    """The per-row bound and partials (row_partials_blas, used at sizes where
    the whole bound is too slow) add up to the whole-problem bound and energy."""
    c = golden["res_150_9_31_2_5"]
    args = (c["b"], c["eps"], c["nocc"], c["nvir"], c["first_occ"], c["nocc_total"])
    whole = oracle_mod.energy_bound(*args)
    e = oracle_mod.spin_components(*args)
    rows = oracle_mod.row_partials_blas(c["b"], c["eps"], c["nocc"], c["nvir"],
                                        range(c["nocc"]), c["first_occ"], c["nocc_total"])
    assert sum(r[2] for r in rows.values()) == pytest.approx(whole[0], rel=1e-12)
    assert sum(r[3] for r in rows.values()) == pytest.approx(whole[1], rel=1e-12)
    assert sum(r[0] for r in rows.values()) == pytest.approx(e[0], rel=1e-12)
    assert sum(r[1] for r in rows.values()) == pytest.approx(e[1], rel=1e-12)


def test_fp32_accumulation_model_bounds_a_sequential_fp32_dot():
    """The element model the reduced-precision tests use, eps_g = (naux + 16)
    2^-24 (Higham's gamma_n for an FP32 dot product of length naux, with slack
    for the rounding of the FP32 sum into FP64), bounds an FP32 dot product
    accumulated in the worst association (one long sequential chain), including
    on the hot-row (x20) distribution; the realised error is within three
    orders of the bound (random rounding grows like sqrt(naux))."""
    rng = np.random.default_rng(5)
    naux = 480
    worst_ratio = 0.0
    for _ in range(200):
        x = rng.standard_normal(naux).astype(np.float32)
        y = rng.standard_normal(naux).astype(np.float32)
        hot = rng.choice(naux, 5, replace=False)
        x[hot] *= np.float32(20.0)
        y[hot] *= np.float32(20.0)
        acc = np.float32(0.0)
        for q in range(naux):
            acc = np.float32(acc + np.float32(x[q] * y[q]))
        exact = math.fsum(float(a) * float(b) for a, b in zip(x, y))
        s = float(np.abs(x.astype(np.float64)) @ np.abs(y.astype(np.float64)))
        bound = (naux + 16) * 2.0 ** -24 * s
        err = abs(float(acc) - exact)
        assert err <= bound
        worst_ratio = max(worst_ratio, err / bound)
    assert worst_ratio >= 1e-3, worst_ratio


# -- adjoint --------------------------------------------------------------------

def test_x_ov_matches_scalar_definition(oracle_mod):
    """zvector.rs:1135-1175: x_ov against its scalar definition with the
    (nov + 2) eps sum|tt B| bar, on the reference's xorshift case."""
    t2, b, nocc, nvir, naux = xorshift_case()
    nov = nocc * nvir
    want = np.zeros((naux, nov))
    absum = np.zeros((naux, nov))
    for p in range(naux):
        for i in range(nocc):
            for a in range(nvir):
                s = sa = 0.0
                for j in range(nocc):
                    for bb in range(nvir):
                        tt = 2.0 * t2[i * nvir + a, j * nvir + bb] - t2[i * nvir + bb, j * nvir + a]
                        s += tt * b[p, j * nvir + bb]
                        sa += abs(tt * b[p, j * nvir + bb])
                want[p, i * nvir + a] = s
                absum[p, i * nvir + a] = sa
    got = oracle_mod.x_ov_weighted(t2, b, nocc, nvir)
    bar = (nov + 2.0) * np.finfo(float).eps
    assert np.all(np.abs(got - want) <= bar * absum)


@pytest.mark.parametrize("scaling", [(1.0, 1.0), (1.2, 1.0 / 3.0), (1.3, 0.0)])
def test_adjoint_matches_golden_closed_form(oracle_mod, golden, scaling):
    c_os, c_ss = scaling
    tag = {(1.0, 1.0): "mp2", (1.2, 1.0 / 3.0): "scs"}.get(scaling)
    for name, c in golden.items():
        if "gamma_mp2" not in c:
            continue
        eos, ess, e, gam, deps = oracle_mod.adjoint(c["b"], c["eps"], c["nocc"], c["nvir"],
                                                    c["first_occ"], c["nocc_total"], c_os, c_ss)
        assert e == pytest.approx(c_os * c["energy"][0] + c_ss * c["energy"][1], rel=1e-12)
        if tag:
            scale = np.max(np.abs(c[f"gamma_{tag}"]))
            assert np.max(np.abs(gam - c[f"gamma_{tag}"])) <= 1e-12 * scale, name
            dscale = np.max(np.abs(c[f"deps_{tag}"]))
            assert np.max(np.abs(deps - c[f"deps_{tag}"])) <= 1e-12 * dscale, name


def test_adjoint_central_differences(oracle_mod, golden):
    """dE/dB and dE/deps against central differences (h = 1e-5) of the oracle energy."""
    c = golden["res_120_7_23_1_3"]
    b, eps = c["b"].copy(), c["eps"].copy()
    dims = (c["nocc"], c["nvir"], c["first_occ"], c["nocc_total"])
    c_os, c_ss = 1.2, 1.0 / 3.0
    _, _, _, gam, deps = oracle_mod.adjoint(b, eps, *dims, c_os, c_ss)

    def energy(bb, ee):
        r = oracle_mod.spin_components(bb, ee, *dims)
        return c_os * r[0] + c_ss * r[1]

    rng = np.random.default_rng(128)
    h = 1e-5
    for _ in range(32):
        p, q = rng.integers(b.shape[0]), rng.integers(b.shape[1])
        bp, bm = b.copy(), b.copy()
        bp[p, q] += h
        bm[p, q] -= h
        fd = (energy(bp, eps) - energy(bm, eps)) / (2 * h)
        assert fd == pytest.approx(gam[p, q], abs=1e-8 * max(1.0, abs(gam[p, q])))
    for q in range(eps.size):
        ep, em = eps.copy(), eps.copy()
        ep[q] += h
        em[q] -= h
        fd = (energy(b, ep) - energy(b, em)) / (2 * h)
        assert fd == pytest.approx(deps[q], abs=1e-8 * max(1.0, abs(deps[q])))
    assert np.all(deps[:c["first_occ"]] == 0.0)


def test_adjoint_invariants(oracle_mod, golden):
    """Euler: sum B dE/dB = 4E (degree 4 in B); sum eps dE/deps = -E (degree -1);
    sum dE/deps = 0 (a common shift leaves every denominator unchanged)."""
    for name in ("res_150_9_31_2_5", "lcg_120_8_40_2"):
        c = golden[name]
        eos, ess, e, gam, deps = oracle_mod.adjoint(c["b"], c["eps"], c["nocc"], c["nvir"],
                                                    c["first_occ"], c["nocc_total"], 1.2, 0.4)
        assert np.sum(c["b"] * gam) == pytest.approx(4 * e, rel=1e-11)
        assert np.sum(c["eps"] * deps) == pytest.approx(-e, rel=1e-11)
        assert abs(np.sum(deps)) < 1e-11 * np.sum(np.abs(deps))


@pytest.mark.parametrize("scaling", [(1.0, 1.0), (1.2, 1.0 / 3.0), (0.0, 1.0)])
def test_blas_density_matches_the_loop_oracle(oracle_mod, golden, scaling):
    """The BLAS restatement of build_mp2_density (used at decane scale) equals
    the C oracle's plain loops on the golden cases, frozen core included."""
    for name in ("res_120_7_23_1_3", "res_150_9_31_2_5", "water_shapes", "lcg_120_8_40_2"):
        c = golden[name]
        args = (c["b"], c["eps"], c["nocc"], c["nvir"], c["first_occ"], c["nocc_total"])
        t2, _ = oracle_mod.t2_and_integrals(*args)
        w_oo, w_vv = oracle_mod.mp2_density(t2, c["nocc"], c["nvir"], *scaling)
        g_oo, g_vv = oracle_mod.mp2_density_blas(*args, *scaling)
        assert np.max(np.abs(g_oo - w_oo)) <= 1e-13 * np.max(np.abs(w_oo)), name
        assert np.max(np.abs(g_vv - w_vv)) <= 1e-13 * np.max(np.abs(w_vv)), name


def test_density_diagonal_is_the_orbital_energy_gradient(oracle_mod, golden):
    """dE/deps_k = 2 P_kk and dE/deps_c = 2 P_cc (the adjoint's closed form) on
    the densities the BLAS restatement forms, SCS weights."""
    c = golden["res_150_9_31_2_5"]
    args = (c["b"], c["eps"], c["nocc"], c["nvir"], c["first_occ"], c["nocc_total"])
    p_oo, p_vv = oracle_mod.mp2_density_blas(*args, 1.2, 0.4)
    _, _, _, _, deps = oracle_mod.adjoint(*args, 1.2, 0.4)
    fc, nt = c["first_occ"], c["nocc_total"]
    assert np.allclose(2 * np.diag(p_oo), deps[fc:fc + c["nocc"]], rtol=1e-12, atol=0)
    assert np.allclose(2 * np.diag(p_vv), deps[nt:nt + c["nvir"]], rtol=1e-12, atol=0)


def test_density_symmetry_and_trace(oracle_mod, golden):
    """rimp2.rs:3291-3350: P_oo, P_vv symmetric; tr P_oo + tr P_vv = 0."""
    c = golden["res_120_7_23_1_3"]
    t2, _ = oracle_mod.t2_and_integrals(c["b"], c["eps"], c["nocc"], c["nvir"], c["first_occ"],
                                        c["nocc_total"])
    p_oo, p_vv = oracle_mod.mp2_density(t2, c["nocc"], c["nvir"])
    assert np.allclose(p_oo, p_oo.T, atol=1e-14 * np.abs(p_oo).max())
    assert np.allclose(p_vv, p_vv.T, atol=1e-14 * np.abs(p_vv).max())
    assert abs(np.trace(p_oo) + np.trace(p_vv)) < 1e-13 * np.abs(np.diag(p_vv)).sum()
    # scaled density is bilinear: c_ss P_mp2 + (c_os - c_ss) P_os
    s_oo, s_vv = oracle_mod.mp2_density(t2, c["nocc"], c["nvir"], 1.2, 0.4)
    assert abs(np.trace(s_oo) + np.trace(s_vv)) < 1e-13 * np.abs(np.diag(s_vv)).sum()
    assert not np.allclose(s_oo, p_oo)
    assert math.isfinite(float(np.sum(s_vv)))
