"""Energies from explicit (ia|jb) blocks on the GPU, run with -m gpu.

* fh_rimp2_energy_from_g[_dev] (spin_components_from_g, rimp2.rs:984-1018):
  with g = B^T B it reproduces the golden energies (their generator's brute
  force) to 1e-12; with non-Gram g = B^T C + 0.3 C^T B (the shape of the robust
  fit's cross terms) it matches a brute force built from the factors by einsum,
  which fixes the (ib|ja) = g[i nvir + b][j nvir + a] convention (a transposed
  read gives another number), and the oracle's restatement; virtual counts
  across tile edges, frozen core, host and device entries.
* fh_pair_energies_from_blocks_dev on domain-fit blocks: each pair against the
  same brute force; an unbounded radius gives the canonical pair energies.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from ferric_amd import LayoutError, rimp2, synthetic  # noqa: E402
from ferric_amd.dlpno import (domain_fit_pair_blocks, estimate_pair_energies_from_b_ov,  # noqa: E402
                              pair_energies_from_blocks)


def _cuda(x):
    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


def _brute(bl, br, eps, nocc, nvir, fo, nt):
    """(e_os, e_ss) with (ia|jb) = sum_P bl[P, ia] br[P, jb] and (ib|ja) from the
    same factors, over i <= j with the mirror factor."""
    n = bl.shape[0]
    x = bl.reshape(n, nocc, nvir)
    y = br.reshape(n, nocc, nvir)
    g = np.einsum("pia,pjb->iajb", x, y)  # (ia|jb)
    eo, ev = eps[fo:fo + nocc], eps[nt:nt + nvir]
    os_ = ss = 0.0
    for i in range(nocc):
        for j in range(i, nocc):
            d = eo[i] + eo[j] - ev[:, None] - ev[None, :]
            iajb = g[i, :, j, :]
            ibja = np.einsum("pb,pa->ab", x[:, i, :], y[:, j, :])
            f = 1.0 if i == j else 2.0
            os_ += f * np.sum(iajb * iajb / d)
            ss += f * np.sum(iajb * (iajb - ibja) / d)
    return os_, ss


def test_gram_g_reproduces_the_golden_energies(golden):
    for name, c in golden.items():
        if c["nocc"] * c["nvir"] > 600:
            continue
        print("from_g golden", name, flush=True)
        g = c["b"].T @ c["b"]
        sc = rimp2.spin_components_from_g(_cuda(g), c["eps"], c["nocc"], c["nvir"],
                                          c["first_occ"], c["nocc_total"])
        os_, ss = c["energy"]
        assert abs(sc.e_os - os_) <= 1e-12 * abs(os_), name
        assert abs(sc.e_ss - ss) <= 1e-12 * abs(os_), name
        host = rimp2.spin_components_from_g(g, c["eps"], c["nocc"], c["nvir"], c["first_occ"],
                                            c["nocc_total"])
        assert (host.e_os, host.e_ss) == (sc.e_os, sc.e_ss)


@pytest.mark.parametrize("nocc,nvir,fo", [(3, 70, 0), (4, 129, 1), (2, 64, 0), (5, 17, 2)])
def test_non_gram_g_against_factor_brute_force(oracle_mod, nocc, nvir, fo):
    rng = np.random.default_rng(nocc * 1000 + nvir)
    naux = 40
    b = rng.standard_normal((naux, nocc * nvir)) * 0.05
    c = rng.standard_normal((naux, nocc * nvir)) * 0.05
    nt = fo + nocc
    eps = np.concatenate([np.sort(rng.uniform(-20.0, -15.0, fo)),
                          np.sort(rng.uniform(-1.3, -0.3, nocc)),
                          np.sort(rng.uniform(0.2, 3.0, nvir))])
    g = b.T @ c + 0.3 * (c.T @ b)
    print(f"from_g non-Gram {nocc}/{nvir} fc {fo}", flush=True)
    sc = rimp2.spin_components_from_g(_cuda(g), eps, nocc, nvir, fo, nt)
    # g = L^T R with L = [B; sqrt(0.3) C], R = [C; sqrt(0.3) B]
    big_l = np.concatenate([b, np.sqrt(0.3) * c])
    big_r = np.concatenate([c, np.sqrt(0.3) * b])
    want_os, want_ss = _brute(big_l, big_r, eps, nocc, nvir, fo, nt)
    assert abs(sc.e_os - want_os) <= 1e-12 * abs(want_os)
    assert abs(sc.e_ss - want_ss) <= 1e-12 * abs(want_os)
    orc = oracle_mod.spin_components_from_g(g, eps, nocc, nvir, fo, nt)
    assert abs(sc.e_os - orc[0]) <= 1e-12 * abs(orc[0])
    assert abs(sc.e_ss - orc[1]) <= 1e-12 * abs(orc[0])
    # the transposed read is a different number (the test can tell them apart)
    wrong = oracle_mod.spin_components_from_g(np.ascontiguousarray(g.T), eps, nocc, nvir, fo, nt)
    assert abs(wrong[1] - orc[1]) > 1e-9 * abs(orc[0])


def test_dense_g_with_positive_and_tiny_denominators(oracle_mod):
    """spin_components_from_g divides without checking the gap: with virtual
    energies below some occupied ones (denominators of both signs) and with all
    energies scaled to 1e-303 (denominators below 2^-1000) the kernel takes 1/D by
    division and still equals the oracle's plain division to 1e-12."""
    rng = np.random.default_rng(4242)
    nocc, nvir, naux = 4, 70, 30
    b = rng.standard_normal((naux, nocc * nvir)) * 0.05
    c = rng.standard_normal((naux, nocc * nvir)) * 0.05
    g = b.T @ c + 0.3 * (c.T @ b)
    eps = np.concatenate([np.sort(rng.uniform(-1.0, 0.5, nocc)),
                          np.sort(rng.uniform(-0.2, 2.0, nvir))])
    for scale in (1.0, 1e-303):
        e = eps * scale
        eo, ev = e[:nocc], e[nocc:]
        d = (eo[:, None, None, None] + eo[None, :, None, None] - ev[None, None, :, None]
             - ev[None, None, None, :])
        assert d.min() < 0 < d.max() and np.all(d != 0)
        print(f"from_g mixed-sign denominators, scale {scale}", flush=True)
        gs = g * math.sqrt(scale)  # keeps g^2 / D of order one
        sc = rimp2.spin_components_from_g(_cuda(gs), e, nocc, nvir)
        orc = oracle_mod.spin_components_from_g(gs, e, nocc, nvir)
        assert math.isfinite(sc.e_os) and math.isfinite(sc.e_ss)
        assert abs(sc.e_os - orc[0]) <= 1e-12 * abs(orc[0])
        assert abs(sc.e_ss - orc[1]) <= 1e-12 * abs(orc[0])


def test_pair_blocks_from_domain_fits():
    a, v, xyz, cen, eps = synthetic.local_fit_chain(nocc=6, nvir=70, aux_per_site=5)
    nocc, nvir = 6, 70
    pairs = [(i, j) for i in range(nocc) for j in range(i, nocc)]
    print("pair energies from domain-fit blocks", flush=True)
    l = np.linalg.cholesky(v)
    b = np.linalg.solve(l, a)
    jd, _ = domain_fit_pair_blocks(_cuda(a), _cuda(v), xyz, cen, nocc, nvir, math.inf, pairs)
    e_os, e_ss = pair_energies_from_blocks(jd, pairs, eps, nocc, nvir)
    pe = estimate_pair_energies_from_b_ov(_cuda(b), eps, nocc, nvir)
    scale = np.max(np.abs(pe.e_os))
    for k, (i, j) in enumerate(pairs):
        assert abs(e_os[k] - pe.e_os[i, j]) <= 1e-11 * scale
        assert abs(e_ss[k] - pe.e_ss[i, j]) <= 1e-11 * scale
    # a finite radius: every pair against the per-block brute force
    jd, _ = domain_fit_pair_blocks(_cuda(a), _cuda(v), xyz, cen, nocc, nvir, 3.0, pairs[::2])
    e_os, e_ss = pair_energies_from_blocks(jd, pairs[::2], eps, nocc, nvir)
    jh = jd.cpu().numpy()
    ev = eps[nocc:]
    for k, (i, j) in enumerate(pairs[::2]):
        blk = jh[k]
        d = eps[i] + eps[j] - ev[:, None] - ev[None, :]
        f = 1.0 if i == j else 2.0
        assert abs(e_os[k] - f * np.sum(blk * blk / d)) <= 1e-12 * scale
        assert abs(e_ss[k] - f * np.sum(blk * (blk - blk.T) / d)) <= 1e-12 * scale
    with pytest.raises(LayoutError):
        pair_energies_from_blocks(jd, pairs[:2], eps, nocc, nvir)
