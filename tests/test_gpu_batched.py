"""Batched local-MP2 domains (one persistent launch over every tile side), run with -m gpu.

Every domain is an independent RI-MP2 problem, so each domain's energy is checked
against the oracle (FP64: 1e-10 Eh; BF16x3: the exact-format energy within the
FP32-accumulation bound). hipGraph replays are compared with the eager result
after every replay, with eager work in between. One launch that mixes 32- and
64-wide tiles is checked against the oracle too.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import ferric_amd  # noqa: E402
from ferric_amd import Precision, synthetic  # noqa: E402
from ferric_amd.batched import domain_energies, domain_set_energies  # noqa: E402
from tests import quant  # noqa: E402

U32 = 2.0 ** -24


@pytest.fixture(scope="module")
def small_set():
    return synthetic.batched_domains(48, seed=77)


@pytest.fixture(scope="module")
def small_dev(small_set):
    return (torch.from_numpy(small_set.b_flat).cuda(), torch.from_numpy(small_set.eps_flat).cuda())


def test_fp64_domains_match_oracle(small_set, small_dev, oracle_mod):
    ds = small_set
    print("batched fp64", len(ds), "domains", flush=True)
    e_os, e_ss = domain_set_energies(ds, Precision.FP64, use_graph=False,
                                     b_dev=small_dev[0], eps_dev=small_dev[1])
    for d in range(len(ds)):
        p = ds.domain(d)
        want = oracle_mod.spin_components(p.b_ov, p.eps, p.nocc, p.nvir)
        assert abs(e_os[d] - want[0]) <= 1e-10 and abs(e_ss[d] - want[1]) <= 1e-10, d


def test_bf16x3_domains_match_exact_format(small_set, small_dev, oracle_mod):
    ds = small_set
    e_os, e_ss = domain_set_energies(ds, Precision.BF16X3, use_graph=False,
                                     b_dev=small_dev[0], eps_dev=small_dev[1])
    worst = 0.0
    for d in range(len(ds)):
        p = ds.domain(d)
        hi, lo = quant.bf16_split(p.b_ov)
        g = hi.T @ hi + hi.T @ lo + lo.T @ hi + np.diag(np.sum(lo * lo, axis=0))
        want = oracle_mod.spin_components_from_g(g, p.eps, p.nocc, p.nvir)
        bnd = oracle_mod.energy_bound(hi + lo, p.eps, p.nocc, p.nvir,
                                      eps_g=(p.naux + 16) * U32)
        assert abs(e_os[d] - want[0]) <= bnd[0] and abs(e_ss[d] - want[1]) <= bnd[1], d
        ref = oracle_mod.spin_components(p.b_ov, p.eps, p.nocc, p.nvir)
        worst = max(worst, abs(e_os[d] + e_ss[d] - ref[2]))
    print(f"bf16x3 per-domain max |dE| vs fp64 {worst:.3e}", flush=True)


@pytest.mark.parametrize("prec", [Precision.FP64, Precision.BF16X3])
def test_graph_replays_match_eager(small_set, small_dev, prec):
    ds = small_set
    eager = domain_set_energies(ds, prec, use_graph=False, b_dev=small_dev[0],
                                eps_dev=small_dev[1])
    for rep in range(4):
        print("graph replay", prec.name, rep, flush=True)
        got = domain_set_energies(ds, prec, use_graph=True, b_dev=small_dev[0],
                                  eps_dev=small_dev[1])
        assert np.array_equal(got[0], eager[0]) and np.array_equal(got[1], eager[1]), rep
        # eager work between replays
        _ = torch.ones(1000, device="cuda").sum().item()
        again = domain_set_energies(ds, prec, use_graph=False, b_dev=small_dev[0],
                                    eps_dev=small_dev[1])
        assert np.array_equal(again[0], eager[0])


@pytest.mark.parametrize("tile", ["32", "64", "128"])   # 128 runs as 64 in the batched kernel
def test_every_tile_class(small_set, small_dev, oracle_mod, tile, monkeypatch):
    monkeypatch.setenv("FERRIC_HIP_TILE", tile)
    ds = small_set
    e_os, e_ss = domain_set_energies(ds, Precision.FP64, use_graph=False,
                                     b_dev=small_dev[0], eps_dev=small_dev[1])
    for d in range(0, len(ds), 5):
        p = ds.domain(d)
        want = oracle_mod.spin_components(p.b_ov, p.eps, p.nocc, p.nvir)
        assert abs(e_os[d] - want[0]) <= 1e-10 and abs(e_ss[d] - want[1]) <= 1e-10, (tile, d)


@pytest.mark.parametrize("prec", [Precision.FP64, Precision.BF16X3])
def test_mixed_tile_sides_in_one_launch(small_set, small_dev, oracle_mod, prec):
    from ferric_amd import autotune

    ds = small_set
    assert ds.nvir.min() <= 60 < ds.nvir.max()
    try:
        autotune.install([{"precision": prec.name, "nvir_max": 60, "naux_max": 1 << 20,
                           "tile": 32},
                          {"precision": prec.name, "nvir_max": 1 << 16, "naux_max": 1 << 20,
                           "tile": 64}])
        e_os, e_ss = domain_set_energies(ds, prec, use_graph=True, b_dev=small_dev[0],
                                         eps_dev=small_dev[1])
    finally:
        autotune.install(autotune.load_table())
    print("mixed tiles", prec.name, flush=True)
    for d in range(len(ds)):
        p = ds.domain(d)
        if prec == Precision.FP64:
            want = oracle_mod.spin_components(p.b_ov, p.eps, p.nocc, p.nvir)
            assert abs(e_os[d] - want[0]) <= 1e-10 and abs(e_ss[d] - want[1]) <= 1e-10, d
        else:
            hi, lo = quant.bf16_split(p.b_ov)
            g = hi.T @ hi + hi.T @ lo + lo.T @ hi + np.diag(np.sum(lo * lo, axis=0))
            want = oracle_mod.spin_components_from_g(g, p.eps, p.nocc, p.nvir)
            bnd = oracle_mod.energy_bound(hi + lo, p.eps, p.nocc, p.nvir,
                                          eps_g=(p.naux + 16) * U32)
            assert abs(e_os[d] - want[0]) <= bnd[0] and abs(e_ss[d] - want[1]) <= bnd[1], d


def _edge_set():
    """Domains at the edges of the tiling: one occupied and one virtual
    orbital, one auxiliary function, virtual counts just past a 16- and a
    64-multiple, auxiliary counts below and just past a panel, domains laid out
    with gaps between their blocks."""
    rng = np.random.default_rng(2024)
    shapes = [(1, 1, 1), (1, 17, 7), (2, 16, 8), (3, 65, 9), (5, 33, 17), (4, 64, 129),
              (12, 80, 3), (1, 129, 40), (7, 20, 100)]
    nocc = np.array([x[0] for x in shapes])
    nvir = np.array([x[1] for x in shapes])
    naux = np.array([x[2] for x in shapes])
    gap = 5
    sizes = naux * nocc * nvir
    b_off = np.concatenate([[0], np.cumsum(sizes + gap)[:-1]]).astype(np.int64)
    e_off = np.concatenate([[0], np.cumsum(nocc + nvir + gap)[:-1]]).astype(np.int64)
    b = np.full(int(b_off[-1] + sizes[-1] + gap), np.nan)
    eps = np.full(int(e_off[-1] + nocc[-1] + nvir[-1] + gap), np.nan)
    for d in range(len(shapes)):
        no, nv, na = shapes[d]
        b[b_off[d]:b_off[d] + na * no * nv] = synthetic._b(rng, na, no, nv).ravel()
        eps[e_off[d]:e_off[d] + no + nv] = np.concatenate(
            [np.sort(rng.uniform(-1.2, -0.35, no)), np.sort(rng.uniform(0.12, 5.0, nv))])
    return synthetic.DomainSet(2024, nocc, nvir, naux, b_off, e_off, b, eps)


@pytest.mark.parametrize("tile", ["", "32", "64"])
def test_edge_shapes_against_oracle(oracle_mod, tile, monkeypatch):
    """NaN-filled gaps between the blocks would poison any read outside a
    domain's own block."""
    if tile:
        monkeypatch.setenv("FERRIC_HIP_TILE", tile)
    ds = _edge_set()
    bd = torch.from_numpy(ds.b_flat).cuda()
    ed = torch.from_numpy(ds.eps_flat).cuda()
    for prec in (Precision.FP64, Precision.BF16X3):
        print("batched edge shapes", tile or "table", prec.name, flush=True)
        e_os, e_ss = domain_set_energies(ds, prec, use_graph=False, b_dev=bd, eps_dev=ed)
        for d in range(len(ds)):
            p = ds.domain(d)
            if prec == Precision.FP64:
                want = oracle_mod.spin_components(p.b_ov, p.eps, p.nocc, p.nvir)
                tol = 1e-12 * max(1.0, abs(want[2]))
                assert abs(e_os[d] - want[0]) <= tol and abs(e_ss[d] - want[1]) <= tol, d
            else:
                hi, lo = quant.bf16_split(p.b_ov)
                g = hi.T @ hi + hi.T @ lo + lo.T @ hi + np.diag(np.sum(lo * lo, axis=0))
                want = oracle_mod.spin_components_from_g(g, p.eps, p.nocc, p.nvir)
                bnd = oracle_mod.energy_bound(hi + lo, p.eps, p.nocc, p.nvir,
                                              eps_g=(p.naux + 32) * U32)
                assert abs(e_os[d] - want[0]) <= bnd[0] + 1e-15, d
                assert abs(e_ss[d] - want[1]) <= bnd[1] + 1e-15, d


def test_full_512_domains_against_single_problem_path():
    ds = synthetic.batched_domains(512)
    b = torch.from_numpy(ds.b_flat).cuda()
    e = torch.from_numpy(ds.eps_flat).cuda()
    e_os, e_ss = domain_set_energies(ds, Precision.FP64, use_graph=True, b_dev=b, eps_dev=e)
    from ferric_amd import rimp2

    worst = 0.0
    for d in range(len(ds)):
        p = ds.domain(d)
        bt = b[ds.b_offset[d]:ds.b_offset[d] + p.naux * p.nocc * p.nvir].view(p.naux, -1)
        sc = rimp2.spin_components(bt, p.eps, p.nocc, p.nvir)
        worst = max(worst, abs(sc.e_os - e_os[d]), abs(sc.e_ss - e_ss[d]))
    print(f"512 domains: max |batched - single| {worst:.2e}", flush=True)
    assert worst <= 1e-11


def test_descriptor_errors(small_set, small_dev):
    ds = small_set
    nv = ds.nvir.copy()
    nv[3] += 10_000   # B block now runs past the buffer
    with pytest.raises(ferric_amd.LayoutError):
        domain_energies(small_dev[0], small_dev[1], ds.nocc, nv, ds.naux, ds.b_offset,
                        ds.eps_offset)
    with pytest.raises(ferric_amd.UnsupportedError):
        domain_energies(small_dev[0], small_dev[1], ds.nocc, ds.nvir, ds.naux, ds.b_offset,
                        ds.eps_offset, precision=Precision.FP8_E4M3)
