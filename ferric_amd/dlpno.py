"""Local (DLPNO) MP2 pair layer on the device: pair energies and the pair screen.

Mirrors the pair-driven path of crates/ferric-mp2/src/dlpno_mp2.rs and
crates/ferric-mp2/src/pair_energy_screen.rs:

* :func:`estimate_pair_energies_from_b_ov` (dlpno_mp2.rs:527-594): the MP2
  energy of every occupied pair from one pass of the fused pair kernel
  (``fh_rimp2_pair_energies[_dev]``), any precision of the energy;
* :class:`DlpnoConfig` (dlpno_mp2.rs:86-170) and
  :func:`build_pair_domains_by_energy` (pair_energy_screen.rs:206-282): pairs
  ``i <= j`` kept when ``|e_ij| >= t_cut_pairs`` (diagonals always), and the
  geometric coupling lists on pair centroids;
* :func:`pno_pair_energies` (``fh_dlpno_pno_pair_energies[_dev]``): the
  per-pair body of dlpno_mp2_from_b_ov with PNO truncation -- pair density,
  its eigenvectors above ``t_cut_pno`` (local_pno.rs:132-200), semicanonical
  rotation and the energy in that basis (dlpno_mp2.rs:249-325), batched over
  pairs on the device;
* :func:`estimate_pair_energies`, :func:`pno_pair_energies_from_g` and
  :func:`dlpno_mp2_spin_components`: the same from an explicit dense (ia|jb)
  matrix (pair_energy_screen.rs:125-190, dlpno_mp2.rs:354-440);
* :func:`domain_fit_pair_blocks` (``fh_domain_fit_pair_blocks[_dev]``): the
  domain-local RI fit J_ij = A_{i,D}^T V_DD^-1 A_{j,D} of each pair over the aux
  functions near either of its centres (lmp2_amplitude.rs:1707-1758);
* :func:`dlpno_mp2_from_b_ov` (dlpno_mp2.rs:670-760). With ``t_cut_pno == 0``
  every pair's PNO transform is a square orthogonal rotation, and the
  semicanonical rotation after it re-diagonalises the virtual Fock block, so
  both cancel out of the pair energy (the reference's exactness contract,
  dlpno_mp2.rs:797-830): the energy is the sum of the retained pairs' canonical
  energies, which the fused pair kernel computes without any eigensolve. With
  ``t_cut_pno > 0`` the retained pairs go through :func:`pno_pair_energies`.
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass, field

import numpy as np

from ._lib import LayoutError, PnoDiagnosticsC, Precision, UnsupportedError, check, lib
from .rimp2 import SpinComponents, _is_torch_cuda

DEFAULT_T_CUT_PAIRS = 1e-5  # dlpno_mp2.rs:121


@dataclass
class DlpnoConfig:
    """dlpno_mp2.rs:86-170. Default: pair-energy screening at 1e-5 Eh, PNO and
    distance screens off."""

    pair_cutoff: float = math.inf
    coupling_cutoff: float = math.inf
    t_cut_pno: float = 0.0
    t_cut_pairs: float = DEFAULT_T_CUT_PAIRS

    @staticmethod
    def exact() -> "DlpnoConfig":
        return DlpnoConfig(math.inf, math.inf, 0.0, 0.0)

    def is_exact(self) -> bool:
        return math.isinf(self.pair_cutoff) and self.t_cut_pno == 0.0 and self.t_cut_pairs == 0.0


@dataclass
class PairEnergies:
    """pair_energy_screen.rs:65-90: ``e`` is the symmetric nocc x nocc matrix of
    pair energies (mirror factor applied); ``e_os`` + ``e_ss`` = ``e``."""

    nocc: int
    e: np.ndarray
    e_os: np.ndarray
    e_ss: np.ndarray

    def e_ij(self, i: int, j: int) -> float:
        return float(self.e[min(i, j), max(i, j)])

    def total(self) -> float:
        return float(sum(self.e[i, j] for i in range(self.nocc) for j in range(i, self.nocc)))

    def discarded_energy(self, t_cut_pairs: float) -> float:
        """pair_energy_screen.rs:90-105: the energy of the off-diagonal pairs a
        threshold drops (|e_ij| < t_cut_pairs)."""
        return float(sum(self.e[i, j] for i in range(self.nocc) for j in range(i + 1, self.nocc)
                         if abs(self.e[i, j]) < t_cut_pairs))


@dataclass
class PairDomains:
    """pair_domains.rs:51-65 (the fields the MP2 layer uses)."""

    nocc: int
    pairs: list
    coupled: list
    centers: np.ndarray
    pair_cutoff: float = math.inf
    coupling_cutoff: float = math.inf

    def pair_retention(self) -> float:
        """pair_domains.rs:68-75."""
        total = self.nocc * (self.nocc + 1) // 2
        return 1.0 if total == 0 else len(self.pairs) / total

    def coupling_retention(self) -> float:
        """pair_domains.rs:77-88."""
        total = len(self.pairs) ** 2
        return 1.0 if total == 0 else sum(len(c) for c in self.coupled) / total

    def is_complete(self) -> bool:
        """pair_domains.rs:90-98."""
        n = len(self.pairs)
        return n == self.nocc * (self.nocc + 1) // 2 and all(len(c) == n for c in self.coupled)


@dataclass
class DlpnoDiagnostics:
    """dlpno_mp2.rs:174-185."""

    pair_retention: float
    virtual_retention: float
    max_discarded_weight: float
    n_pairs: int
    extra: dict = field(default_factory=dict)


def estimate_pair_energies_from_b_ov(b_ov, eps, nocc: int, nvir: int, first_occ: int = 0,
                                     nocc_total: int | None = None,
                                     precision: Precision = Precision.FP64) -> PairEnergies:
    """dlpno_mp2.rs:527-594 on the device. ``b_ov``: numpy array (host entry) or
    contiguous CUDA tensor (device entry on the current torch stream)."""
    if nocc_total is None:
        nocc_total = first_occ + nocc
    e = np.ascontiguousarray(np.asarray(eps, dtype=np.float64))
    e_os = np.zeros((nocc, nocc))
    e_ss = np.zeros((nocc, nocc))
    L = lib()
    dp = ctypes.POINTER(ctypes.c_double)
    if _is_torch_cuda(b_ov):
        import torch

        if b_ov.dim() != 2 or not b_ov.is_contiguous():
            raise LayoutError("b_ov must be a contiguous 2-D (naux, nocc*nvir) tensor")
        if b_ov.dtype not in (torch.float64, torch.float32):
            raise LayoutError(f"b_ov dtype {b_ov.dtype} is not float64/float32")
        if b_ov.shape[1] != nocc * nvir:
            raise LayoutError(f"b_ov is {tuple(b_ov.shape)}, expected width {nocc * nvir}")
        dt = 0 if b_ov.dtype == torch.float64 else 1
        stream = torch.cuda.current_stream(b_ov.device).cuda_stream
        st = L.fh_rimp2_pair_energies_dev(
            ctypes.c_void_p(b_ov.data_ptr()), dt, b_ov.shape[0], nocc, nvir, e.ctypes.data_as(dp),
            e.size, first_occ, nocc_total, int(precision), ctypes.c_void_p(stream),
            e_os.ctypes.data_as(dp), e_ss.ctypes.data_as(dp))
    else:
        b = np.ascontiguousarray(np.asarray(b_ov, dtype=np.float64))
        if b.ndim != 2:
            raise LayoutError("b_ov must be 2-D (naux, nocc*nvir)")
        if b.shape[1] != nocc * nvir:
            raise LayoutError(f"b_ov is {b.shape}, expected width {nocc * nvir}")
        st = L.fh_rimp2_pair_energies(b.ctypes.data_as(dp), b.shape[0], nocc, nvir,
                                      e.ctypes.data_as(dp), e.size, first_occ, nocc_total,
                                      int(precision), e_os.ctypes.data_as(dp),
                                      e_ss.ctypes.data_as(dp))
    check(st, "fh_rimp2_pair_energies")
    return PairEnergies(nocc, e_os + e_ss, e_os, e_ss)


def build_pair_domains_by_energy(centers, pair_energies: PairEnergies, t_cut_pairs: float,
                                 coupling_cutoff_bohr: float = math.inf) -> PairDomains:
    """pair_energy_screen.rs:206-282: pairs i <= j with |e_ij| >= t_cut_pairs,
    diagonals always kept; coupled[p] = the pairs whose centroids lie within
    the coupling cutoff of p's (p itself always)."""
    c = np.asarray(centers, dtype=np.float64)
    if c.ndim != 2 or c.shape[1] != 3:
        raise LayoutError(f"Boys centers must be (nocc, 3); got {c.shape}")
    if c.shape[0] != pair_energies.nocc:
        raise LayoutError(f"centers has {c.shape[0]} rows but pair energies were built for "
                          f"nocc={pair_energies.nocc}")
    if not t_cut_pairs >= 0.0:
        raise LayoutError(f"t_cut_pairs must be >= 0 (got {t_cut_pairs}); use 0.0 to disable")
    if coupling_cutoff_bohr < 0.0:
        raise LayoutError(f"coupling cutoff must be >= 0 (got {coupling_cutoff_bohr})")
    nocc = c.shape[0]
    iu, ju = np.triu_indices(nocc)  # row-major i <= j, the reference's loop order
    e = np.asarray(pair_energies.e, dtype=np.float64)
    keep = (iu == ju) | (np.abs(e[iu, ju]) >= t_cut_pairs)
    pi, pj = iu[keep], ju[keep]
    pairs = list(zip(pi.tolist(), pj.tolist()))
    cents = 0.5 * (c[pi] + c[pj])
    cut_sq = coupling_cutoff_bohr * coupling_cutoff_bohr
    n = len(pairs)
    if cut_sq == math.inf:  # every centroid distance is within an infinite cutoff
        everything = list(range(n))
        coupled = [everything.copy() for _ in range(n)]
    else:
        coupled = []
        for r0 in range(0, n, 256):
            blk = cents[r0:r0 + 256]
            d = np.zeros((blk.shape[0], n))
            for ax in range(3):  # the reference's axis order of the sum
                d += (blk[:, ax:ax + 1] - cents[None, :, ax]) ** 2
            m = d <= cut_sq
            m[np.arange(blk.shape[0]), np.arange(r0, r0 + blk.shape[0])] = True
            coupled.extend(np.flatnonzero(row).tolist() for row in m)
    return PairDomains(nocc, pairs, coupled, c.copy(), math.inf, coupling_cutoff_bohr)


def pno_pair_energies(b_ov, eps, nocc: int, nvir: int, first_occ: int, nocc_total: int, pairs,
                      t_cut_pno: float):
    """Per-pair energies in each pair's truncated, semicanonical PNO basis
    (dlpno_mp2.rs:700-740 per pair): (e_os, e_ss, kept_pnos, max_discarded_weight)
    with e_os, e_ss arrays over ``pairs`` (mirror factor applied), FP64.
    ``b_ov``: numpy array (host entry) or contiguous float64 CUDA tensor (device
    entry on the current torch stream)."""
    pl = np.ascontiguousarray(np.asarray(pairs, dtype=np.int32).reshape(-1, 2))
    npairs = pl.shape[0]
    e = np.ascontiguousarray(np.asarray(eps, dtype=np.float64))
    e_os = np.zeros(npairs)
    e_ss = np.zeros(npairs)
    diag = PnoDiagnosticsC()
    L = lib()
    dp = ctypes.POINTER(ctypes.c_double)
    if _is_torch_cuda(b_ov):
        import torch

        if b_ov.dim() != 2 or not b_ov.is_contiguous() or b_ov.dtype != torch.float64:
            raise LayoutError("b_ov must be a contiguous 2-D float64 (naux, nocc*nvir) tensor")
        if b_ov.shape[1] != nocc * nvir:
            raise LayoutError(f"b_ov is {tuple(b_ov.shape)}, expected width {nocc * nvir}")
        stream = torch.cuda.current_stream(b_ov.device).cuda_stream
        st = L.fh_dlpno_pno_pair_energies_dev(
            ctypes.c_void_p(b_ov.data_ptr()), b_ov.shape[0], nocc, nvir, e.ctypes.data_as(dp),
            e.size, first_occ, nocc_total, ctypes.c_void_p(pl.ctypes.data), npairs,
            float(t_cut_pno), ctypes.c_void_p(stream), e_os.ctypes.data_as(dp),
            e_ss.ctypes.data_as(dp), ctypes.byref(diag))
    else:
        b = np.ascontiguousarray(np.asarray(b_ov, dtype=np.float64))
        if b.ndim != 2 or b.shape[1] != nocc * nvir:
            raise LayoutError(f"b_ov is {b.shape}, expected (naux, {nocc * nvir})")
        st = L.fh_dlpno_pno_pair_energies(
            b.ctypes.data_as(dp), b.shape[0], nocc, nvir, e.ctypes.data_as(dp), e.size,
            first_occ, nocc_total, ctypes.c_void_p(pl.ctypes.data), npairs, float(t_cut_pno),
            e_os.ctypes.data_as(dp), e_ss.ctypes.data_as(dp), ctypes.byref(diag))
    check(st, "fh_dlpno_pno_pair_energies")
    return e_os, e_ss, int(diag.kept_pnos), float(diag.max_discarded_weight)


def domain_fit_pair_blocks(a, v, aux_xyz, occ_centers, nocc: int, nvir: int, radius: float,
                           pairs=None):
    """lmp2_amplitude.rs:1707-1758 for every listed pair (all i <= j when
    ``pairs`` is None): (J, max_domain) with J[k] the nvir x nvir block of pair
    k, J[k][a, b] = (ia|jb) fitted in the pair's aux domain. ``a`` (naux,
    nocc*nvir) unwhitened (P|ia) and ``v`` (naux, naux): both numpy (host entry,
    J numpy) or both contiguous float64 CUDA tensors (device entry on the
    current torch stream, J a CUDA tensor)."""
    if pairs is None:
        pairs = [(i, j) for i in range(nocc) for j in range(i, nocc)]
    pl = np.ascontiguousarray(np.asarray(pairs, dtype=np.int32).reshape(-1, 2))
    npairs = pl.shape[0]
    xyz = np.ascontiguousarray(np.asarray(aux_xyz, dtype=np.float64))
    cen = np.ascontiguousarray(np.asarray(occ_centers, dtype=np.float64))
    if xyz.ndim != 2 or xyz.shape[1] != 3 or cen.shape != (nocc, 3):
        raise LayoutError(f"aux_xyz must be (naux, 3) and centers ({nocc}, 3); got {xyz.shape}, "
                          f"{cen.shape}")
    L = lib()
    dp = ctypes.POINTER(ctypes.c_double)
    dmax = ctypes.c_size_t()
    if _is_torch_cuda(a):
        import torch

        for t, shape in ((a, (xyz.shape[0], nocc * nvir)), (v, (xyz.shape[0], xyz.shape[0]))):
            if (not _is_torch_cuda(t) or t.dtype != torch.float64 or not t.is_contiguous()
                    or tuple(t.shape) != shape):
                raise LayoutError(f"expected a contiguous float64 CUDA tensor of shape {shape}")
        out = torch.empty((npairs, nvir, nvir), dtype=torch.float64, device=a.device)
        stream = torch.cuda.current_stream(a.device).cuda_stream
        st = L.fh_domain_fit_pair_blocks_dev(
            ctypes.c_void_p(a.data_ptr()), ctypes.c_void_p(v.data_ptr()), xyz.shape[0], nocc,
            nvir, xyz.ctypes.data_as(dp), cen.ctypes.data_as(dp), float(radius),
            ctypes.c_void_p(pl.ctypes.data), npairs, ctypes.c_void_p(stream),
            ctypes.c_void_p(out.data_ptr()), ctypes.byref(dmax))
    else:
        an = np.ascontiguousarray(np.asarray(a, dtype=np.float64))
        vn = np.ascontiguousarray(np.asarray(v, dtype=np.float64))
        if an.shape != (xyz.shape[0], nocc * nvir) or vn.shape != (xyz.shape[0], xyz.shape[0]):
            raise LayoutError(f"a {an.shape} / v {vn.shape} do not match naux={xyz.shape[0]}, "
                              f"nocc*nvir={nocc * nvir}")
        out = np.empty((npairs, nvir, nvir))
        st = L.fh_domain_fit_pair_blocks(
            an.ctypes.data_as(dp), vn.ctypes.data_as(dp), xyz.shape[0], nocc, nvir,
            xyz.ctypes.data_as(dp), cen.ctypes.data_as(dp), float(radius),
            ctypes.c_void_p(pl.ctypes.data), npairs, out.ctypes.data_as(dp), ctypes.byref(dmax))
    check(st, "fh_domain_fit_pair_blocks")
    return out, int(dmax.value)


def pair_energies_from_blocks(blocks, pairs, eps, nocc: int, nvir: int, first_occ: int = 0,
                              nocc_total: int | None = None):
    """(e_os, e_ss) arrays over ``pairs`` from per-pair (ia|jb) blocks (a
    contiguous float64 CUDA tensor (npairs, nvir, nvir), the layout of
    :func:`domain_fit_pair_blocks`), the pair terms of spin_components_from_g
    (rimp2.rs:984-1018) with the mirror factor applied. FP64."""
    import torch

    if nocc_total is None:
        nocc_total = first_occ + nocc
    pl = np.ascontiguousarray(np.asarray(pairs, dtype=np.int32).reshape(-1, 2))
    npairs = pl.shape[0]
    if (not _is_torch_cuda(blocks) or blocks.dtype != torch.float64
            or not blocks.is_contiguous() or tuple(blocks.shape) != (npairs, nvir, nvir)):
        raise LayoutError(f"blocks must be a contiguous float64 CUDA tensor ({npairs}, {nvir}, "
                          f"{nvir})")
    e = np.ascontiguousarray(np.asarray(eps, dtype=np.float64))
    e_os, e_ss = np.zeros(npairs), np.zeros(npairs)
    dp = ctypes.POINTER(ctypes.c_double)
    stream = torch.cuda.current_stream(blocks.device).cuda_stream
    st = lib().fh_pair_energies_from_blocks_dev(
        ctypes.c_void_p(blocks.data_ptr()), ctypes.c_void_p(pl.ctypes.data), npairs, nocc, nvir,
        e.ctypes.data_as(dp), e.size, first_occ, nocc_total, ctypes.c_void_p(stream),
        e_os.ctypes.data_as(dp), e_ss.ctypes.data_as(dp))
    check(st, "fh_pair_energies_from_blocks_dev")
    return e_os, e_ss


def _dense_g(g, nocc, nvir):
    import torch

    nov = nocc * nvir
    if (not _is_torch_cuda(g) or g.dtype != torch.float64 or not g.is_contiguous()
            or tuple(g.shape) != (nov, nov)):
        raise LayoutError(f"g must be a contiguous float64 CUDA tensor ({nov}, {nov})")
    return ctypes.c_void_p(torch.cuda.current_stream(g.device).cuda_stream)


def estimate_pair_energies(g, eps, nocc: int, nvir: int, first_occ: int = 0,
                           nocc_total: int | None = None) -> PairEnergies:
    """pair_energy_screen.rs:125-190 on the device: the pair energies of an
    explicit dense (ia|jb) matrix ``g`` (a CUDA tensor, not necessarily B^T B)."""
    if nocc_total is None:
        nocc_total = first_occ + nocc
    stream = _dense_g(g, nocc, nvir)
    e = np.ascontiguousarray(np.asarray(eps, dtype=np.float64))
    e_os, e_ss = np.zeros((nocc, nocc)), np.zeros((nocc, nocc))
    dp = ctypes.POINTER(ctypes.c_double)
    st = lib().fh_pair_energies_from_g_dev(ctypes.c_void_p(g.data_ptr()), nocc, nvir,
                                           e.ctypes.data_as(dp), e.size, first_occ, nocc_total,
                                           stream, e_os.ctypes.data_as(dp),
                                           e_ss.ctypes.data_as(dp))
    check(st, "fh_pair_energies_from_g_dev")
    return PairEnergies(nocc, e_os + e_ss, e_os, e_ss)


def pno_pair_energies_from_g(g, eps, nocc: int, nvir: int, first_occ: int, nocc_total: int,
                             pairs, t_cut_pno: float):
    """:func:`pno_pair_energies` with each pair's block read from a dense device
    ``g`` (dlpno_mp2.rs:390-430)."""
    stream = _dense_g(g, nocc, nvir)
    pl = np.ascontiguousarray(np.asarray(pairs, dtype=np.int32).reshape(-1, 2))
    e = np.ascontiguousarray(np.asarray(eps, dtype=np.float64))
    e_os, e_ss = np.zeros(pl.shape[0]), np.zeros(pl.shape[0])
    diag = PnoDiagnosticsC()
    dp = ctypes.POINTER(ctypes.c_double)
    st = lib().fh_dlpno_pno_pair_energies_from_g_dev(
        ctypes.c_void_p(g.data_ptr()), nocc, nvir, e.ctypes.data_as(dp), e.size, first_occ,
        nocc_total, ctypes.c_void_p(pl.ctypes.data), pl.shape[0], float(t_cut_pno), stream,
        e_os.ctypes.data_as(dp), e_ss.ctypes.data_as(dp), ctypes.byref(diag))
    check(st, "fh_dlpno_pno_pair_energies_from_g_dev")
    return e_os, e_ss, int(diag.kept_pnos), float(diag.max_discarded_weight)


def _sum_pairs(domains, nvir, pos, pss, kept, worst):
    e_os = e_ss = 0.0
    for k in range(len(domains.pairs)):  # the reference's order of accumulation
        e_os += pos[k]
        e_ss += pss[k]
    n = len(domains.pairs)
    vret = 1.0 if n == 0 or nvir == 0 else kept / (n * nvir)
    return (SpinComponents(e_os, e_ss, e_os + e_ss),
            DlpnoDiagnostics(domains.pair_retention(), vret, worst, n))


def dlpno_mp2_spin_components(g, eps, nocc: int, nvir: int, first_occ: int, nocc_total: int,
                              domains: PairDomains, cfg: DlpnoConfig = DlpnoConfig()):
    """dlpno_mp2.rs:354-440: (SpinComponents, DlpnoDiagnostics) over the retained
    pairs of ``domains`` from an explicit dense ``g`` (a CUDA tensor). FP64."""
    if cfg.t_cut_pno < 0.0:
        raise LayoutError(f"t_cut_pno must be >= 0 (got {cfg.t_cut_pno}); use 0.0 to disable "
                          "truncation")
    if domains.nocc != nocc:
        raise LayoutError(f"domains were built for nocc={domains.nocc}, but nocc={nocc}")
    if not cfg.t_cut_pno == 0.0:
        return _sum_pairs(domains, nvir, *pno_pair_energies_from_g(
            g, eps, nocc, nvir, first_occ, nocc_total, domains.pairs, cfg.t_cut_pno))
    pe = estimate_pair_energies(g, eps, nocc, nvir, first_occ, nocc_total)
    e_os = math.fsum(pe.e_os[i, j] for i, j in domains.pairs)
    e_ss = math.fsum(pe.e_ss[i, j] for i, j in domains.pairs)
    return (SpinComponents(e_os, e_ss, e_os + e_ss),
            DlpnoDiagnostics(domains.pair_retention(), 1.0, 0.0, len(domains.pairs)))


def dlpno_mp2_from_b_ov(b_ov, eps, nocc: int, nvir: int, first_occ: int, nocc_total: int,
                        domains: PairDomains, cfg: DlpnoConfig = DlpnoConfig(),
                        precision: Precision = Precision.FP64,
                        pair_energies: PairEnergies | None = None):
    """dlpno_mp2.rs:670-760: (SpinComponents, DlpnoDiagnostics) summed over the
    retained pairs of ``domains`` in their order. At ``t_cut_pno == 0`` the
    canonical pair energies serve (pass the ``pair_energies`` already computed
    for the screen to reuse them); above it the PNO path runs, in FP64."""
    if cfg.t_cut_pno < 0.0:
        raise LayoutError(f"t_cut_pno must be >= 0 (got {cfg.t_cut_pno}); use 0.0 to disable "
                          "truncation")
    if domains.nocc != nocc:
        raise LayoutError(f"domains were built for nocc={domains.nocc}, but nocc={nocc}")
    if not cfg.t_cut_pno == 0.0:
        if precision != Precision.FP64:
            raise UnsupportedError("PNO truncation runs in FP64")
        return _sum_pairs(domains, nvir, *pno_pair_energies(
            b_ov, eps, nocc, nvir, first_occ, nocc_total, domains.pairs, cfg.t_cut_pno))
    pe = pair_energies
    if pe is None:
        pe = estimate_pair_energies_from_b_ov(b_ov, eps, nocc, nvir, first_occ, nocc_total,
                                              precision)
    e_os = math.fsum(pe.e_os[i, j] for i, j in domains.pairs)
    e_ss = math.fsum(pe.e_ss[i, j] for i, j in domains.pairs)
    diag = DlpnoDiagnostics(domains.pair_retention(), 1.0, 0.0, len(domains.pairs))
    return SpinComponents(e_os, e_ss, e_os + e_ss), diag
