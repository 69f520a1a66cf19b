"""Seeded synthetic inputs of the benchmark workloads (numpy PCG64 streams).

Each generator returns a :class:`Problem` in the reference layout: ``b_ov`` is
row-major (naux, nocc*nvir) and ``eps`` the full orbital-energy vector
[occupied..., virtual...] with ``first_occ = 0``, ``nocc_total = nocc``.
Orbital energies are sorted ascending within each block.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np


@dataclass
class Problem:
    name: str
    seed: int
    b_ov: np.ndarray
    eps: np.ndarray
    nocc: int
    nvir: int
    naux: int
    first_occ: int = 0

    @property
    def nocc_total(self) -> int:
        return self.first_occ + self.nocc

    @property
    def eps_occ(self) -> np.ndarray:
        return self.eps[self.first_occ:self.first_occ + self.nocc]

    @property
    def eps_vir(self) -> np.ndarray:
        return self.eps[self.nocc_total:self.nocc_total + self.nvir]

    @property
    def flops(self) -> float:
        """Unique-pair GEMM work nocc(nocc+1) nvir^2 naux (2 FLOP per multiply-add)."""
        return float(self.nocc) * (self.nocc + 1) * self.nvir ** 2 * self.naux


def _loguniform(rng, lo, hi, n):
    return np.exp(rng.uniform(math.log(lo), math.log(hi), n))


def _b(rng, naux, nocc, nvir, dtype=np.float64):
    b = rng.standard_normal((naux, nocc * nvir), dtype=dtype)
    b *= dtype(math.sqrt(0.09 / naux)) if dtype is np.float32 else math.sqrt(0.09 / naux)
    return b


def water_shapes(seed: int = 1001) -> Problem:
    """Water cc-pVDZ / cc-pVDZ-RI shapes (nocc 5, nvir 19, naux 84): one core
    orbital energy U[-20.6, -20.4], four valence U[-1.4, -0.45], virtuals
    log-uniform on [0.15, 4.0]."""
    rng = np.random.default_rng(seed)
    nocc, nvir, naux = 5, 19, 84
    b = _b(rng, naux, nocc, nvir)
    occ = np.sort(np.concatenate([rng.uniform(-20.6, -20.4, 1), rng.uniform(-1.4, -0.45, 4)]))
    vir = np.sort(_loguniform(rng, 0.15, 4.0, nvir))
    return Problem("water-shapes 5/19/84", seed, b, np.concatenate([occ, vir]), nocc, nvir, naux)


def decane_scale(seed: int = 1002, dtype=np.float64) -> Problem:
    """Decane-scale shapes (nocc 41, nvir 209, naux 960): ten core energies
    U[-11.3, -11.1], 31 valence U[-1.2, -0.35], virtuals log-uniform [0.12, 5.0]."""
    rng = np.random.default_rng(seed)
    nocc, nvir, naux = 41, 209, 960
    b = _b(rng, naux, nocc, nvir, dtype)
    occ = np.sort(np.concatenate([rng.uniform(-11.3, -11.1, 10), rng.uniform(-1.2, -0.35, 31)]))
    vir = np.sort(_loguniform(rng, 0.12, 5.0, nvir))
    return Problem("decane-scale 41/209/960", seed, b, np.concatenate([occ, vir]), nocc, nvir,
                   naux)


def mixed_large(seed: int = 1003, nocc: int = 100, nvir: int = 900, naux: int = 3600,
                hot_fraction: float = 0.01, hot_scale: float = 20.0) -> Problem:
    """The mixed-precision workload (default 100/900/3600): B stored FP32 from
    N(0, 0.09/naux) with 1% of the Q rows scaled x20; occupied U[-1.3, -0.3];
    virtuals log-uniform on [0.1, 6.0]."""
    rng = np.random.default_rng(seed)
    b = _b(rng, naux, nocc, nvir, np.float32)
    nhot = max(1, int(math.ceil(hot_fraction * naux)))
    hot = np.sort(rng.choice(naux, size=nhot, replace=False))
    b[hot, :] *= np.float32(hot_scale)
    occ = np.sort(rng.uniform(-1.3, -0.3, nocc))
    vir = np.sort(_loguniform(rng, 0.1, 6.0, nvir))
    return Problem(f"mixed {nocc}/{nvir}/{naux}", seed, b, np.concatenate([occ, vir]), nocc,
                   nvir, naux)


def local_chain(seed: int = 1006, nocc: int = 24, nvir: int = 72, rows_per_site: int = 24,
                reach: int = 3, spacing: float = 2.5):
    """(Problem, centers): localized occupied orbitals on a line (x = spacing * i
    bohr) whose B columns live on the aux rows of the sites within ``reach`` of
    their own, with weight exp(-|k - i|). Pairs further apart than 2 * reach
    have exactly zero pair energy; the nearer ones decay with distance, so an
    energy screen keeps a band of pairs. Occupied U[-1.2, -0.4], virtuals
    log-uniform on [0.2, 4.0]."""
    rng = np.random.default_rng(seed)
    naux = nocc * rows_per_site
    b = np.zeros((naux, nocc * nvir))
    for i in range(nocc):
        for k in range(max(0, i - reach), min(nocc, i + reach + 1)):
            blk = rng.standard_normal((rows_per_site, nvir)) * math.exp(-abs(k - i))
            b[k * rows_per_site:(k + 1) * rows_per_site, i * nvir:(i + 1) * nvir] = blk
    b *= math.sqrt(0.0081 / rows_per_site)
    occ = np.sort(rng.uniform(-1.2, -0.4, nocc))
    vir = np.sort(_loguniform(rng, 0.2, 4.0, nvir))
    centers = np.zeros((nocc, 3))
    centers[:, 0] = spacing * np.arange(nocc)
    prob = Problem(f"local chain {nocc}/{nvir}/{naux}", seed, b, np.concatenate([occ, vir]),
                   nocc, nvir, naux)
    return prob, centers


def local_fit_chain(seed: int = 1007, nocc: int = 8, nvir: int = 16, aux_per_site: int = 6,
                    spacing: float = 2.5):
    """(a, v, aux_xyz, centers, eps) for domain-local fits: occupied centres on a
    line (x = spacing * i bohr), aux functions spread around each site, an SPD
    metric v[P, Q] = exp(-(x_P - x_Q)^2 / 4) + 0.5 delta_PQ, and an unwhitened
    (P|ia) whose rows decay as exp(-|x_P - x_i| / spacing) away from centre i."""
    rng = np.random.default_rng(seed)
    naux = nocc * aux_per_site
    xs = np.repeat(spacing * np.arange(nocc), aux_per_site) + rng.uniform(-0.8, 0.8, naux)
    aux_xyz = np.zeros((naux, 3))
    aux_xyz[:, 0] = xs
    aux_xyz[:, 1] = rng.uniform(-0.3, 0.3, naux)
    centers = np.zeros((nocc, 3))
    centers[:, 0] = spacing * np.arange(nocc)
    dx = xs[:, None] - xs[None, :]
    v = np.exp(-dx * dx / 4.0) + 0.5 * np.eye(naux)
    a = rng.standard_normal((naux, nocc * nvir)) * 0.1
    for i in range(nocc):
        a[:, i * nvir:(i + 1) * nvir] *= np.exp(-np.abs(xs - centers[i, 0]) / spacing)[:, None]
    occ = np.sort(rng.uniform(-1.2, -0.4, nocc))
    vir = np.sort(_loguniform(rng, 0.2, 4.0, nvir))
    return a, v, aux_xyz, centers, np.concatenate([occ, vir])


@dataclass
class DomainSet:
    """Independent local-MP2 pair domains packed into flat buffers."""

    seed: int
    nocc: np.ndarray
    nvir: np.ndarray
    naux: np.ndarray
    b_offset: np.ndarray
    eps_offset: np.ndarray
    b_flat: np.ndarray      # concatenated row-major (naux_d, nocc_d*nvir_d) blocks
    eps_flat: np.ndarray    # concatenated [eps_occ_d..., eps_vir_d...]

    def __len__(self):
        return len(self.nocc)

    def domain(self, d: int) -> Problem:
        no, nv, na = int(self.nocc[d]), int(self.nvir[d]), int(self.naux[d])
        b = self.b_flat[self.b_offset[d]:self.b_offset[d] + na * no * nv].reshape(na, no * nv)
        e = self.eps_flat[self.eps_offset[d]:self.eps_offset[d] + no + nv]
        return Problem(f"domain {d}", self.seed, b, e, no, nv, na)

    @property
    def flops(self) -> float:
        return float(np.sum(self.nocc * (self.nocc + 1.0) * self.nvir ** 2 * self.naux))


def batched_domains(ndom: int = 512, seed: int = 1005) -> DomainSet:
    """Uneven pair domains: nocc_d uniform on {2..12}; nvir_d, naux_d
    log-uniform on [20, 240] and [80, 720]; B ~ N(0, 0.09/naux_d); occupied
    energies U[-1.2, -0.35] (the valence band of the decane-scale workload),
    virtuals log-uniform on [0.12, 5.0]."""
    rng = np.random.default_rng(seed)
    nocc = rng.integers(2, 13, ndom)
    nvir = np.rint(_loguniform(rng, 20, 240, ndom)).astype(np.int64)
    naux = np.rint(_loguniform(rng, 80, 720, ndom)).astype(np.int64)
    sizes = naux * nocc * nvir
    b_off = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
    e_off = np.concatenate([[0], np.cumsum(nocc + nvir)[:-1]]).astype(np.int64)
    b_flat = np.empty(int(sizes.sum()))
    eps_flat = np.empty(int((nocc + nvir).sum()))
    for d in range(ndom):
        no, nv, na = int(nocc[d]), int(nvir[d]), int(naux[d])
        b_flat[b_off[d]:b_off[d] + sizes[d]] = _b(rng, na, no, nv).ravel()
        occ = np.sort(rng.uniform(-1.2, -0.35, no))
        vir = np.sort(_loguniform(rng, 0.12, 5.0, nv))
        eps_flat[e_off[d]:e_off[d] + no + nv] = np.concatenate([occ, vir])
    return DomainSet(seed, nocc, nvir, naux, b_off, e_off, b_flat, eps_flat)
