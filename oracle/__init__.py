"""CPU oracle for the RI-MP2 pair-energy stage (test infrastructure only).

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg
import this package, and only as the checker or the timed CPU reference. The
product (``ferric_amd``) never imports it: there is no CPU fallback.

``rimp2_oracle.c`` restates the reference's algorithm loop for loop (file:line
citations in its header); this module is a thin ctypes binding plus two drivers:

* :func:`spin_components` -- the C oracle end to end (own blocked GEMM);
* :func:`spin_components_blas` -- the reference's CPU algorithm with the tail
  block ``G_i = B_i^T B_tail`` formed by OpenBLAS (numpy's BLAS, the library the
  reference links), one task per ``i`` on a thread pool with single-threaded
  BLAS inside tasks, the C pair loop, and the i-ordered serial fold
  (crates/ferric-mp2/src/rimp2.rs:1445-1544). This is the timed CPU baseline.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from concurrent.futures import ThreadPoolExecutor

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "build", "librimp2_oracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_sz = ctypes.c_size_t


def build(force: bool = False) -> str:
    """Compile the C oracle with gcc (``oracle/Makefile``)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-B" if force else "-s", "-C", _HERE], check=True,
                       stdout=subprocess.DEVNULL)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_active_occ.argtypes = [_sz, _sz, ctypes.POINTER(_sz)]
        L.orc_tail_block.argtypes = [_dp, _sz, _sz, _sz, _sz, _dp]
        L.orc_pair_energy.argtypes = [_dp, _sz, _sz, _sz, _sz, _dp, _sz, _sz, ctypes.c_int,
                                      ctypes.c_double, _dp, _dp]
        L.orc_i_partial.argtypes = [_dp, _sz, _sz, _sz, _dp, _sz, _sz, ctypes.c_int,
                                    ctypes.c_double, _dp, _dp]
        L.orc_spin_components_from_b_ov.argtypes = [_dp, _sz, _sz, _sz, _dp, _sz, _sz, _sz,
                                                    ctypes.c_int, ctypes.c_double,
                                                    ctypes.c_int, _dp]
        L.orc_spin_components_from_g.argtypes = [_dp, _sz, _sz, _dp, _sz, _sz, _dp]
        L.orc_t2_and_integrals.argtypes = [_dp, _sz, _sz, _sz, _dp, _sz, _sz, _dp, _dp]
        L.orc_x_ov_weighted.argtypes = [_dp, _dp, _sz, _sz, _sz, ctypes.c_double,
                                        ctypes.c_double, _dp]
        L.orc_mp2_density.argtypes = [_dp, _sz, _sz, _dp, _dp]
        L.orc_opposite_spin_density.argtypes = [_dp, _sz, _sz, _dp, _dp]
        L.orc_scaled_density.argtypes = [_dp, _sz, _sz, ctypes.c_double, ctypes.c_double,
                                         _dp, _dp]
        L.orc_rimp2_adjoint.argtypes = [_dp, _sz, _sz, _sz, _dp, _sz, _sz, _sz,
                                        ctypes.c_double, ctypes.c_double, _dp, _dp, _dp]
        L.orc_energy_bound.argtypes = [_dp, _sz, _sz, _sz, _dp, _sz, _sz, ctypes.c_int,
                                       ctypes.c_double, ctypes.c_double, ctypes.c_double, _dp]
        L.orc_energy_bound_row.argtypes = [_dp, _dp, _sz, _sz, _sz, _dp, _sz, _sz,
                                           ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                           ctypes.c_double, _dp]
        L.orc_energy_bound_finish.argtypes = [_dp, _sz, _sz, _dp]
        L.orc_u_same_spin.argtypes = [_dp, _sz, _sz, _sz, _dp, _sz, _sz, _sz, ctypes.c_int, _dp]
        L.orc_u_opposite_spin.argtypes = [_dp, _dp, _sz, _sz, _sz, _sz, _sz, _dp, _sz, _sz, _sz,
                                          _dp, _sz, _sz, _sz, ctypes.c_int, _dp]
        L.orc_signed_spin_components.argtypes = [_dp, _sz, _dp, _sz, _sz, _sz, _dp, _sz, _sz,
                                                 _sz, ctypes.c_int, _dp]
        L.orc_dress_b_ov.argtypes = [_dp, _sz, _sz, _dp, _sz, _dp, _sz, _dp, _sz, _dp]
        _lib = L
    return _lib


def _p(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


class OracleError(RuntimeError):
    pass


def active_occ(nocc_total: int, frozen_core: int) -> int:
    """rimp2.rs:99-106."""
    out = _sz(0)
    if lib().orc_active_occ(nocc_total, frozen_core, ctypes.byref(out)):
        raise OracleError(f"frozen_core = {frozen_core} freezes all {nocc_total} occupied orbitals")
    return out.value


def _kappa_args(kappa):
    return (0, 0.0) if kappa is None else (1, float(kappa))


def spin_components(b_ov, eps, nocc, nvir, first_occ=0, nocc_total=None, kappa=None,
                    nthreads: int = 0):
    """(e_os, e_ss, e_total) from the C oracle (rimp2.rs:1445-1544)."""
    b = _f64(b_ov)
    e = _f64(eps)
    if nocc_total is None:
        nocc_total = first_occ + nocc
    naux = b.shape[0] if b.ndim == 2 else 0
    out = np.zeros(3)
    hk, k = _kappa_args(kappa)
    rc = lib().orc_spin_components_from_b_ov(_p(b), naux, nocc, nvir, _p(e), e.size, first_occ,
                                             nocc_total, hk, k, nthreads, _p(out))
    if rc:
        raise OracleError(f"oracle error {rc}")
    return tuple(float(x) for x in out)


def tail_block(b_ov, nocc, nvir, i):
    b = _f64(b_ov)
    g = np.zeros((nvir, (nocc - i) * nvir))
    lib().orc_tail_block(_p(b), b.shape[0], nocc, nvir, i, _p(g))
    return g


def i_partial(g_i, i, nocc, nvir, eps, first_occ, nocc_total, kappa=None):
    g = _f64(g_i)
    e = _f64(eps)
    os_ = ctypes.c_double(0.0)
    ss_ = ctypes.c_double(0.0)
    hk, k = _kappa_args(kappa)
    lib().orc_i_partial(_p(g), i, nocc, nvir, _p(e), first_occ, nocc_total, hk, k,
                        ctypes.byref(os_), ctypes.byref(ss_))
    return os_.value, ss_.value


def spin_components_blas(b_ov, eps, nocc, nvir, first_occ=0, nocc_total=None, kappa=None,
                         nthreads: int | None = None, i_rows=None):
    """The reference's CPU algorithm with OpenBLAS tail GEMMs (the timed baseline).

    ``i_rows`` restricts the work to a subset of occupied rows (a bounded sample
    for timing); the default is every row, which gives the full energy.
    Returns (e_os, e_ss, e_total) over the rows done.
    """
    from threadpoolctl import threadpool_limits

    b = _f64(b_ov)
    e = _f64(eps)
    if nocc_total is None:
        nocc_total = first_occ + nocc
    if nthreads is None:
        nthreads = len(os.sched_getaffinity(0))
    rows = list(range(nocc)) if i_rows is None else list(i_rows)
    parts = {}

    def task(i):
        bi = b[:, i * nvir:(i + 1) * nvir]
        tail = b[:, i * nvir:]
        g = np.ascontiguousarray(bi.T @ tail)
        parts[i] = i_partial(g, i, nocc, nvir, e, first_occ, nocc_total, kappa)

    with threadpool_limits(limits=1, user_api="blas"):
        with ThreadPoolExecutor(max_workers=nthreads) as ex:
            list(ex.map(task, rows))
    os_ = ss_ = 0.0
    for i in sorted(parts):
        os_ += parts[i][0]
        ss_ += parts[i][1]
    return os_, ss_, os_ + ss_


def spin_components_from_g(g, eps, nocc, nvir, first_occ=0, nocc_total=None):
    """rimp2.rs:984-1018."""
    gg = _f64(g)
    e = _f64(eps)
    if nocc_total is None:
        nocc_total = first_occ + nocc
    out = np.zeros(3)
    lib().orc_spin_components_from_g(_p(gg), nocc, nvir, _p(e), first_occ, nocc_total, _p(out))
    return tuple(float(x) for x in out)


def t2_and_integrals(b_ov, eps, nocc, nvir, first_occ=0, nocc_total=None):
    b = _f64(b_ov)
    e = _f64(eps)
    if nocc_total is None:
        nocc_total = first_occ + nocc
    nov = nocc * nvir
    t2 = np.zeros((nov, nov))
    eri = np.zeros((nov, nov))
    lib().orc_t2_and_integrals(_p(b), b.shape[0], nocc, nvir, _p(e), nocc_total, first_occ,
                               _p(t2), _p(eri))
    return t2, eri


def x_ov_weighted(t2, b_ov, nocc, nvir, c_os=1.0, c_ss=1.0):
    t = _f64(t2)
    b = _f64(b_ov)
    x = np.zeros_like(b)
    rc = lib().orc_x_ov_weighted(_p(t), _p(b), nocc, nvir, b.shape[0], c_os, c_ss, _p(x))
    if rc:
        raise OracleError(f"oracle error {rc}")
    return x


def mp2_density(t2, nocc, nvir, c_os=1.0, c_ss=1.0):
    t = _f64(t2)
    p_oo = np.zeros((nocc, nocc))
    p_vv = np.zeros((nvir, nvir))
    lib().orc_scaled_density(_p(t), nocc, nvir, c_os, c_ss, _p(p_oo), _p(p_vv))
    return p_oo, p_vv


def mp2_density_blas(b_ov, eps, nocc, nvir, first_occ=0, nocc_total=None, c_os=1.0, c_ss=1.0):
    """build_mp2_density (oo_rimp2.rs:1192-1300) with the spin scaling of
    rimp2.rs:1917-2052, as numpy BLAS contractions over the dense amplitudes
    t[i, a, j, b] = (ia|jb) / D (the C oracle's orc_t2_and_integrals and
    orc_scaled_density restated for sizes where its plain loops are slow):
    P_oo = -sum_kab t_ik^ab T~_jk^ab, P_vv = sum_ijc t_ij^ac T~_ij^bc with
    T~ = (c_os + c_ss) t - c_ss t^T (t^T swapping a and b)."""
    b = _f64(b_ov)
    e = _f64(eps)
    if nocc_total is None:
        nocc_total = first_occ + nocc
    eo = e[first_occ:first_occ + nocc]
    ev = e[nocc_total:nocc_total + nvir]
    g = (b.T @ b).reshape(nocc, nvir, nocc, nvir)
    d = eo[:, None, None, None] + eo[None, None, :, None] - ev[None, :, None, None] \
        - ev[None, None, None, :]
    t = g / d
    del g, d
    tt = (c_os + c_ss) * t - c_ss * t.transpose(0, 3, 2, 1)
    # P_oo[i, j] = -sum_{k,a,b} t[i,a,k,b] tt[j,a,k,b]
    p_oo = -(t.reshape(nocc, -1) @ tt.reshape(nocc, -1).T)
    # P_vv[a, b] = sum_{i,j,c} t[i,a,j,c] tt[i,b,j,c]
    ta = np.ascontiguousarray(t.transpose(1, 0, 2, 3)).reshape(nvir, -1)
    tb = np.ascontiguousarray(tt.transpose(1, 0, 2, 3)).reshape(nvir, -1)
    p_vv = ta @ tb.T
    return p_oo, p_vv


def adjoint(b_ov, eps, nocc, nvir, first_occ=0, nocc_total=None, c_os=1.0, c_ss=1.0):
    """(E_os, E_ss, E, Gamma = dE/dB, dE/deps) through the reference's intermediates."""
    b = _f64(b_ov)
    e = _f64(eps)
    if nocc_total is None:
        nocc_total = first_occ + nocc
    gamma = np.zeros_like(b)
    deps = np.zeros_like(e)
    out = np.zeros(3)
    rc = lib().orc_rimp2_adjoint(_p(b), b.shape[0], nocc, nvir, _p(e), e.size, first_occ,
                                 nocc_total, c_os, c_ss, _p(gamma), _p(deps), _p(out))
    if rc:
        raise OracleError(f"oracle error {rc}")
    return out[0], out[1], out[2], gamma, deps


def energy_bound(b_ov, eps, nocc, nvir, first_occ=0, nocc_total=None, kappa=None,
                 eps_g=None, eta=0.0):
    """Worst-case |dE_os|, |dE_ss| for per-element G perturbations eps_g*S_ab + eta.

    Default eps_g is the f64 device-vs-CPU factor 2*gamma_naux(u64).
    """
    b = _f64(b_ov)
    e = _f64(eps)
    if nocc_total is None:
        nocc_total = first_occ + nocc
    naux = b.shape[0]
    if eps_g is None:
        u = 2.0 ** -53
        eps_g = 2.0 * naux * u / (1.0 - naux * u)
    out = np.zeros(2)
    hk, k = _kappa_args(kappa)
    lib().orc_energy_bound(_p(b), naux, nocc, nvir, _p(e), first_occ, nocc_total, hk, k,
                           float(eps_g), float(eta), _p(out))
    return float(out[0]), float(out[1])


def row_partials_blas(b_ov, eps, nocc, nvir, rows, first_occ=0, nocc_total=None, kappa=None,
                      eps_g=None, eta=0.0, bounds=True):
    """Per occupied row i in ``rows``: the oracle's partial (e_os_i, e_ss_i) of
    spin_components_from_blocks (rimp2.rs:1520-1530) from a BLAS tail block
    G_i = B_i^T B_tail, and the energy bound of that row alone (the bound of
    ``energy_bound`` restricted to the row's pairs, with |B|^T|B| as S_ab).

    Returns {i: (e_os_i, e_ss_i, bound_os_i, bound_ss_i)}. Rows are computed one
    at a time, so the memory is two (nvir, (nocc - i) nvir) blocks."""
    b = _f64(b_ov)
    e = _f64(eps)
    if nocc_total is None:
        nocc_total = first_occ + nocc
    naux = b.shape[0]
    if eps_g is None:
        u = 2.0 ** -53
        eps_g = 2.0 * naux * u / (1.0 - naux * u)
    hk, k = _kappa_args(kappa)
    out = {}
    for i in rows:
        bi = b[:, i * nvir:(i + 1) * nvir]
        tail = b[:, i * nvir:]
        g = np.ascontiguousarray(bi.T @ tail)
        eos, ess = i_partial(g, i, nocc, nvir, e, first_occ, nocc_total, kappa)
        bos = bss = float("nan")
        if bounds:
            s = np.ascontiguousarray(np.abs(bi).T @ np.abs(tail))
            acc = np.zeros(4)
            lib().orc_energy_bound_row(_p(g), _p(s), i, nocc, nvir, _p(e), first_occ, nocc_total,
                                       hk, k, float(eps_g), float(eta), _p(acc))
            bnd = np.zeros(2)
            lib().orc_energy_bound_finish(_p(acc), nocc, nvir, _p(bnd))
            bos, bss = float(bnd[0]), float(bnd[1])
            del s
        out[int(i)] = (eos, ess, bos, bss)
        del g
    return out


def u_same_spin(b_ov, eps, nocc, nvir, first_occ=0, nocc_total=None, nthreads: int = 0):
    """Same-spin energy of one spin channel (u_rimp2.rs:693-729, 766-797)."""
    b = _f64(b_ov)
    e = _f64(eps)
    if nocc_total is None:
        nocc_total = first_occ + nocc
    out = np.zeros(1)
    rc = lib().orc_u_same_spin(_p(b), b.shape[0], nocc, nvir, _p(e), e.size, first_occ,
                               nocc_total, nthreads, _p(out))
    if rc:
        raise OracleError(f"oracle error {rc}")
    return float(out[0])


def u_opposite_spin(ch_a, ch_b, nthreads: int = 0):
    """Opposite-spin energy of two channels (u_rimp2.rs:731-760, 978-995).

    Each channel is (b_ov, eps, nocc, nvir, first_occ, nocc_total), both fitted
    in one aux basis (same row count)."""
    ba, ea, noa, nva, foa, nta = ch_a
    bb, eb, nob, nvb, fob, ntb = ch_b
    ba, ea, bb, eb = _f64(ba), _f64(ea), _f64(bb), _f64(eb)
    if ba.shape[0] != bb.shape[0]:
        raise OracleError("the two channels must share the aux basis")
    out = np.zeros(1)
    rc = lib().orc_u_opposite_spin(_p(ba), _p(bb), ba.shape[0], noa, nva, nob, nvb, _p(ea),
                                   ea.size, foa, nta, _p(eb), eb.size, fob, ntb, nthreads,
                                   _p(out))
    if rc:
        raise OracleError(f"oracle error {rc}")
    return float(out[0])


def signed_spin_components(pos, neg, eps, nocc, nvir, first_occ=0, nocc_total=None,
                           nthreads: int = 0):
    """(e_os, e_ss, e_total) of a signed factor K = pos^T pos - neg^T neg
    (rimp2.rs:1559-1607)."""
    bp, bn, e = _f64(pos), _f64(neg), _f64(eps)
    if nocc_total is None:
        nocc_total = first_occ + nocc
    out = np.zeros(3)
    rc = lib().orc_signed_spin_components(_p(bp), bp.shape[0], _p(bn), bn.shape[0], nocc, nvir,
                                          _p(e), e.size, first_occ, nocc_total, nthreads,
                                          _p(out))
    if rc:
        raise OracleError(f"oracle error {rc}")
    return tuple(float(x) for x in out)


def dress_b_ov(j3, v_inv_sqrt, c_left, c_right, output_band=None):
    """B = M (C_left^T (Q|mu nu) C_right) of the output band (rimp2.rs:648-756).

    j3 (naux, nao, nao); v_inv_sqrt (naux, naux); c_left (nao, nleft);
    c_right (nao, nright). Returns (band, nleft*nright)."""
    a = _f64(j3)
    naux, nao = a.shape[0], a.shape[1]
    p0, p1 = output_band if output_band is not None else (0, naux)
    m = _f64(np.asarray(v_inv_sqrt)[p0:p1])
    cl, cr = _f64(c_left), _f64(c_right)
    out = np.zeros((p1 - p0, cl.shape[1] * cr.shape[1]))
    rc = lib().orc_dress_b_ov(_p(a), naux, nao, _p(cl), cl.shape[1], _p(cr), cr.shape[1], _p(m),
                              p1 - p0, _p(out))
    if rc:
        raise OracleError(f"oracle error {rc}")
    return out


def pair_energies(b_ov, eps, nocc, nvir, first_occ=0, nocc_total=None):
    """estimate_pair_energies_from_b_ov (crates/ferric-mp2/src/dlpno_mp2.rs:527-594)
    split into its parts: (e_os, e_ss), each nocc x nocc symmetric with
    e_os[i, j] = fac sum_ab g_ab^2 / D and e_ss[i, j] = fac sum_ab g_ab (g_ab - g_ba) / D,
    g = B_i^T B_j, fac = 1 (i == j) or 2; their sum is the reference's e_ij. A
    non-negative denominator raises ValueError, as the reference refuses it.
    numpy, one nvir x nvir block per pair."""
    b = _f64(b_ov)
    e = _f64(eps)
    if nocc_total is None:
        nocc_total = first_occ + nocc
    eo = e[first_occ:first_occ + nocc]
    ev = e[nocc_total:nocc_total + nvir]
    e_os = np.zeros((nocc, nocc))
    e_ss = np.zeros((nocc, nocc))
    for i in range(nocc):
        bi = b[:, i * nvir:(i + 1) * nvir]
        for j in range(i, nocc):
            g = bi.T @ b[:, j * nvir:(j + 1) * nvir]
            d = eo[i] + eo[j] - ev[:, None] - ev[None, :]
            if np.any(d >= 0.0):
                raise ValueError(f"non-negative MP2 denominator for pair ({i}, {j})")
            fac = 1.0 if i == j else 2.0
            e_os[i, j] = e_os[j, i] = fac * float(np.sum(g * g / d))
            e_ss[i, j] = e_ss[j, i] = fac * float(np.sum(g * (g - g.T) / d))
    return e_os, e_ss


def pair_energies_from_g(g, eps, nocc, nvir, first_occ=0, nocc_total=None):
    """estimate_pair_energies (crates/ferric-mp2/src/pair_energy_screen.rs:125-190)
    split into (e_os, e_ss) like pair_energies, from an explicit dense g:
    (ia|jb) = g[i nvir + a, j nvir + b], (ib|ja) = g[i nvir + b, j nvir + a].
    A non-negative denominator raises ValueError."""
    gg = _f64(g)
    e = _f64(eps)
    if nocc_total is None:
        nocc_total = first_occ + nocc
    eo = e[first_occ:first_occ + nocc]
    ev = e[nocc_total:nocc_total + nvir]
    e_os = np.zeros((nocc, nocc))
    e_ss = np.zeros((nocc, nocc))
    for i in range(nocc):
        for j in range(i, nocc):
            blk = gg[i * nvir:(i + 1) * nvir, j * nvir:(j + 1) * nvir]
            d = eo[i] + eo[j] - ev[:, None] - ev[None, :]
            if np.any(d >= 0.0):
                raise ValueError(f"non-negative MP2 denominator for pair ({i}, {j})")
            fac = 1.0 if i == j else 2.0
            e_os[i, j] = e_os[j, i] = fac * float(np.sum(blk * blk / d))
            e_ss[i, j] = e_ss[j, i] = fac * float(np.sum(blk * (blk - blk.T) / d))
    return e_os, e_ss


def pno_pair_energies(b_ov, eps, nocc, nvir, first_occ, nocc_total, pairs, t_cut_pno, g=None):
    """The pair loop of dlpno_mp2_from_b_ov (crates/ferric-mp2/src/dlpno_mp2.rs:718-744)
    per pair: amplitudes T = g / D (:700-712), pno_for_pair (local_pno.rs:132-200:
    D = sym(T T^T + T^T T), eigh, keep |n| >= t_cut_pno in descending order, never
    none), pair_energy_in_pno_basis (dlpno_mp2.rs:249-340: Q^T g Q, the PNO Fock
    block re-diagonalised, the energy with the semicanonical denominators).
    Returns (e_os, e_ss, kept, discarded, kept_occupation) arrays over pairs
    (the last, the sum of the kept occupations, is for the trace identity of
    the tests); numpy eigh. With ``g`` (a dense (nocc nvir)^2 matrix) each
    pair's block is g[i nvir:(i+1) nvir, j nvir:(j+1) nvir] instead of
    B_i^T B_j (the dense-g path, dlpno_mp2.rs:354-440); b_ov is then unused."""
    b = _f64(b_ov) if g is None else None
    gg = None if g is None else _f64(g)
    e = _f64(eps)
    pairs = [tuple(int(x) for x in p) for p in pairs]
    n = len(pairs)
    e_os, e_ss = np.zeros(n), np.zeros(n)
    kept, disc, kocc = np.zeros(n, dtype=np.int64), np.zeros(n), np.zeros(n)
    ev = e[nocc_total:nocc_total + nvir]
    for k, (i, j) in enumerate(pairs):
        if gg is None:
            g = b[:, i * nvir:(i + 1) * nvir].T @ b[:, j * nvir:(j + 1) * nvir]
        else:
            g = gg[i * nvir:(i + 1) * nvir, j * nvir:(j + 1) * nvir]
        eij = e[first_occ + i] + e[first_occ + j]
        t = g / (eij - ev[:, None] - ev[None, :])
        d = t @ t.T + t.T @ t
        d = 0.5 * (d + d.T)
        occ, vec = np.linalg.eigh(d)
        order = sorted(range(nvir), key=lambda x: -occ[x])  # stable, descending
        keep = [x for x in order if abs(occ[x]) >= t_cut_pno] or [order[0]]
        disc[k] = sum(abs(occ[x]) for x in order if x not in keep)
        kept[k] = len(keep)
        kocc[k] = sum(occ[x] for x in keep)
        q = vec[:, keep]
        tij, tji = q.T @ g @ q, q.T @ g.T @ q
        f = q.T @ (ev[:, None] * q)
        e_pno, u = np.linalg.eigh(f)
        tij, tji = u.T @ tij @ u, u.T @ tji @ u
        dd = eij - e_pno[:, None] - e_pno[None, :]
        fac = 1.0 if i == j else 2.0
        e_os[k] = fac * float(np.sum(tij * tij / dd))
        e_ss[k] = fac * float(np.sum(tij * (tij - tji) / dd))
    return e_os, e_ss, kept, disc, kocc


def domain_fit_pair(a, v, aux_xyz, occ_centers, i, j, nvir, radius):
    """domain_fit_pair (crates/ferric-mp2/src/lmp2_amplitude.rs:1707-1758): the
    aux domain D = {P : |x_P - c_i|^2 <= r^2 or |x_P - c_j|^2 <= r^2} in aux
    order, V_DD inverted through its Cholesky factor (ndarray-linalg invc), then
    A_{i,D}^T (V_DD^-1 A_{j,D}). Raises ValueError on an empty domain."""
    import scipy.linalg

    a = _f64(a)
    v = _f64(v)
    x = _f64(aux_xyz)
    c = _f64(occ_centers)
    r2 = radius * radius
    d2i = np.sum((x - c[i]) ** 2, axis=1)
    d2j = np.sum((x - c[j]) ** 2, axis=1)
    dom = [p for p in range(x.shape[0]) if d2i[p] <= r2 or d2j[p] <= r2]
    if not dom:
        raise ValueError(f"empty aux domain for pair ({i},{j}) at radius {radius} bohr")
    vdd = v[np.ix_(dom, dom)]
    cf = scipy.linalg.cho_factor(vdd, lower=True)
    vinv = scipy.linalg.cho_solve(cf, np.eye(len(dom)))
    a_i = a[dom, i * nvir:(i + 1) * nvir]
    a_j = a[dom, j * nvir:(j + 1) * nvir]
    return a_i.T @ (vinv @ a_j), len(dom)
