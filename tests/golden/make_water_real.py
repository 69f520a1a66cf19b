"""Real-water RI-MP2 inputs from this project's own integrals (fixture generator).

Writes tests/golden/water_ccpvdz_real.npz: the dressed three-index tensor
B^Q_ia (cc-pVDZ / cc-pVDZ-RI, RHF, no frozen core) and the orbital energies of
water at the geometry of the reference's testdata/molecules/water.xyz, so that
the GPU kernel can be checked against the reference's published RI-MP2
correlation energy (testdata/reference/h2o_cc-pvdz_rimp2.json). Run in the
build container, where the reference checkout is readable; the fixture, not
this script, is what the tests load.

Method (this file's own code):
  * basis sets: the Basis Set Exchange JSON files bundled with the reference
    (crates/ferric-core/src/basis/bundled/{cc-pvdz,cc-pvdz-ri}.json), read as
    data; pure (spherical) d and f shells;
  * integrals: McMurchie-Davidson Hermite expansions (overlap, kinetic,
    nuclear attraction, exact four-centre ERIs for the SCF, three- and two-centre
    Coulomb integrals for the fit), Boys function from the regularised
    incomplete gamma function;
  * RHF: core-Hamiltonian guess, DIIS, converged to 1e-12 Eh and 1e-10 in the
    orbital gradient;
  * RI: (P|Q) = L L^T (Cholesky), B = L^-1 (P|ia) -- the reference's dressing
    (rimp2.rs SignedBOv doc, "dressed by the Cholesky factor"). For the
    distorted geometry the fixture also keeps (Q|mu nu), the MO coefficients
    and (P|Q), the inputs of the dressing / MO-transform stage
    (stream_dressed_mo_band, rimp2.rs:648-756).

The equilibrium water also in conventional (unfitted) MP2: B is then an exact
factor of the four-centre (ia|jb) matrix, and the energy must meet the
reference's testdata/reference/h2o_cc-pvdz_mp2.json (PySCF).

Three molecules. H2 at 0.74 Angstrom, the reference's inline known-answer test
(rimp2.rs:3187-3201: RI-MP2 within 1e-4 of the canonical MP2 -0.0263715576).
The equilibrium water once more with a density-fitted SCF (J and K fitted in
def2-universal-jkfit, the reference's scf_jk = "rijk"; known answer
testdata/reference/h2o_cc-pvdz_rijk-def2-universal-jkfit_dfmp2.json, PySCF).
Two water geometries: the equilibrium water of testdata/molecules/water.xyz (known
answer testdata/reference/h2o_cc-pvdz_rimp2.json) and the distorted water of
testdata/molecules/validation/h2o_distorted.xyz, whose ORCA 6.1.1 and PySCF
2.13.1 results agree to 5e-11 Eh (testdata/reference/validation/
{rimp2_gradient,scs_mp2_gradient}/h2o_distorted_cc-pvdz.json: E_RHF, E_corr,
E_os / E_ss with and without one frozen core orbital). Before anything is
written, the nuclear repulsion and the RHF energies are checked against those
files (|dE| <= 1e-8 Eh), which tests the integrals, geometry and units end to
end. Energies do not depend on the spherical-harmonic convention or
MO phases, only on the spanned spaces, so B is one valid factor among many.

    python tests/golden/make_water_real.py
"""
from __future__ import annotations

import json
import math
import os
import sys

import numpy as np
from scipy.special import gamma as gamma_fn
from scipy.special import gammainc

REF = "/root/reference"
BASIS_DIR = os.path.join(REF, "crates", "ferric-core", "src", "basis", "bundled")
XYZ = os.path.join(REF, "testdata", "molecules", "water.xyz")
XYZ_DISTORTED = os.path.join(REF, "testdata", "molecules", "validation", "h2o_distorted.xyz")
REF_JSON = os.path.join(REF, "testdata", "reference", "h2o_cc-pvdz_rimp2.json")
REF_GRAD = os.path.join(REF, "testdata", "reference", "validation", "rimp2_gradient",
                        "h2o_distorted_cc-pvdz.json")
REF_SCS = os.path.join(REF, "testdata", "reference", "validation", "scs_mp2_gradient",
                       "h2o_distorted_cc-pvdz.json")
REF_RIJK = os.path.join(REF, "testdata", "reference",
                        "h2o_cc-pvdz_rijk-def2-universal-jkfit_dfmp2.json")
REF_CONV = os.path.join(REF, "testdata", "reference", "h2o_cc-pvdz_mp2.json")
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "water_ccpvdz_real.npz")
BOHR_ANGSTROM = 0.52917721092   # CODATA 2010, the reference's Å <-> Bohr constant
Z = {"H": 1, "O": 8}


# ---- basis -----------------------------------------------------------------------
def cart_components(l):
    return [(lx, ly, l - lx - ly) for lx in range(l, -1, -1) for ly in range(l - lx, -1, -1)]


def spherical_transform(l):
    """Rows: pure real solid harmonics of degree l as combinations of the
    Cartesian monomials of cart_components(l) (unnormalised; any basis of the
    pure space gives the same energies)."""
    comps = cart_components(l)
    idx = {c: k for k, c in enumerate(comps)}
    if l < 2:
        return np.eye(len(comps))
    if l == 2:
        rows = [{(1, 1, 0): 1}, {(0, 1, 1): 1}, {(1, 0, 1): 1}, {(2, 0, 0): 1, (0, 2, 0): -1},
                {(0, 0, 2): 2, (2, 0, 0): -1, (0, 2, 0): -1}]
    elif l == 3:
        rows = [{(0, 0, 3): 2, (2, 0, 1): -3, (0, 2, 1): -3},
                {(1, 0, 2): 4, (3, 0, 0): -1, (1, 2, 0): -1},
                {(0, 1, 2): 4, (2, 1, 0): -1, (0, 3, 0): -1},
                {(2, 0, 1): 1, (0, 2, 1): -1},
                {(1, 1, 1): 1},
                {(3, 0, 0): 1, (1, 2, 0): -3},
                {(2, 1, 0): 3, (0, 3, 0): -1}]
    else:
        # higher l: the harmonic polynomials of degree l, the null space of the
        # Laplacian from degree l to degree l - 2 (any basis of it will do)
        low = cart_components(l - 2)
        lidx = {c: k for k, c in enumerate(low)}
        lap = np.zeros((len(low), len(comps)))
        for k, (a, b, c) in enumerate(comps):
            for (n, t) in ((a, (a - 2, b, c)), (b, (a, b - 2, c)), (c, (a, b, c - 2))):
                if n >= 2:
                    lap[lidx[t], k] += n * (n - 1)
        _, sv, vt = np.linalg.svd(lap)
        return vt[len(low):]
    m = np.zeros((len(rows), len(comps)))
    for r, d in enumerate(rows):
        for c, v in d.items():
            m[r, idx[c]] = v
    return m


def dfact(n):
    return 1.0 if n <= 0 else float(np.prod(np.arange(n, 0, -2)))


def prim_norm(a, l):
    return (2 * a / math.pi) ** 0.75 * (4 * a) ** (l / 2) / math.sqrt(dfact(2 * l - 1))


class Shell:
    def __init__(self, l, center, exps, coefs):
        self.l = l
        self.center = np.asarray(center, float)
        self.exps = np.asarray(exps, float)
        self.coefs = np.asarray(coefs, float).reshape(len(exps), -1)   # (nprim, ncontr)
        self.coefs = self.coefs * np.array([prim_norm(a, l) for a in self.exps])[:, None]
        self.ncart = len(cart_components(l))
        self.sph = spherical_transform(l)

    @property
    def nfunc(self):
        return self.coefs.shape[1] * self.sph.shape[0]


def load_basis(name, atoms):
    data = json.load(open(os.path.join(BASIS_DIR, name + ".json")))
    shells = []
    for sym, xyz in atoms:
        for sh in data["elements"][str(Z[sym])]["electron_shells"]:
            exps = [float(x) for x in sh["exponents"]]
            coefs = np.array([[float(c) for c in row] for row in sh["coefficients"]]).T
            for l in sh["angular_momentum"]:
                shells.append(Shell(l, xyz, exps, coefs))
    return shells


# ---- McMurchie-Davidson ---------------------------------------------------------------
def hermite_e(la, lb, a, b, xa, xb):
    """E[i][j][t] over primitive pairs (arrays), 1-D, including exp(-mu X^2)."""
    p = a + b
    mu = a * b / p
    xab = xa - xb
    xpa = (a * xa + b * xb) / p - xa
    xpb = (a * xa + b * xb) / p - xb
    E = [[[None] * (la + lb + 2) for _ in range(lb + 1)] for _ in range(la + 1)]
    zero = np.zeros_like(p)
    E[0][0][0] = np.exp(-mu * xab * xab)

    def g(i, j, t):
        if t < 0 or t > i + j or i < 0 or j < 0:
            return zero
        return E[i][j][t]

    for i in range(la + 1):
        for j in range(lb + 1):
            if i == 0 and j == 0:
                continue
            for t in range(i + j + 1):
                if i > 0:
                    v = g(i - 1, j, t - 1) / (2 * p) + xpa * g(i - 1, j, t) + (t + 1) * g(i - 1, j, t + 1)
                else:
                    v = g(i, j - 1, t - 1) / (2 * p) + xpb * g(i, j - 1, t) + (t + 1) * g(i, j - 1, t + 1)
                E[i][j][t] = v
    return E


def boys(nmax, T):
    T = np.asarray(T, float)
    out = np.empty((nmax + 1,) + T.shape)
    small = T < 1e-10
    Ts = np.where(small, 1.0, T)
    for n in range(nmax + 1):
        v = gamma_fn(n + 0.5) * gammainc(n + 0.5, Ts) / (2.0 * Ts ** (n + 0.5))
        out[n] = np.where(small, 1.0 / (2 * n + 1) - T / (2 * n + 3), v)
    return out


def hermite_r(L, alpha, X, Y, Zc):
    """R[t,u,v] (t+u+v <= L) over arrays, the Hermite Coulomb integrals R^0_tuv."""
    T = alpha * (X * X + Y * Y + Zc * Zc)
    F = boys(L, T)
    Rn = {}
    for n in range(L + 1):
        Rn[(0, 0, 0, n)] = (-2 * alpha) ** n * F[n]
    # build up t, then u, then v with the n index descending
    def R(t, u, v, n):
        key = (t, u, v, n)
        if key in Rn:
            return Rn[key]
        if t < 0 or u < 0 or v < 0:
            return 0.0
        if t > 0:
            val = (t - 1) * R(t - 2, u, v, n + 1) + X * R(t - 1, u, v, n + 1)
        elif u > 0:
            val = (u - 1) * R(t, u - 2, v, n + 1) + Y * R(t, u - 1, v, n + 1)
        else:
            val = (v - 1) * R(t, u, v - 2, n + 1) + Zc * R(t, u, v - 1, n + 1)
        Rn[key] = val
        return val

    out = {}
    for t in range(L + 1):
        for u in range(L + 1 - t):
            for v in range(L + 1 - t - u):
                out[(t, u, v)] = R(t, u, v, 0)
    return out


class Pair:
    """Primitive-pair data of two shells (or of one shell with a unit s-function)."""

    def __init__(self, sa, sb=None):
        a = sa.exps[:, None]
        if sb is None:
            b = np.zeros((1,))[None, :]
            cb = np.ones((1, 1))
            lb, B = 0, sa.center
        else:
            b = sb.exps[None, :]
            cb = sb.coefs
            lb, B = sb.l, sb.center
        A = sa.center
        self.la, self.lb = sa.l, lb
        self.p = (a + b).ravel()
        self.P = np.stack([((a * A[k] + b * B[k]) / (a + b)).ravel() for k in range(3)])
        self.E = [hermite_e(sa.l, lb, np.repeat(a.ravel(), b.size), np.tile(b.ravel(), a.size),
                            A[k], B[k]) for k in range(3)]
        # contraction weights (nprim_a*nprim_b, nca, ncb)
        self.w = np.einsum("ix,jy->ijxy", sa.coefs, cb).reshape(a.size * b.size, sa.coefs.shape[1],
                                                                 cb.shape[1])
        self.ca = cart_components(sa.l)
        self.cb = cart_components(lb)

    def hermite_coef(self):
        """coef[(ca, cb)] -> dict (t,u,v) -> array over primitive pairs."""
        out = {}
        for ia, (ax, ay, az) in enumerate(self.ca):
            for ib, (bx, by, bz) in enumerate(self.cb):
                d = {}
                for t in range(ax + bx + 1):
                    et = self.E[0][ax][bx][t]
                    for u in range(ay + by + 1):
                        eu = self.E[1][ay][by][u]
                        for v in range(az + bz + 1):
                            d[(t, u, v)] = et * eu * self.E[2][az][bz][v]
                out[(ia, ib)] = d
        return out


def two_electron(bra: Pair, ket: Pair):
    """(ab|cd) over Cartesian components and contracted functions:
    array (nca, ncb, ncc, ncd, fa, fb, fc, fd)."""
    p, q = bra.p[:, None], ket.p[None, :]
    alpha = p * q / (p + q)
    pref = 2 * math.pi ** 2.5 / (p * q * np.sqrt(p + q))
    X = bra.P[0][:, None] - ket.P[0][None, :]
    Y = bra.P[1][:, None] - ket.P[1][None, :]
    Zc = bra.P[2][:, None] - ket.P[2][None, :]
    L = bra.la + bra.lb + ket.la + ket.lb
    R = hermite_r(L, alpha, X, Y, Zc)
    hb, hk = bra.hermite_coef(), ket.hermite_coef()
    nca, ncb, ncc, ncd = len(bra.ca), len(bra.cb), len(ket.ca), len(ket.cb)
    fa, fb = bra.w.shape[1:]
    fc, fd = ket.w.shape[1:]
    out = np.zeros((nca, ncb, ncc, ncd, fa, fb, fc, fd))
    for (ia, ib), db in hb.items():
        for (ic, id_), dk in hk.items():
            acc = np.zeros_like(alpha)
            for (t, u, v), eb in db.items():
                for (tt, uu, vv), ek in dk.items():
                    s = -1.0 if (tt + uu + vv) % 2 else 1.0
                    acc = acc + s * np.outer(eb, ek) * R[(t + tt, u + uu, v + vv)]
            prim = pref * acc
            out[ia, ib, ic, id_] = np.einsum("pq,pab,qcd->abcd", prim, bra.w, ket.w)
    return out


def one_electron(sa, sb, atoms_bohr, charges):
    """(S, T, V) blocks over Cartesian components and contracted functions."""
    pr = Pair(sa, sb)
    a = np.repeat(sa.exps, sb.exps.size)
    b = np.tile(sb.exps, sa.exps.size)
    p = pr.p
    # overlap pieces with the j index raised by up to 2 for the kinetic term
    Ebig = [hermite_e(sa.l, sb.l + 2, a, b, sa.center[k], sb.center[k]) for k in range(3)]
    sq = (math.pi / p) ** 0.5

    def s1(k, i, j):
        if j < 0:
            return np.zeros_like(p)
        return Ebig[k][i][j][0] * sq

    nca, ncb = len(pr.ca), len(pr.cb)
    fa, fb = pr.w.shape[1:]
    S = np.zeros((nca, ncb, fa, fb))
    T = np.zeros_like(S)
    V = np.zeros_like(S)
    for ia, ca in enumerate(pr.ca):
        for ib, cb in enumerate(pr.cb):
            sx = [s1(k, ca[k], cb[k]) for k in range(3)]
            dx = [cb[k] * (cb[k] - 1) * s1(k, ca[k], cb[k] - 2)
                  - 2 * b * (2 * cb[k] + 1) * s1(k, ca[k], cb[k])
                  + 4 * b * b * s1(k, ca[k], cb[k] + 2) for k in range(3)]
            s = sx[0] * sx[1] * sx[2]
            t = -0.5 * (dx[0] * sx[1] * sx[2] + sx[0] * dx[1] * sx[2] + sx[0] * sx[1] * dx[2])
            S[ia, ib] = np.einsum("p,pab->ab", s, pr.w)
            T[ia, ib] = np.einsum("p,pab->ab", t, pr.w)
    L = sa.l + sb.l
    hc = pr.hermite_coef()
    for C, zc in zip(atoms_bohr, charges):
        R = hermite_r(L, p, pr.P[0] - C[0], pr.P[1] - C[1], pr.P[2] - C[2])
        for (ia, ib), d in hc.items():
            acc = np.zeros_like(p)
            for (t, u, v), e in d.items():
                acc = acc + e * R[(t, u, v)]
            V[ia, ib] += np.einsum("p,pab->ab", -zc * 2 * math.pi / p * acc, pr.w)
    return S, T, V


# ---- assembling matrices over contracted spherical functions -----------------------
def offsets(shells):
    off, n = [], 0
    for s in shells:
        off.append(n)
        n += s.nfunc
    return off, n


def to_sph(block, shells_axes):
    """block (cart..., contr...) with one (cart, contr) pair per shell ->
    (func...) with func = contr-major, spherical-minor."""
    out = block
    k = len(shells_axes)
    for ax, sh in enumerate(shells_axes):
        out = np.tensordot(sh.sph, out, axes=([1], [ax]))   # new axis 0 = sph
        out = np.moveaxis(out, 0, ax)
    # now axes: sph_0..sph_{k-1}, contr_0..contr_{k-1}; interleave contr-major
    perm = []
    for i in range(k):
        perm += [k + i, i]
    out = np.transpose(out, perm)
    shape = []
    for sh in shells_axes:
        shape.append(sh.coefs.shape[1] * sh.sph.shape[0])
    return out.reshape(shape)


def four_centre(ao, aoff, nao, pairs, nrm):
    """The exact (mu nu|lambda sigma), normalised like the one-electron integrals."""
    eri = np.zeros((nao,) * 4)
    for i in range(len(ao)):
        for j in range(i + 1):
            for k in range(len(ao)):
                for l in range(k + 1):
                    if (i, j) < (k, l):
                        continue
                    raw = two_electron(pairs[(i, j)], pairs[(k, l)])   # (ca,cb,cc,cd,fa,fb,fc,fd)
                    blk = to_sph(raw, [ao[i], ao[j], ao[k], ao[l]])
                    si, sj, sk, sl_ = (np.s_[aoff[x]:aoff[x] + ao[x].nfunc] for x in (i, j, k, l))
                    for (a_, b_, c_, d_, arr) in (
                            (si, sj, sk, sl_, blk), (sj, si, sk, sl_, blk.transpose(1, 0, 2, 3)),
                            (si, sj, sl_, sk, blk.transpose(0, 1, 3, 2)),
                            (sj, si, sl_, sk, blk.transpose(1, 0, 3, 2))):
                        eri[a_, b_, c_, d_] = arr
                        eri[c_, d_, a_, b_] = arr.transpose(2, 3, 0, 1)
    eri *= np.einsum("p,q,r,s->pqrs", nrm, nrm, nrm, nrm)
    return eri


def fit_integrals(aux, ao, pairs, nrm):
    """(Q|mu nu) (naux, nao, nao) and (P|Q) of an aux basis."""
    aoff, nao = offsets(ao)
    xoff, naux = offsets(aux)
    j3 = np.zeros((naux, nao, nao))
    for x, sx in enumerate(aux):
        px = Pair(sx)
        for i in range(len(ao)):
            for j in range(len(ao)):
                raw = two_electron(px, pairs[(i, j)])   # (cx, 1, ca, cb, fx, 1, fa, fb)
                raw = raw[:, 0][:, :, :, :, 0]           # (cx, ca, cb, fx, fa, fb)
                j3[xoff[x]:xoff[x] + sx.nfunc, aoff[i]:aoff[i] + ao[i].nfunc,
                   aoff[j]:aoff[j] + ao[j].nfunc] = to_sph(raw, [sx, ao[i], ao[j]])
    j3 *= nrm[None, :, None] * nrm[None, None, :]
    j2 = np.zeros((naux, naux))
    for x, sx in enumerate(aux):
        for y, sy in enumerate(aux):
            raw = two_electron(Pair(sx), Pair(sy))[:, 0, :, 0][:, :, :, 0, :, 0]   # (cx, cy, fx, fy)
            j2[xoff[x]:xoff[x] + sx.nfunc, xoff[y]:xoff[y] + sy.nfunc] = to_sph(raw, [sx, sy])
    return j3, j2


def build(xyz_path, text=None, jk_fit=None):
    """RHF + RI factor at one geometry (an .xyz file, or its text): dict of B,
    eps and the energies. jk_fit names a JK fitting basis: the SCF then uses
    density-fitted (mu nu|lambda sigma) = sum_PQ (mu nu|P) (P|Q)^-1 (Q|lambda sigma)
    (the reference's scf_jk = "rijk") instead of the exact four-centre integrals."""
    lines = (text if text is not None else open(xyz_path).read()).split("\n")[2:]
    atoms = []
    for ln in lines:
        f = ln.split()
        if len(f) == 4:
            atoms.append((f[0], np.array([float(x) for x in f[1:]]) / BOHR_ANGSTROM))
    coords = [x for _, x in atoms]
    charges = [Z[s] for s, _ in atoms]
    ao = load_basis("cc-pvdz", atoms)
    aux = load_basis("cc-pvdz-ri", atoms)
    aoff, nao = offsets(ao)
    xoff, naux = offsets(aux)

    S = np.zeros((nao, nao))
    T = np.zeros_like(S)
    V = np.zeros_like(S)
    for i, si in enumerate(ao):
        for j, sj in enumerate(ao):
            s, t, v = one_electron(si, sj, coords, charges)
            sl = np.s_[aoff[i]:aoff[i] + si.nfunc, aoff[j]:aoff[j] + sj.nfunc]
            S[sl] = to_sph(s, [si, sj])
            T[sl] = to_sph(t, [si, sj])
            V[sl] = to_sph(v, [si, sj])
    # normalise the AO functions (energies are invariant; this keeps S well scaled)
    nrm = 1.0 / np.sqrt(np.diag(S))
    S, T, V = (M * nrm[:, None] * nrm[None, :] for M in (S, T, V))

    pairs = {(i, j): Pair(ao[i], ao[j]) for i in range(len(ao)) for j in range(len(ao))}
    if jk_fit is not None:
        jk3, jk2 = fit_integrals(load_basis(jk_fit, atoms), ao, pairs, nrm)
        xf = np.linalg.solve(np.linalg.cholesky(jk2), jk3.reshape(jk3.shape[0], -1))
        eri = (xf.T @ xf).reshape((nao,) * 4)
    else:
        eri = four_centre(ao, aoff, nao, pairs, nrm)

    enuc = 0.0
    for a in range(len(atoms)):
        for b in range(a):
            enuc += charges[a] * charges[b] / np.linalg.norm(coords[a] - coords[b])
    nocc = sum(charges) // 2
    Hc = T + V
    s_val, s_vec = np.linalg.eigh(S)
    X = s_vec @ np.diag(s_val ** -0.5) @ s_vec.T
    e_orb, Cp = np.linalg.eigh(X.T @ Hc @ X)
    C = X @ Cp
    e_old = 0.0
    focks, errs = [], []
    for it in range(200):
        D = 2 * C[:, :nocc] @ C[:, :nocc].T
        J = np.einsum("pqrs,rs->pq", eri, D)
        K = np.einsum("prqs,rs->pq", eri, D)
        F = Hc + J - 0.5 * K
        e_el = 0.5 * np.sum(D * (Hc + F))
        err = X.T @ (F @ D @ S - S @ D @ F) @ X
        focks.append(F)
        errs.append(err)
        focks, errs = focks[-8:], errs[-8:]
        if len(focks) > 1:
            n = len(focks)
            Bm = -np.ones((n + 1, n + 1))
            Bm[n, n] = 0
            for x in range(n):
                for y in range(n):
                    Bm[x, y] = np.sum(errs[x] * errs[y])
            rhs = np.zeros(n + 1)
            rhs[n] = -1
            c = np.linalg.solve(Bm, rhs)[:n]
            F = sum(ci * Fi for ci, Fi in zip(c, focks))
        e_orb, Cp = np.linalg.eigh(X.T @ F @ X)
        C = X @ Cp
        if abs(e_el - e_old) < 1e-13 and np.max(np.abs(err)) < 1e-11:
            break
        e_old = e_el
    # final Fock of the converged density, its orbitals and energies
    D = 2 * C[:, :nocc] @ C[:, :nocc].T
    F = Hc + np.einsum("pqrs,rs->pq", eri, D) - 0.5 * np.einsum("prqs,rs->pq", eri, D)
    e_orb, Cp = np.linalg.eigh(X.T @ F @ X)
    C = X @ Cp
    e_rhf = 0.5 * np.sum(D * (Hc + F)) + enuc

    # fit integrals
    j3, j2 = fit_integrals(aux, ao, pairs, nrm)
    nvir = nao - nocc
    co, cv = C[:, :nocc], C[:, nocc:]
    pia = np.einsum("Pmn,mi,na->Pia", j3, co, cv).reshape(naux, nocc * nvir)
    b = np.linalg.solve(np.linalg.cholesky(j2), pia)
    # exact (ia|jb) of the SCF's own four-centre integrals, an (ov x ov) matrix
    ovov = np.einsum("pqrs,pi,qa,rj,sb->iajb", eri, co, cv, co, cv,
                     optimize=True).reshape(nocc * nvir, nocc * nvir)
    return dict(b_ov=b, eps=e_orb, nocc=nocc, nvir=nvir, naux=naux, e_rhf=e_rhf, enuc=enuc,
                iterations=it + 1, j3=j3, c=C, j2=j2, ovov=ovov)


def exact_factor(m):
    """B (rank x n) with B^T B = m for a symmetric positive semidefinite m: the
    eigenvectors scaled by the square roots of the eigenvalues, dropping those at
    or below 1e-14 of the largest (round-off of a Gram matrix)."""
    w, v = np.linalg.eigh(0.5 * (m + m.T))
    keep = w > 1e-14 * w[-1]
    return (v[:, keep] * np.sqrt(w[keep])).T.copy()


def numpy_mp2(b, eps, nocc, nvir, fc=0):
    """(E_os, E_ss) by a plain full (i, j) double loop over the active occupied."""
    B3 = b.reshape(b.shape[0], nocc, nvir)
    eo, ev = eps[:nocc], eps[nocc:]
    e_os = e_ss = 0.0
    for i in range(fc, nocc):
        for j in range(fc, nocc):
            g = B3[:, i, :].T @ B3[:, j, :]
            d = eo[i] + eo[j] - ev[:, None] - ev[None, :]
            e_os += np.sum(g * g / d)
            e_ss += np.sum(g * (g - g.T) / d)
    return e_os, e_ss


def main():
    out = {}
    ref = json.load(open(REF_JSON))
    w = build(XYZ)
    print(f"water: RHF {w['e_rhf']:.12f} (reference {ref['rhf_energy']:.12f}, diff "
          f"{w['e_rhf'] - ref['rhf_energy']:.2e}) after {w['iterations']} iterations", flush=True)
    if abs(w["e_rhf"] - ref["rhf_energy"]) > 1e-8:
        sys.exit("RHF energy does not match the reference: integrals, geometry or units are off")
    os_, ss_ = numpy_mp2(w["b_ov"], w["eps"], w["nocc"], w["nvir"])
    print(f"water: RI-MP2 E_corr {os_ + ss_:.12f} (reference {ref['mp2_corr']:.12f}, diff "
          f"{os_ + ss_ - ref['mp2_corr']:.2e})", flush=True)
    out.update(b_ov=w["b_ov"], eps=w["eps"], nocc=w["nocc"], nvir=w["nvir"], naux=w["naux"],
               e_rhf=w["e_rhf"], e_corr_numpy=os_ + ss_, e_corr_reference=ref["mp2_corr"],
               e_rhf_reference=ref["rhf_energy"])

    grad = json.load(open(REF_GRAD))
    scs = json.load(open(REF_SCS))
    d = build(XYZ_DISTORTED)
    print(f"distorted: E_nuc {d['enuc']:.12f} (reference {grad['nuclear_repulsion']:.12f}); RHF "
          f"{d['e_rhf']:.12f} (PySCF {grad['pyscf']['e_rhf']:.12f}, ORCA {grad['orca']['e_rhf']:.12f})",
          flush=True)
    if abs(d["enuc"] - grad["nuclear_repulsion"]) > 1e-9 or \
            abs(d["e_rhf"] - grad["pyscf"]["e_rhf"]) > 1e-8:
        sys.exit("distorted water: nuclear repulsion or RHF energy does not match")
    os_, ss_ = numpy_mp2(d["b_ov"], d["eps"], d["nocc"], d["nvir"])
    os_fc, ss_fc = numpy_mp2(d["b_ov"], d["eps"], d["nocc"], d["nvir"], fc=1)
    print(f"distorted: E_corr {os_ + ss_:.12f} (PySCF {grad['pyscf']['e_corr']:.12f}, diff "
          f"{os_ + ss_ - grad['pyscf']['e_corr']:.2e}); os {os_:.12f} (PySCF {scs['pyscf']['e_os']:.12f}) "
          f"ss {ss_:.12f} (PySCF {scs['pyscf']['e_ss']:.12f}); fc os {os_fc:.12f} "
          f"(PySCF {scs['pyscf']['e_os_fc']:.12f}) ss {ss_fc:.12f} (PySCF {scs['pyscf']['e_ss_fc']:.12f})",
          flush=True)
    out.update(dist_b_ov=d["b_ov"], dist_eps=d["eps"], dist_e_rhf=d["e_rhf"],
               dist_e_os_numpy=os_, dist_e_ss_numpy=ss_,
               dist_e_corr_ref=grad["pyscf"]["e_corr"], dist_e_corr_orca=grad["orca"]["e_corr"],
               dist_e_corr_fc_orca=grad["orca_frozen_core"]["e_corr"],
               dist_e_os_ref=scs["pyscf"]["e_os"], dist_e_ss_ref=scs["pyscf"]["e_ss"],
               dist_e_os_fc_ref=scs["pyscf"]["e_os_fc"], dist_e_ss_fc_ref=scs["pyscf"]["e_ss_fc"],
               dist_e_rhf_ref=grad["pyscf"]["e_rhf"],
               # inputs of the dressing / MO transform: (Q|mu nu), MO coefficients, (P|Q)
               dist_j3=d["j3"], dist_c=d["c"], dist_j2=d["j2"])
    # Water with a density-fitted SCF (def2-universal-jkfit for J and K) and the
    # cc-pVDZ-RI fit for MP2: the reference's RI-JK known answer
    # (testdata/reference/h2o_cc-pvdz_rijk-def2-universal-jkfit_dfmp2.json, PySCF).
    rj = json.load(open(REF_RIJK))
    r = build(XYZ, jk_fit=rj["jk_auxbasis"])
    os_, ss_ = numpy_mp2(r["b_ov"], r["eps"], r["nocc"], r["nvir"])
    print(f"RI-JK water: RHF {r['e_rhf']:.12f} (PySCF {rj['rhf_energy']:.12f}, diff "
          f"{r['e_rhf'] - rj['rhf_energy']:.2e}); E_corr {os_ + ss_:.12f} (PySCF "
          f"{rj['mp2_corr']:.12f}, diff {os_ + ss_ - rj['mp2_corr']:.2e})", flush=True)
    if abs(r["e_rhf"] - rj["rhf_energy"]) > 1e-8:
        sys.exit("RI-JK RHF energy does not match the reference")
    out.update(rijk_b_ov=r["b_ov"], rijk_eps=r["eps"], rijk_e_rhf=r["e_rhf"],
               rijk_e_rhf_reference=rj["rhf_energy"], rijk_e_corr_numpy=os_ + ss_,
               rijk_e_corr_reference=rj["mp2_corr"])

    # H2 / cc-pVDZ at 0.74 Angstrom, the reference's own known-answer test
    # (rimp2.rs:3187-3201): RI-MP2 within 1e-4 of PySCF's canonical MP2
    # -0.0263715576 (the RI error is of that order).
    h = build(None, text="2\nH2\nH 0 0 0\nH 0 0 0.74\n")
    os_, ss_ = numpy_mp2(h["b_ov"], h["eps"], h["nocc"], h["nvir"])
    print(f"H2: RHF {h['e_rhf']:.10f}, RI-MP2 E_corr {os_ + ss_:.10f} (canonical MP2 "
          f"-0.0263715576, diff {os_ + ss_ + 0.0263715576:.2e})", flush=True)
    if abs(os_ + ss_ + 0.0263715576) > 1e-4:
        sys.exit("H2 RI-MP2 energy is not within the reference test's 1e-4 of canonical MP2")
    out.update(h2_b_ov=h["b_ov"], h2_eps=h["eps"], h2_nocc=h["nocc"], h2_nvir=h["nvir"],
               h2_e_rhf=h["e_rhf"], h2_e_corr_numpy=os_ + ss_, h2_e_os_numpy=os_,
               h2_e_corr_canonical=-0.0263715576)

    # Conventional MP2 of the equilibrium water (testdata/reference/h2o_cc-pvdz_mp2.json,
    # PySCF): B is an exact factor of the four-centre (ia|jb), so the same pair-energy
    # path computes the unfitted energy.
    cv_ = json.load(open(REF_CONV))
    bx = exact_factor(w["ovov"])
    err = np.max(np.abs(bx.T @ bx - w["ovov"]))
    os_, ss_ = numpy_mp2(bx, w["eps"], w["nocc"], w["nvir"])
    print(f"conventional water: rank {bx.shape[0]} of {bx.shape[1]}, |B^T B - (ia|jb)| {err:.1e}; "
          f"E_corr {os_ + ss_:.12f} (PySCF {cv_['correlation_energy']:.12f}, diff "
          f"{os_ + ss_ - cv_['correlation_energy']:.2e})", flush=True)
    if abs(cv_["scf_energy"] - ref["rhf_energy"]) > 1e-9 or \
            abs(os_ + ss_ - cv_["correlation_energy"]) > 1e-8:
        sys.exit("conventional MP2 of the exact factor does not match the reference")
    out.update(conv_b_ov=bx, conv_e_corr_numpy=os_ + ss_,
               conv_e_corr_reference=cv_["correlation_energy"],
               conv_e_rhf_reference=cv_["scf_energy"])
    np.savez_compressed(OUT, **out)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
