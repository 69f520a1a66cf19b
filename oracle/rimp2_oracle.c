/*
 * CPU oracle for the closed-shell RI-MP2 pair-energy stage and its adjoint.
 *
 * TEST INFRASTRUCTURE ONLY. This file is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it, and only to
 * compare against or time beside the HIP library. No product code path links,
 * loads or calls it.
 *
 * It restates, loop for loop, the reference's CPU algorithm (ferric, Rust):
 *   - active_occ                           crates/ferric-mp2/src/rimp2.rs:99-106
 *   - pair_energy (a-major, then b; one divide per term; kappa damping)
 *                                          crates/ferric-mp2/src/rimp2.rs:1399-1441
 *   - spin_components_from_b_ov_kappa_cpu  (per-i tail block G_i = B_i^T B_tail)
 *                                          crates/ferric-mp2/src/rimp2.rs:1445-1499
 *   - spin_components_from_blocks          (j >= i, fac 1/2, i-ordered serial fold)
 *                                          crates/ferric-mp2/src/rimp2.rs:1506-1544
 *   - spin_components_from_g               crates/ferric-mp2/src/rimp2.rs:984-1018
 *   - compute_t2_and_integrals             crates/ferric-mp2/src/oo_rimp2.rs:1079-1122
 *   - build_x_ov_weighted / tt_block       crates/ferric-mp2/src/zvector.rs:651-697
 *   - build_mp2_density                    crates/ferric-mp2/src/oo_rimp2.rs:1192-1300
 *   - opposite_spin_density                crates/ferric-mp2/src/rimp2.rs:2026-2052
 *   - apply_spin_scaling / SpinScaling     crates/ferric-mp2/src/rimp2.rs:1917-2011
 *   - energy error bound (Higham model)    crates/ferric-mp2/tests/common/rimp2_error_bound.rs
 *   - same_spin_block_energy / same_spin_energy_only
 *                                          crates/ferric-mp2/src/u_rimp2.rs:693-729, 766-797
 *   - opposite_spin_block_energy / opposite_spin_energy_only
 *                                          crates/ferric-mp2/src/u_rimp2.rs:731-760, 978-995
 *   - signed_pair_block / spin_components_from_signed_b_ov
 *                                          crates/ferric-mp2/src/rimp2.rs:1559-1607
 *   - stream_dressed_mo_band(_budgeted)    crates/ferric-mp2/src/rimp2.rs:648-756
 *
 * The reference forms G_i with an OpenBLAS dgemm; here it is a plain blocked
 * loop in the same (P-outer) association class. The two agree to the
 * 2*gamma_naux*S_ab element bound the reference's own device tests use, not
 * bit for bit (no test of the reference pins BLAS bits either).
 *
 * Pinning (tests/test_oracle.py): the reference's own unit-test generators
 * and shapes (rimp2.rs:2856-2927 LCG, gpu_rimp2_resident.rs:92-120,
 * zvector.rs:1135-1175 xorshift), an independent numpy brute force over the
 * full (i,j) double loop, the spin-orbital identity of the closed-shell
 * split, the reference's own mutation (testdata/mutations/ri-mp2.patch must
 * fail), the kappa limits, central finite differences of the energy for the
 * adjoint, Euler homogeneity (sum B*dE/dB = 4E, sum eps*dE/deps = -E) and
 * tr P_oo + tr P_vv = 0.
 *
 * Layout: b is row-major (naux, nocc*nvir), b[P*nov + i*nvir + a] = B^P_ia
 * over ACTIVE occupied i; eps is the full MO energy vector, occupied i at
 * eps[first_occ + i], virtual a at eps[nocc_total + a].
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_OK 0
#define ORC_LAYOUT 1
#define ORC_ALLOC 2

/* rimp2.rs:99-106: an error (not an underflow) when the freeze covers every
 * occupied orbital; frozen_core == 0 with nocc_total == 0 is allowed. */
int orc_active_occ(size_t nocc_total, size_t frozen_core, size_t *nocc) {
    if (frozen_core != 0 && frozen_core >= nocc_total) return ORC_LAYOUT;
    *nocc = nocc_total - frozen_core;
    return ORC_OK;
}

/* G_i over the j >= i tail only: g_i[a*ld + (j-i)*nvir + b] = (ia|jb), with
 * ld = (nocc-i)*nvir (rimp2.rs:1489-1494). Blocked over output columns so the
 * block stays in cache; within an element the sum runs over P ascending. */
void orc_tail_block(const double *b, size_t naux, size_t nocc, size_t nvir, size_t i,
                    double *g_i) {
    const size_t nov = nocc * nvir;
    const size_t ld = (nocc - i) * nvir;
    const size_t col0 = i * nvir;
    const size_t CB = 512;
    memset(g_i, 0, sizeof(double) * nvir * ld);
    for (size_t c0 = 0; c0 < ld; c0 += CB) {
        size_t c1 = c0 + CB < ld ? c0 + CB : ld;
        for (size_t p = 0; p < naux; ++p) {
            const double *row = b + p * nov + col0;
            for (size_t a = 0; a < nvir; ++a) {
                const double s = row[a];
                double *g = g_i + a * ld;
                for (size_t c = c0; c < c1; ++c) g[c] += s * row[c];
            }
        }
    }
}

/* rimp2.rs:1399-1441, arithmetic and loop order as written there. */
void orc_pair_energy(const double *g_i, size_t ld, size_t i, size_t j, size_t nvir,
                     const double *eps, size_t first_occ, size_t nocc_total, int has_kappa,
                     double kappa, double *e_os_out, double *e_ss_out) {
    const size_t jcol = (j - i) * nvir;
    const double e_ij = eps[first_occ + i] + eps[first_occ + j];
    double e_os = 0.0, e_ss = 0.0;
    if (!has_kappa) {
        for (size_t a = 0; a < nvir; ++a) {
            for (size_t bb = 0; bb < nvir; ++bb) {
                const double g_ab = g_i[a * ld + jcol + bb];
                const double g_ba = g_i[bb * ld + jcol + a];
                const double denom = e_ij - eps[nocc_total + a] - eps[nocc_total + bb];
                e_os += g_ab * g_ab / denom;
                e_ss += g_ab * (g_ab - g_ba) / denom;
            }
        }
    } else {
        for (size_t a = 0; a < nvir; ++a) {
            for (size_t bb = 0; bb < nvir; ++bb) {
                const double g_ab = g_i[a * ld + jcol + bb];
                const double g_ba = g_i[bb * ld + jcol + a];
                const double denom = e_ij - eps[nocc_total + a] - eps[nocc_total + bb];
                const double d1 = 1.0 - exp(kappa * denom);
                const double damp = d1 * d1;
                e_os += damp * g_ab * g_ab / denom;
                e_ss += damp * g_ab * (g_ab - g_ba) / denom;
            }
        }
    }
    *e_os_out = e_os;
    *e_ss_out = e_ss;
}

/* The per-i body of spin_components_from_blocks (rimp2.rs:1520-1530): j
 * ascending from i, fac = 1 on the diagonal and 2 off it. */
void orc_i_partial(const double *g_i, size_t i, size_t nocc, size_t nvir, const double *eps,
                   size_t first_occ, size_t nocc_total, int has_kappa, double kappa,
                   double *e_os_i, double *e_ss_i) {
    const size_t ld = (nocc - i) * nvir;
    double os = 0.0, ss = 0.0;
    for (size_t j = i; j < nocc; ++j) {
        const double fac = (i == j) ? 1.0 : 2.0;
        double pos, pss;
        orc_pair_energy(g_i, ld, i, j, nvir, eps, first_occ, nocc_total, has_kappa, kappa,
                        &pos, &pss);
        os += fac * pos;
        ss += fac * pss;
    }
    *e_os_i = os;
    *e_ss_i = ss;
}

static int check_layout(size_t naux, size_t nocc, size_t nvir, size_t n_eps, size_t first_occ,
                        size_t nocc_total) {
    (void)naux;
    size_t need_o = first_occ + nocc, need_v = nocc_total + nvir;
    size_t need = need_o > need_v ? need_o : need_v;
    if (n_eps < need) return ORC_LAYOUT;
    return ORC_OK;
}

/* spin_components_from_b_ov_kappa_cpu + spin_components_from_blocks
 * (rimp2.rs:1445-1544): one task per i (OpenMP stands in for rayon), each with
 * its own tail block, partials collected per i and folded serially in i order,
 * so the result does not depend on the thread count. out = {e_os, e_ss, e_total}. */
int orc_spin_components_from_b_ov(const double *b, size_t naux, size_t nocc, size_t nvir,
                                  const double *eps, size_t n_eps, size_t first_occ,
                                  size_t nocc_total, int has_kappa, double kappa, int nthreads,
                                  double *out) {
    if (check_layout(naux, nocc, nvir, n_eps, first_occ, nocc_total)) return ORC_LAYOUT;
    out[0] = out[1] = out[2] = 0.0;
    if (nocc == 0 || nvir == 0) return ORC_OK;
    double *part = (double *)calloc(2 * nocc, sizeof(double));
    if (!part) return ORC_ALLOC;
    int failed = 0;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
#endif
    for (long ii = 0; ii < (long)nocc; ++ii) {
        size_t i = (size_t)ii;
        double *g = (double *)malloc(sizeof(double) * nvir * (nocc - i) * nvir);
        if (!g) {
            failed = 1;
            continue;
        }
        orc_tail_block(b, naux, nocc, nvir, i, g);
        orc_i_partial(g, i, nocc, nvir, eps, first_occ, nocc_total, has_kappa, kappa,
                      &part[2 * i], &part[2 * i + 1]);
        free(g);
    }
    (void)nthreads;
    if (failed) {
        free(part);
        return ORC_ALLOC;
    }
    double os = 0.0, ss = 0.0;
    for (size_t i = 0; i < nocc; ++i) {
        os += part[2 * i];
        ss += part[2 * i + 1];
    }
    free(part);
    out[0] = os;
    out[1] = ss;
    out[2] = os + ss;
    return ORC_OK;
}

/* rimp2.rs:984-1018: the same energy from an explicit (nov x nov) (ia|jb)
 * matrix, g[(i*nvir+a)*nov + j*nvir+b]. */
void orc_spin_components_from_g(const double *g, size_t nocc, size_t nvir, const double *eps,
                                size_t first_occ, size_t nocc_total, double *out) {
    const size_t nov = nocc * nvir;
    double e_os = 0.0, e_ss = 0.0;
    for (size_t i = 0; i < nocc; ++i) {
        const double e_i = eps[first_occ + i];
        for (size_t j = i; j < nocc; ++j) {
            const double e_j = eps[first_occ + j];
            const double fac = (i == j) ? 1.0 : 2.0;
            for (size_t a = 0; a < nvir; ++a) {
                const double e_a = eps[nocc_total + a];
                for (size_t bb = 0; bb < nvir; ++bb) {
                    const double e_b = eps[nocc_total + bb];
                    const double d = e_i + e_j - e_a - e_b;
                    const double iajb = g[(i * nvir + a) * nov + j * nvir + bb];
                    const double ibja = g[(i * nvir + bb) * nov + j * nvir + a];
                    e_os += fac * iajb * iajb / d;
                    e_ss += fac * iajb * (iajb - ibja) / d;
                }
            }
        }
    }
    out[0] = e_os;
    out[1] = e_ss;
    out[2] = e_os + e_ss;
}

/* ---------------------------------------------------------------------------
 * Adjoint intermediates.
 * ------------------------------------------------------------------------- */

/* oo_rimp2.rs:1079-1122: eri[ia*nov + jb] = sum_P B^P_ia B^P_jb and
 * t2[ia*nov + jb] = eri / (e_i + e_j - e_a - e_b). eri may be NULL. */
int orc_t2_and_integrals(const double *b, size_t naux, size_t nocc, size_t nvir,
                         const double *eps, size_t nocc_total, size_t first_occ, double *t2,
                         double *eri) {
    const size_t nov = nocc * nvir;
    double *k = eri ? eri : (double *)malloc(sizeof(double) * nov * nov);
    if (!k) return ORC_ALLOC;
    memset(k, 0, sizeof(double) * nov * nov);
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (long r = 0; r < (long)nov; ++r) {
        double *row = k + (size_t)r * nov;
        for (size_t p = 0; p < naux; ++p) {
            const double s = b[p * nov + (size_t)r];
            const double *bp = b + p * nov;
            for (size_t c = 0; c < nov; ++c) row[c] += s * bp[c];
        }
    }
    for (size_t ia = 0; ia < nov; ++ia) {
        const size_t i = ia / nvir, a = ia % nvir;
        for (size_t j = 0; j < nocc; ++j) {
            for (size_t bb = 0; bb < nvir; ++bb) {
                const size_t jb = j * nvir + bb;
                const double denom = eps[first_occ + i] + eps[first_occ + j] -
                                     eps[nocc_total + a] - eps[nocc_total + bb];
                t2[ia * nov + jb] = k[ia * nov + jb] / denom;
            }
        }
    }
    if (!eri) free(k);
    return ORC_OK;
}

/* zvector.rs:651-697: x[P, ia] = sum_jb TT_i[a, jb] B^P_jb with
 * TT_i[a, jb] = d t_ij^ab - e t_ij^ba, (d, e) = (c_os + c_ss, c_ss). */
int orc_x_ov_weighted(const double *t2, const double *b, size_t nocc, size_t nvir, size_t naux,
                      double c_os, double c_ss, double *x) {
    const size_t nov = nocc * nvir;
    const double d = c_os + c_ss, e = c_ss;
    memset(x, 0, sizeof(double) * naux * nov);
    if (nvir == 0) return ORC_OK;
    int failed = 0;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1)
#endif
    for (long ii = 0; ii < (long)nocc; ++ii) {
        const size_t i = (size_t)ii;
        double *tt = (double *)malloc(sizeof(double) * nvir * nov);
        if (!tt) {
            failed = 1;
            continue;
        }
        for (size_t a = 0; a < nvir; ++a)
            for (size_t j = 0; j < nocc; ++j)
                for (size_t bb = 0; bb < nvir; ++bb) {
                    const double t_ab = t2[(i * nvir + a) * nov + j * nvir + bb];
                    const double t_ba = t2[(i * nvir + bb) * nov + j * nvir + a];
                    tt[a * nov + j * nvir + bb] = d * t_ab - e * t_ba;
                }
        for (size_t p = 0; p < naux; ++p) {
            const double *bp = b + p * nov;
            for (size_t a = 0; a < nvir; ++a) {
                const double *tr = tt + a * nov;
                double s = 0.0;
                for (size_t c = 0; c < nov; ++c) s += bp[c] * tr[c];
                x[p * nov + i * nvir + a] = s;
            }
        }
        free(tt);
    }
    return failed ? ORC_ALLOC : ORC_OK;
}

/* oo_rimp2.rs:1192-1300 restated with plain loops:
 *   P_ij = - sum_kab t_ik^ab (d t_jk^ab - e t_jk^ba)
 *   P_ab =   sum_ijc t_ij^ac (d t_ij^bc - e t_ij^cb)
 * (d, e) = (2, 1) is build_mp2_density, (1, 0) is opposite_spin_density
 * (rimp2.rs:2026-2052). */
static void mp2_density_de(const double *t2, size_t nocc, size_t nvir, double d, double e,
                           double *p_oo, double *p_vv) {
    const size_t nov = nocc * nvir;
#define T2(i, a, j, b) t2[((i) * nvir + (a)) * nov + (j) * nvir + (b)]
    for (size_t i = 0; i < nocc; ++i)
        for (size_t j = 0; j < nocc; ++j) {
            double s = 0.0;
            for (size_t k = 0; k < nocc; ++k)
                for (size_t a = 0; a < nvir; ++a)
                    for (size_t bb = 0; bb < nvir; ++bb)
                        s += T2(i, a, k, bb) * (d * T2(j, a, k, bb) - e * T2(j, bb, k, a));
            p_oo[i * nocc + j] = -s;
        }
    for (size_t a = 0; a < nvir; ++a)
        for (size_t bb = 0; bb < nvir; ++bb) {
            double s = 0.0;
            for (size_t i = 0; i < nocc; ++i)
                for (size_t j = 0; j < nocc; ++j)
                    for (size_t c = 0; c < nvir; ++c)
                        s += T2(i, a, j, c) * (d * T2(i, bb, j, c) - e * T2(i, c, j, bb));
            p_vv[a * nvir + bb] = s;
        }
#undef T2
}

void orc_mp2_density(const double *t2, size_t nocc, size_t nvir, double *p_oo, double *p_vv) {
    mp2_density_de(t2, nocc, nvir, 2.0, 1.0, p_oo, p_vv);
}

void orc_opposite_spin_density(const double *t2, size_t nocc, size_t nvir, double *p_oo,
                               double *p_vv) {
    mp2_density_de(t2, nocc, nvir, 1.0, 0.0, p_oo, p_vv);
}

/* apply_spin_scaling (rimp2.rs:1917-1950): MP2 unchanged; c_os == c_ss
 * rescales; otherwise P = c_ss P_mp2 + (c_os - c_ss) P_os. */
int orc_scaled_density(const double *t2, size_t nocc, size_t nvir, double c_os, double c_ss,
                       double *p_oo, double *p_vv) {
    orc_mp2_density(t2, nocc, nvir, p_oo, p_vv);
    if (c_os == 1.0 && c_ss == 1.0) return ORC_OK;
    if (c_os == c_ss) {
        for (size_t q = 0; q < nocc * nocc; ++q) p_oo[q] *= c_ss;
        for (size_t q = 0; q < nvir * nvir; ++q) p_vv[q] *= c_ss;
        return ORC_OK;
    }
    double *os_oo = (double *)malloc(sizeof(double) * (nocc * nocc + nvir * nvir + 1));
    if (!os_oo) return ORC_ALLOC;
    double *os_vv = os_oo + nocc * nocc;
    orc_opposite_spin_density(t2, nocc, nvir, os_oo, os_vv);
    for (size_t q = 0; q < nocc * nocc; ++q) p_oo[q] = c_ss * p_oo[q] + (c_os - c_ss) * os_oo[q];
    for (size_t q = 0; q < nvir * nvir; ++q) p_vv[q] = c_ss * p_vv[q] + (c_os - c_ss) * os_vv[q];
    free(os_oo);
    return ORC_OK;
}

/* The adjoint of E = c_os E_os + c_ss E_ss through the reference's gradient
 * intermediates: Gamma = dE/dB = 4 x_ov (x = 1/4 dE/dB, rimp2.rs:1832-1840),
 * dE/deps_k = 2 P_kk on the active occupied block, 2 P_cc on the virtual
 * block, 0 for frozen core (P + P^T = dE/dF). out = {E_os, E_ss, E}.
 * Needs the dense nov x nov amplitudes, as the reference does. */
int orc_rimp2_adjoint(const double *b, size_t naux, size_t nocc, size_t nvir, const double *eps,
                      size_t n_eps, size_t first_occ, size_t nocc_total, double c_os, double c_ss,
                      double *gamma, double *deps, double *out) {
    if (check_layout(naux, nocc, nvir, n_eps, first_occ, nocc_total)) return ORC_LAYOUT;
    const size_t nov = nocc * nvir;
    memset(gamma, 0, sizeof(double) * naux * nov);
    memset(deps, 0, sizeof(double) * n_eps);
    out[0] = out[1] = out[2] = 0.0;
    if (nov == 0) return ORC_OK;
    double *t2 = (double *)malloc(sizeof(double) * nov * nov);
    double *eri = (double *)malloc(sizeof(double) * nov * nov);
    double *p = (double *)malloc(sizeof(double) * (nocc * nocc + nvir * nvir));
    if (!t2 || !eri || !p) {
        free(t2);
        free(eri);
        free(p);
        return ORC_ALLOC;
    }
    int rc = orc_t2_and_integrals(b, naux, nocc, nvir, eps, nocc_total, first_occ, t2, eri);
    if (!rc) {
        double sc[3];
        orc_spin_components_from_g(eri, nocc, nvir, eps, first_occ, nocc_total, sc);
        out[0] = sc[0];
        out[1] = sc[1];
        out[2] = c_os * sc[0] + c_ss * sc[1];
        rc = orc_x_ov_weighted(t2, b, nocc, nvir, naux, c_os, c_ss, gamma);
    }
    if (!rc) {
        for (size_t q = 0; q < naux * nov; ++q) gamma[q] *= 4.0;
        rc = orc_scaled_density(t2, nocc, nvir, c_os, c_ss, p, p + nocc * nocc);
    }
    if (!rc) {
        for (size_t k = 0; k < nocc; ++k) deps[first_occ + k] = 2.0 * p[k * nocc + k];
        for (size_t c = 0; c < nvir; ++c)
            deps[nocc_total + c] = 2.0 * p[nocc * nocc + c * nvir + c];
    }
    free(t2);
    free(eri);
    free(p);
    return rc;
}

/* ---------------------------------------------------------------------------
 * Error bound (tests/common/rimp2_error_bound.rs restated): with every G
 * element perturbed by at most d_ab = eps_g * S_ab + eta,
 * S_ab = sum_P |B^P_ia||B^P_jb|, the (OS, SS) energies move by at most
 *   OS: fac*damp*d_ab(2|g_ab| + d_ab)/|D|
 *   SS: fac*damp*[d_ab(2|g_ab| + d_ab) + d_ab|g_ba| + |g_ab| d_ba + d_ab d_ba]/|D|
 * plus the f64 summation term gamma_m (sum|t| + sum|t^|), m = nvir^2 + 2 nocc + 8.
 * bound = {os, ss}.
 * ------------------------------------------------------------------------- */
/* The bound's sums for one occupied row i from its tail blocks g_i
 * (= B_i^T B_tail) and s_i (= |B_i|^T |B_tail|), both laid out as
 * orc_tail_block writes them: acc = {bos, bss, abs_os, abs_ss} of that row. */
void orc_energy_bound_row(const double *g, const double *s, size_t i, size_t nocc, size_t nvir,
                          const double *eps, size_t first_occ, size_t nocc_total, int has_kappa,
                          double kappa, double eps_g, double eta, double *acc) {
    const size_t ld = (nocc - i) * nvir;
    double bos = 0.0, bss = 0.0, abs_os = 0.0, abs_ss = 0.0;
    for (size_t j = i; j < nocc; ++j) {
        const double fac = (i == j) ? 1.0 : 2.0;
        const size_t jcol = (j - i) * nvir;
        const double e_ij = eps[first_occ + i] + eps[first_occ + j];
        for (size_t a = 0; a < nvir; ++a)
            for (size_t bb = 0; bb < nvir; ++bb) {
                const double gab = fabs(g[a * ld + jcol + bb]);
                const double gba = fabs(g[bb * ld + jcol + a]);
                const double dab = eps_g * s[a * ld + jcol + bb] + eta;
                const double dba = eps_g * s[bb * ld + jcol + a] + eta;
                const double denom = e_ij - eps[nocc_total + a] - eps[nocc_total + bb];
                double damp = 1.0;
                if (has_kappa) {
                    const double d1 = 1.0 - exp(kappa * denom);
                    damp = d1 * d1;
                }
                const double ad = fabs(denom);
                bos += fac * damp * dab * (2.0 * gab + dab) / ad;
                bss += fac * damp * (dab * (2.0 * gab + dab) + dab * gba + gab * dba +
                                     dab * dba) / ad;
                abs_os += fac * damp * gab * gab / ad;
                abs_ss += fac * damp * gab * fabs(gab - gba) / ad;
            }
    }
    acc[0] = bos;
    acc[1] = bss;
    acc[2] = abs_os;
    acc[3] = abs_ss;
}

/* The f64 summation term gamma_m with m = nvir^2 + 2 nocc + 8 applied to the
 * sums acc = {bos, bss, abs_os, abs_ss}: bound = {os, ss}. */
void orc_energy_bound_finish(const double *acc, size_t nocc, size_t nvir, double *bound) {
    const double u = ldexp(1.0, -53);
    const double m = (double)(nvir * nvir + 2 * nocc + 8);
    const double gam = m * u / (1.0 - m * u);
    bound[0] = acc[0] + gam * (2.0 * acc[2] + acc[0]);
    bound[1] = acc[1] + gam * (2.0 * acc[3] + acc[1]);
}

int orc_energy_bound(const double *b, size_t naux, size_t nocc, size_t nvir, const double *eps,
                     size_t first_occ, size_t nocc_total, int has_kappa, double kappa,
                     double eps_g, double eta, double *bound) {
    const size_t nov = nocc * nvir;
    double *absb = (double *)malloc(sizeof(double) * naux * nov);
    if (!absb) return ORC_ALLOC;
    for (size_t q = 0; q < naux * nov; ++q) absb[q] = fabs(b[q]);
    double tot[4] = {0.0, 0.0, 0.0, 0.0};
    for (size_t i = 0; i < nocc; ++i) {
        const size_t ld = (nocc - i) * nvir;
        double *g = (double *)malloc(sizeof(double) * nvir * ld);
        double *s = (double *)malloc(sizeof(double) * nvir * ld);
        if (!g || !s) {
            free(g);
            free(s);
            free(absb);
            return ORC_ALLOC;
        }
        orc_tail_block(b, naux, nocc, nvir, i, g);
        orc_tail_block(absb, naux, nocc, nvir, i, s);
        double acc[4];
        orc_energy_bound_row(g, s, i, nocc, nvir, eps, first_occ, nocc_total, has_kappa, kappa,
                             eps_g, eta, acc);
        for (int k = 0; k < 4; ++k) tot[k] += acc[k];
        free(g);
        free(s);
    }
    free(absb);
    orc_energy_bound_finish(tot, nocc, nvir, bound);
    return ORC_OK;
}

/* ---- Unrestricted blocks and the signed factor -------------------------- */

/* same_spin_block_energy (u_rimp2.rs:693-729) on the j >= i tail block of one
 * spin channel: 0.25 * sum_{j>=i} fac * sum_ab t*K, K = g_ab - g_ba,
 * t = K/denom, loops j, a, b ascending. */
static double u_same_spin_block(const double *g_i, size_t i, size_t nocc, size_t nvir,
                                const double *eps, size_t first_occ, size_t nocc_total) {
    const size_t ld = (nocc - i) * nvir;
    const double eps_i = eps[first_occ + i];
    double energy_i = 0.0;
    for (size_t j = i; j < nocc; ++j) {
        const double fac = (i == j) ? 1.0 : 2.0;
        const size_t jcol = (j - i) * nvir;
        const double eps_j = eps[first_occ + j];
        double energy_ij = 0.0;
        for (size_t a = 0; a < nvir; ++a) {
            const double eps_a = eps[nocc_total + a];
            for (size_t bb = 0; bb < nvir; ++bb) {
                const double eps_b = eps[nocc_total + bb];
                const double k = g_i[a * ld + jcol + bb] - g_i[bb * ld + jcol + a];
                const double denom = eps_i + eps_j - eps_a - eps_b;
                const double t_val = k / denom;
                energy_ij += t_val * k;
            }
        }
        energy_i += fac * energy_ij;
    }
    return 0.25 * energy_i;
}

/* same_spin_energy_only (u_rimp2.rs:766-797): per-i tail blocks, partials
 * summed in ascending i. */
int orc_u_same_spin(const double *b, size_t naux, size_t nocc, size_t nvir, const double *eps,
                    size_t n_eps, size_t first_occ, size_t nocc_total, int nthreads,
                    double *energy) {
    if (check_layout(naux, nocc, nvir, n_eps, first_occ, nocc_total)) return ORC_LAYOUT;
    *energy = 0.0;
    if (nocc == 0 || nvir == 0) return ORC_OK;
    double *part = (double *)calloc(nocc, sizeof(double));
    if (!part) return ORC_ALLOC;
    int failed = 0;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
#endif
    for (long ii = 0; ii < (long)nocc; ++ii) {
        const size_t i = (size_t)ii;
        double *g = (double *)malloc(sizeof(double) * nvir * (nocc - i) * nvir);
        if (!g) {
            failed = 1;
            continue;
        }
        orc_tail_block(b, naux, nocc, nvir, i, g);
        part[i] = u_same_spin_block(g, i, nocc, nvir, eps, first_occ, nocc_total);
        free(g);
    }
    (void)nthreads;
    if (!failed)
        for (size_t i = 0; i < nocc; ++i) *energy += part[i];
    free(part);
    return failed ? ORC_ALLOC : ORC_OK;
}

/* opposite_spin_energy_only + opposite_spin_block_energy (u_rimp2.rs:978-995,
 * 731-760): g_i = B_a[:, i*nvir_a..]^T B_b (nvir_a x nocc_b*nvir_b), then
 * sum over a, J, B ascending of (eri/denom)*eri; per-i partials summed in
 * ascending i. Both channels share one aux basis (naux rows). */
int orc_u_opposite_spin(const double *b_a, const double *b_b, size_t naux, size_t nocc_a,
                        size_t nvir_a, size_t nocc_b, size_t nvir_b, const double *eps_a,
                        size_t n_eps_a, size_t first_occ_a, size_t nocc_total_a,
                        const double *eps_b, size_t n_eps_b, size_t first_occ_b,
                        size_t nocc_total_b, int nthreads, double *energy) {
    if (check_layout(naux, nocc_a, nvir_a, n_eps_a, first_occ_a, nocc_total_a) ||
        check_layout(naux, nocc_b, nvir_b, n_eps_b, first_occ_b, nocc_total_b))
        return ORC_LAYOUT;
    *energy = 0.0;
    if (nocc_a == 0 || nvir_a == 0 || nocc_b == 0 || nvir_b == 0) return ORC_OK;
    const size_t nov_a = nocc_a * nvir_a, nov_b = nocc_b * nvir_b;
    double *part = (double *)calloc(nocc_a, sizeof(double));
    if (!part) return ORC_ALLOC;
    int failed = 0;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
#endif
    for (long ii = 0; ii < (long)nocc_a; ++ii) {
        const size_t i = (size_t)ii;
        double *g = (double *)calloc(nvir_a * nov_b, sizeof(double));
        if (!g) {
            failed = 1;
            continue;
        }
        for (size_t p = 0; p < naux; ++p) {
            const double *ra = b_a + p * nov_a + i * nvir_a;
            const double *rb = b_b + p * nov_b;
            for (size_t a = 0; a < nvir_a; ++a) {
                const double s = ra[a];
                double *gr = g + a * nov_b;
                for (size_t c = 0; c < nov_b; ++c) gr[c] += s * rb[c];
            }
        }
        const double eps_i = eps_a[first_occ_a + i];
        double energy_i = 0.0;
        for (size_t a = 0; a < nvir_a; ++a) {
            const double eps_av = eps_a[nocc_total_a + a];
            for (size_t jj = 0; jj < nocc_b; ++jj) {
                const double eps_j = eps_b[first_occ_b + jj];
                for (size_t bb = 0; bb < nvir_b; ++bb) {
                    const double eps_bv = eps_b[nocc_total_b + bb];
                    const double eri = g[a * nov_b + jj * nvir_b + bb];
                    const double denom = eps_i + eps_j - eps_av - eps_bv;
                    const double t_val = eri / denom;
                    energy_i += t_val * eri;
                }
            }
        }
        part[i] = energy_i;
        free(g);
    }
    (void)nthreads;
    if (!failed)
        for (size_t i = 0; i < nocc_a; ++i) *energy += part[i];
    free(part);
    return failed ? ORC_ALLOC : ORC_OK;
}

/* signed_pair_block + spin_components_from_signed_b_ov (rimp2.rs:1559-1607):
 * G_i = pos_i^T pos_tail - neg_i^T neg_tail over the j >= i tail, then the
 * closed-shell pair loop and ascending fold of spin_components_from_blocks
 * (no kappa). pos is (naux_p, nocc*nvir), neg is (naux_n, nocc*nvir). */
int orc_signed_spin_components(const double *pos, size_t naux_p, const double *neg,
                               size_t naux_n, size_t nocc, size_t nvir, const double *eps,
                               size_t n_eps, size_t first_occ, size_t nocc_total, int nthreads,
                               double *out) {
    if (check_layout(naux_p, nocc, nvir, n_eps, first_occ, nocc_total)) return ORC_LAYOUT;
    out[0] = out[1] = out[2] = 0.0;
    if (nocc == 0 || nvir == 0) return ORC_OK;
    double *part = (double *)calloc(2 * nocc, sizeof(double));
    if (!part) return ORC_ALLOC;
    int failed = 0;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
#endif
    for (long ii = 0; ii < (long)nocc; ++ii) {
        const size_t i = (size_t)ii;
        const size_t n = nvir * (nocc - i) * nvir;
        double *g = (double *)malloc(sizeof(double) * n);
        double *gn = (double *)malloc(sizeof(double) * n);
        if (!g || !gn) {
            free(g);
            free(gn);
            failed = 1;
            continue;
        }
        orc_tail_block(pos, naux_p, nocc, nvir, i, g);
        orc_tail_block(neg, naux_n, nocc, nvir, i, gn);
        for (size_t k = 0; k < n; ++k) g[k] -= gn[k];
        orc_i_partial(g, i, nocc, nvir, eps, first_occ, nocc_total, 0, 0.0, &part[2 * i],
                      &part[2 * i + 1]);
        free(g);
        free(gn);
    }
    (void)nthreads;
    double os = 0.0, ss = 0.0;
    if (!failed)
        for (size_t i = 0; i < nocc; ++i) {
            os += part[2 * i];
            ss += part[2 * i + 1];
        }
    free(part);
    out[0] = os;
    out[1] = ss;
    out[2] = os + ss;
    return failed ? ORC_ALLOC : ORC_OK;
}

/* stream_dressed_mo_band_budgeted (rimp2.rs:666-756) with no budget: per aux
 * row Q, half = C_left^T (Q|mu nu) (nleft x nao), mo = half C_right
 * (nleft x nright); chunks of 256 rows of Q dressed into the output band with
 * beta = 1 accumulation, out[P, :] += m[P, Q-chunk] . mo-chunk. j3 is
 * (naux, nao, nao) row-major, c_left (nao, nleft), c_right (nao, nright), m
 * (nband, naux) -- the band's rows of the dressing matrix -- and out
 * (nband, nleft*nright), all row-major. */
int orc_dress_b_ov(const double *j3, size_t naux, size_t nao, const double *c_left, size_t nleft,
                   const double *c_right, size_t nright, const double *m, size_t nband,
                   double *out) {
    const size_t width = nleft * nright, chunk = 256;
    double *half = (double *)malloc(sizeof(double) * nleft * nao);
    double *mo = (double *)malloc(sizeof(double) * chunk * width);
    if (!half || !mo) {
        free(half);
        free(mo);
        return ORC_ALLOC;
    }
    memset(out, 0, sizeof(double) * nband * width);
    for (size_t q0 = 0; q0 < naux; q0 += chunk) {
        const size_t q1 = q0 + chunk < naux ? q0 + chunk : naux;
        for (size_t q = q0; q < q1; ++q) {
            const double *a = j3 + q * nao * nao;
            for (size_t l = 0; l < nleft; ++l)
                for (size_t nu = 0; nu < nao; ++nu) {
                    double s = 0.0;
                    for (size_t mu = 0; mu < nao; ++mu) s += c_left[mu * nleft + l] * a[mu * nao + nu];
                    half[l * nao + nu] = s;
                }
            double *row = mo + (q - q0) * width;
            for (size_t l = 0; l < nleft; ++l)
                for (size_t r = 0; r < nright; ++r) {
                    double s = 0.0;
                    for (size_t nu = 0; nu < nao; ++nu) s += half[l * nao + nu] * c_right[nu * nright + r];
                    row[l * nright + r] = s;
                }
        }
        for (size_t p = 0; p < nband; ++p)
            for (size_t q = q0; q < q1; ++q) {
                const double w = m[p * naux + q];
                const double *row = mo + (q - q0) * width;
                double *o = out + p * width;
                for (size_t c = 0; c < width; ++c) o[c] += w * row[c];
            }
    }
    free(half);
    free(mo);
    return ORC_OK;
}
