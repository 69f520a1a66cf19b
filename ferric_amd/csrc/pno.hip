// Pair natural orbitals (PNO) with truncation on the device: the per-pair body of
// dlpno_mp2_from_b_ov (crates/ferric-mp2/src/dlpno_mp2.rs:670-760) for any
// t_cut_pno >= 0. For every listed pair (i, j), batched over pairs:
//
//   g    = B_i^T B_j                                    (ia|jb), nvir x nvir
//   T    = g / (e_i + e_j - e_a - e_b)                  first-order amplitudes
//   D    = sym(T T^T + T^T T)                           pair density (local_pno.rs:146-154)
//   D    = W diag(n) W^T                                eigh; keep |n_k| >= t_cut_pno,
//                                                       never none (local_pno.rs:156-185)
//   F    = Q^T diag(e_vir) Q,  F = U diag(e~) U^T       semicanonical PNOs (dlpno_mp2.rs:281-309)
//   R    = Q U,  g~ = R^T g R
//   e_os = f sum_ab g~_ab^2 / d~_ab,  e_ss = f sum_ab g~_ab (g~_ab - g~_ba) / d~_ab
//          with d~_ab = e_i + e_j - e~_a - e~_b, f = 1 (i == j) or 2  (dlpno_mp2.rs:311-325)
//
// Every pair keeps the fixed nvir x nvir shape, so that one strided batch serves
// all of them: the discarded PNO columns of Q are zero, which makes the rows and
// columns of F that belong to them exactly zero; their diagonal is set to a
// shift above every virtual energy (the kept block's eigenvalues lie inside
// [min e_vir, max e_vir]), so the eigensolver keeps them apart from the kept
// block, and the energy skips them. The plain dense products run on rocBLAS and
// the two symmetric eigensolves on rocSOLVER (divide and conquer); the
// elementwise steps, the selection and the energy are the kernels below.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <mutex>
#include <vector>

#include "fh_device.h"
#include "fh_host.h"
#include "fh_kernels.h"

namespace fh {
namespace {

constexpr int kThreads = 256;

fh_status blas_ok(rocblas_status st, const char *what) {
    if (st == rocblas_status_success) return FH_OK;
    return fail(FH_KERNEL, "%s: rocBLAS/rocSOLVER status %d", what, (int)st);
}

// Device pointer arrays of the g GEMM: A[p] = B + i_p nvir, B[p] = B + j_p nvir,
// C[p] = the pair's g.
__global__ void pno_pointers_kernel(const int32_t *__restrict__ pairs, int np, const double *b,
                                    int nvir, double *g0, int64_t mat, const double **pa,
                                    const double **pb, double **pc) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    pa[p] = b + (int64_t)pairs[2 * p] * nvir;
    pb[p] = b + (int64_t)pairs[2 * p + 1] * nvir;
    pc[p] = g0 + (int64_t)p * mat;
}

// g of each pair from a dense row-major (nocc nvir)^2 matrix (the reference's
// dense-g entry, dlpno_mp2.rs:354-420): G[p][a + b nvir] = g[i nvir + a][j nvir + b].
__global__ __launch_bounds__(256) void pno_gather_g_kernel(const int32_t *__restrict__ pairs,
                                                           const double *__restrict__ g,
                                                           int64_t nov, int nvir,
                                                           double *__restrict__ out) {
    const int p = blockIdx.x;
    const int64_t mat = (int64_t)nvir * nvir;
    const double *src = g + (int64_t)pairs[2 * p] * nvir * nov + (int64_t)pairs[2 * p + 1] * nvir;
    double *dst = out + p * mat;
    // b runs fastest: the reads follow g's rows
    for (int64_t e = threadIdx.x; e < mat; e += 256) {
        const int a = (int)(e / nvir), b = (int)(e % nvir);
        dst[a + (int64_t)b * nvir] = src[(int64_t)a * nov + b];
    }
}

// T = g / D (column-major, element (a, b) at a + b nvir), one workgroup per pair.
__global__ __launch_bounds__(kThreads) void pno_amplitude_kernel(
    const int32_t *__restrict__ pairs, const double *__restrict__ eps_occ,
    const double *__restrict__ eps_vir, int nvir, const double *__restrict__ g,
    double *__restrict__ t) {
    const int p = blockIdx.x;
    const int64_t mat = (int64_t)nvir * nvir;
    const double eij = eps_occ[pairs[2 * p]] + eps_occ[pairs[2 * p + 1]];
    const double *gp = g + p * mat;
    double *tp = t + p * mat;
    for (int64_t e = threadIdx.x; e < mat; e += kThreads) {
        const int a = (int)(e % nvir), b = (int)(e / nvir);
        tp[e] = gp[e] / (eij - eps_vir[a] - eps_vir[b]);
    }
}

// D <- (D + D^T) / 2 in place, each (a, b), a < b, handled once.
__global__ __launch_bounds__(kThreads) void pno_symmetrize_kernel(int nvir, double *__restrict__ d) {
    double *dp = d + (int64_t)blockIdx.x * nvir * nvir;
    const int64_t mat = (int64_t)nvir * nvir;
    for (int64_t e = threadIdx.x; e < mat; e += kThreads) {
        const int a = (int)(e % nvir), b = (int)(e / nvir);
        if (a >= b) continue;
        const double v = 0.5 * (dp[a + (int64_t)b * nvir] + dp[b + (int64_t)a * nvir]);
        dp[a + (int64_t)b * nvir] = v;
        dp[b + (int64_t)a * nvir] = v;
    }
}

// Selection (local_pno.rs:156-185): occupations n (ascending from the
// eigensolver), eigenvectors W in the columns of w. The kept PNOs are those
// with |n_k| >= cut; when none passes, the largest occupation is kept (the
// first of equal ones in descending order, i.e. the lowest index among the
// largest values). They are written compacted, in descending order of n, to
// the leading columns of q (the rest of q is zero), with their count and the
// discarded weight (sum of |n_k| over the discarded k, in descending order).
__global__ __launch_bounds__(kThreads) void pno_select_kernel(
    int nvir, double cut, const double *__restrict__ occ, const double *__restrict__ w,
    double *__restrict__ q, int32_t *__restrict__ kept, double *__restrict__ discarded) {
    const int p = blockIdx.x;
    const int64_t mat = (int64_t)nvir * nvir;
    const double *n = occ + (int64_t)p * nvir;
    int mine = 0;
    for (int k = threadIdx.x; k < nvir; k += kThreads) mine += fabs(n[k]) >= cut ? 1 : 0;
    const int any = __syncthreads_or(mine);
    int top = -1;
    if (!any) {
        top = nvir - 1;
        while (top > 0 && n[top - 1] == n[nvir - 1]) --top;
    }
    __shared__ int npno_s;
    if (threadIdx.x == 0) {
        int cnt = 0;
        double dw = 0.0;
        for (int k = nvir - 1; k >= 0; --k) {
            if (fabs(n[k]) >= cut || k == top)
                ++cnt;
            else
                dw += fabs(n[k]);
        }
        npno_s = cnt;
        kept[p] = cnt;
        discarded[p] = dw;
    }
    __syncthreads();
    const int npno = npno_s;
    const double *wp = w + p * mat;
    double *qp = q + p * mat;
    // Column k of W goes to column slot(k) = #{kept j > k} of q; one wave per
    // column, its lanes counting over j and copying the rows.
    const int wid = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    for (int k = wid; k < nvir; k += kThreads / kWave) {
        if (!(fabs(n[k]) >= cut || k == top)) continue;  // uniform over the wave
        int c = 0;
        for (int j = k + 1 + lane; j < nvir; j += kWave) c += (fabs(n[j]) >= cut || j == top);
        for (int o = kWave / 2; o > 0; o >>= 1) c += __shfl_xor(c, o);
        double *dst = qp + (int64_t)c * nvir;
        const double *src = wp + (int64_t)k * nvir;
        for (int a = lane; a < nvir; a += kWave) dst[a] = src[a];
    }
    for (int64_t e = (int64_t)npno * nvir + threadIdx.x; e < mat; e += kThreads) qp[e] = 0.0;
}

// Gathers the pairs of one width class: qc[k] = the first c columns of q[idx[k]]
// (nvir x c), sc[k] = diag(e_vir) qc[k].
__global__ __launch_bounds__(kThreads) void pno_gather_kernel(
    int nvir, int c, const int32_t *__restrict__ idx, const double *__restrict__ q,
    const double *__restrict__ eps_vir, double *__restrict__ qc, double *__restrict__ sc) {
    const int k = blockIdx.x;
    const int64_t n = (int64_t)nvir * c;
    const double *src = q + (int64_t)idx[k] * nvir * nvir;
    for (int64_t e = threadIdx.x; e < n; e += kThreads) {
        const double v = src[e];
        qc[k * n + e] = v;
        sc[k * n + e] = eps_vir[e % nvir] * v;
    }
}

// F[j][j] = shift for the columns j >= kept of each pair's c x c block (their
// rows and columns are exactly zero).
__global__ __launch_bounds__(kThreads) void pno_shift_kernel(int c, double shift,
                                                            const int32_t *__restrict__ idx,
                                                            const int32_t *__restrict__ kept,
                                                            double *__restrict__ f) {
    const int k = blockIdx.x;
    double *fp = f + (int64_t)k * c * c;
    for (int j = kept[idx[k]] + threadIdx.x; j < c; j += kThreads) fp[j + (int64_t)j * c] = shift;
}

// Device pointer arrays of X = g R for one class: g of pair idx[k], R and X
// of class slot k (nvir x c each).
__global__ void pno_class_pointers_kernel(int m, const int32_t *__restrict__ idx, const double *g,
                                          int64_t mat, const double *r, double *x, int64_t vc,
                                          const double **pa, const double **pb, double **pc) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m) return;
    pa[k] = g + (int64_t)idx[k] * mat;
    pb[k] = r + (int64_t)k * vc;
    pc[k] = x + (int64_t)k * vc;
}

// Energy in the semicanonical PNO basis (dlpno_mp2.rs:311-325) of the class's
// pairs: gt = g~ (c x c, column-major), et the semicanonical energies;
// directions with e~ above `skip` are padding and contribute nothing.
__global__ __launch_bounds__(kThreads) void pno_energy_kernel(
    const int32_t *__restrict__ pairs, const int32_t *__restrict__ idx,
    const double *__restrict__ eps_occ, int c, double skip, const double *__restrict__ et,
    const double *__restrict__ gt, double *__restrict__ out) {
    const int k = blockIdx.x, p = idx[k];
    const int i = pairs[2 * p], j = pairs[2 * p + 1];
    const double eij = eps_occ[i] + eps_occ[j];
    const int64_t mat = (int64_t)c * c;
    const double *e = et + (int64_t)k * c;
    const double *g = gt + k * mat;
    KSum os, ss;
    for (int64_t x = threadIdx.x; x < mat; x += kThreads) {
        const int a = (int)(x % c), b = (int)(x / c);
        if (e[a] > skip || e[b] > skip) continue;
        const double d = eij - e[a] - e[b];
        const double gab = g[x], gba = g[b + (int64_t)a * c];
        os.add(gab * gab / d);
        ss.add(gab * (gab - gba) / d);
    }
    wave_merge(os);
    wave_merge(ss);
    __shared__ double red[kThreads / kWave][4];
    const int wid = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    if (lane == 0) {
        red[wid][0] = os.s;
        red[wid][1] = os.c;
        red[wid][2] = ss.s;
        red[wid][3] = ss.c;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        KSum o, s;
        for (int w = 0; w < kThreads / kWave; ++w) {
            o.merge(red[w][0], red[w][1]);
            s.merge(red[w][2], red[w][3]);
        }
        const double f = i == j ? 1.0 : 2.0;
        out[2 * p] = f * o.value();
        out[2 * p + 1] = f * s.value();
    }
}

// Width class of a pair with n kept PNOs: the smallest of 128, 256, 512, ...
// that holds n, or nvir. Below 128 the batched divide-and-conquer eigensolver
// was measured slower than at 128 (n = 64 took 3-13x the time of n = 128 for
// the same pairs), so no class is narrower.
int pno_class(int n, int nvir) {
    int c = 128;
    while (c < n) c *= 2;
    return std::min(c, nvir);
}

// Pairs per batch: the five nvir x nvir matrices of a pair in about 4 GiB
// (FERRIC_PNO_BATCH overrides).
size_t pno_batch(size_t npairs, size_t nvir) {
    size_t nb = 0;
    if (const char *v = std::getenv("FERRIC_PNO_BATCH")) nb = (size_t)std::strtoull(v, nullptr, 10);
    if (nb == 0) {
        const size_t per = 5 * nvir * nvir * sizeof(double) + 64 * nvir;
        nb = std::max<size_t>(1, ((size_t)4 << 30) / per);
    }
    return std::min(nb, npairs);
}

// b: B (naux, nocc nvir) row-major, or, when dense_g, the dense g of
// spin_components_from_g ((nocc nvir)^2 row-major; naux unused).
fh_status pno_core(State &s, const double *b, bool dense_g, size_t naux, size_t nocc, size_t nvir,
                   const double *eps, size_t first_occ, size_t nocc_total,
                   const int32_t *pairs, size_t npairs, double t_cut_pno, hipStream_t stream,
                   double *e_os, double *e_ss, fh_pno_diagnostics *diag) {
    // The largest denominator of pair (i, j) is e_i + e_j - 2 min e_vir.
    double emax = -INFINITY, vmin = INFINITY, vabs = 0.0;
    for (size_t k = 0; k < npairs; ++k) {
        const double eij = eps[first_occ + pairs[2 * k]] + eps[first_occ + pairs[2 * k + 1]];
        emax = std::isnan(eij) ? NAN : std::max(emax, eij);
        if (std::isnan(emax)) break;
    }
    for (size_t a = 0; a < nvir; ++a) {
        const double v = eps[nocc_total + a];
        if (!std::isfinite(v)) return fail(FH_LAYOUT, "non-finite virtual orbital energy %g", v);
        vmin = std::min(vmin, v);
        vabs = std::max(vabs, std::fabs(v));
    }
    if (!(emax - 2.0 * vmin < 0.0))
        return fail(FH_LAYOUT,
                    "non-negative MP2 denominator %g (max e_i + e_j %g, min eps_vir %g) in the "
                    "listed pairs",
                    emax - 2.0 * vmin, emax, vmin);
    // Above every kept semicanonical energy by at least max|e_vir| + 1; the
    // energy skips directions above half of that gap.
    const double shift = 2.0 * vabs + 2.0;
    const double skip = vabs + 1.0;

    const size_t nb = pno_batch(npairs, nvir);
    const size_t mat = nvir * nvir;
    const size_t small = 3 * nvir;  // occupations, semicanonical energies, eigensolver scratch
    fh_status st = reserve(
        s, {{kSlotPnoWork, sizeof(double) * nb * (5 * mat + small)},
            {kSlotPnoMeta, sizeof(int32_t) * 2 * npairs + sizeof(double) * 3 * npairs +
                               sizeof(int32_t) * (npairs + 3 * nb) + 3 * sizeof(void *) * nb +
                               sizeof(double) * (nocc + nvir) + 256}});
    if (st != FH_OK) return st;
    double *G = (double *)s.slots[kSlotPnoWork].ptr;
    double *M = G + nb * mat;
    double *Q = M + nb * mat;
    double *S = Q + nb * mat;
    double *F = S + nb * mat;
    double *Wd = F + nb * mat;   // occupations
    double *We = Wd + nb * nvir;  // semicanonical energies
    double *Ex = We + nb * nvir;  // eigensolver off-diagonal scratch
    char *mp = (char *)s.slots[kSlotPnoMeta].ptr;
    double *out = (double *)mp;  // e_os, e_ss per pair
    double *dw = out + 2 * npairs;
    double *deps = dw + npairs;
    int32_t *dpairs = (int32_t *)(deps + nocc + nvir);
    int32_t *kept = dpairs + 2 * npairs;
    int32_t *info = kept + npairs;  // first eigensolve, per pair of the batch
    int32_t *info2 = info + nb;     // second eigensolve, per pair of a class
    int32_t *cls = info2 + nb;      // batch indices of the pairs of a class
    const size_t off = (size_t)((char *)(cls + nb) - mp);
    const double **pa = (const double **)(mp + (off + 7) / 8 * 8);
    const double **pb = pa + nb;
    double **pc = (double **)(pb + nb);

    std::vector<double> he(nocc + nvir);
    for (size_t k = 0; k < nocc; ++k) he[k] = eps[first_occ + k];
    for (size_t a = 0; a < nvir; ++a) he[nocc + a] = eps[nocc_total + a];
    st = hip_check(hipMemcpyAsync(deps, he.data(), sizeof(double) * he.size(),
                                  hipMemcpyHostToDevice, stream),
                   "H2D eps");
    if (st != FH_OK) return st;
    st = hip_check(hipMemcpyAsync(dpairs, pairs, sizeof(int32_t) * 2 * npairs,
                                  hipMemcpyHostToDevice, stream),
                   "H2D pairs");
    if (st != FH_OK) return st;
    s.pending.h2d_bytes += sizeof(double) * he.size() + sizeof(int32_t) * 2 * npairs;

    void *hv = nullptr;
    st = blas_handle(s, stream, &hv);
    if (st != FH_OK) return st;
    rocblas_handle h = (rocblas_handle)hv;
    const double one = 1.0, zero = 0.0;
    const int V = (int)nvir, nov = (int)(nocc * nvir);
    const rocblas_stride sm = (rocblas_stride)mat;
    const auto N = rocblas_operation_none, T = rocblas_operation_transpose;
    std::vector<int32_t> hinfo(nb);
    mark(s, 0, stream);
    mark(s, 1, stream);
    for (size_t p0 = 0; p0 < npairs; p0 += nb) {
        const int n = (int)std::min(nb, npairs - p0);
        const int32_t *bp = dpairs + 2 * p0;
        if (dense_g) {
            hipLaunchKernelGGL(pno_gather_g_kernel, dim3(n), dim3(256), 0, stream, bp, b,
                               (int64_t)nov, V, G);
        } else {
            hipLaunchKernelGGL(pno_pointers_kernel, dim3((n + 63) / 64), dim3(64), 0, stream, bp,
                               n, b, V, G, (int64_t)mat, pa, pb, pc);
            // g = B_i^T B_j: column-major, B is (nov x naux) with leading dimension nov.
            st = blas_ok(rocblas_dgemm_batched(h, N, T, V, V, (int)naux, &one, pa, nov, pb, nov,
                                               &zero, pc, V, n),
                         "PNO g GEMM");
            if (st != FH_OK) return st;
        }
        hipLaunchKernelGGL(pno_amplitude_kernel, dim3(n), dim3(kThreads), 0, stream, bp, deps,
                           deps + nocc, V, G, Q);
        // D = T T^T + T^T T
        st = blas_ok(rocblas_dgemm_strided_batched(h, N, T, V, V, V, &one, Q, V, sm, Q, V, sm,
                                                   &zero, M, V, sm, n),
                     "PNO density GEMM 1");
        if (st != FH_OK) return st;
        st = blas_ok(rocblas_dgemm_strided_batched(h, T, N, V, V, V, &one, Q, V, sm, Q, V, sm,
                                                   &one, M, V, sm, n),
                     "PNO density GEMM 2");
        if (st != FH_OK) return st;
        hipLaunchKernelGGL(pno_symmetrize_kernel, dim3(n), dim3(kThreads), 0, stream, V, M);
        st = blas_ok(rocsolver_dsyevd_strided_batched(h, rocblas_evect_original,
                                                      rocblas_fill_upper, V, M, V, sm, Wd, V,
                                                      Ex, V, info, n),
                     "PNO eigensolver");
        if (st != FH_OK) return st;
        hipLaunchKernelGGL(pno_select_kernel, dim3(n), dim3(kThreads), 0, stream, V, t_cut_pno,
                           Wd, M, Q, kept + p0, dw + p0);
        st = hip_check(hipGetLastError(), "PNO kernels");
        if (st != FH_OK) return st;
        // The kept counts set each pair's width class: the rest of the pair's
        // work runs at c x c (c >= kept) instead of nvir x nvir.
        std::vector<int32_t> hk(n);
        st = hip_check(hipMemcpyAsync(hk.data(), kept + p0, sizeof(int32_t) * n,
                                      hipMemcpyDeviceToHost, stream),
                       "D2H kept counts");
        if (st != FH_OK) return st;
        st = hip_check(hipMemcpyAsync(hinfo.data(), info, sizeof(int32_t) * n,
                                      hipMemcpyDeviceToHost, stream),
                       "D2H eigensolver info");
        if (st != FH_OK) return st;
        st = hip_check(hipStreamSynchronize(stream), "stream sync");
        if (st != FH_OK) return st;
        for (int k = 0; k < n; ++k)
            if (hinfo[k] != 0)
                return fail(FH_KERNEL, "PNO eigensolver failed for pair (%d, %d): info %d",
                            pairs[2 * (p0 + k)], pairs[2 * (p0 + k) + 1], hinfo[k]);
        std::vector<std::vector<int32_t>> byc(V + 1);
        for (int k = 0; k < n; ++k) byc[pno_class(hk[k], V)].push_back(k);
        for (int c = 1; c <= V; ++c) {
            const std::vector<int32_t> &idx = byc[c];
            const int m = (int)idx.size();
            if (m == 0) continue;
            const rocblas_stride svc = (rocblas_stride)V * c, scc = (rocblas_stride)c * c;
            st = hip_check(hipMemcpyAsync(cls, idx.data(), sizeof(int32_t) * m,
                                          hipMemcpyHostToDevice, stream),
                           "H2D class");
            if (st != FH_OK) return st;
            // QC (in M) and SC = diag(e_vir) QC (in S), nvir x c per pair
            hipLaunchKernelGGL(pno_gather_kernel, dim3(m), dim3(kThreads), 0, stream, V, c, cls,
                               Q, deps + nocc, M, S);
            // F = QC^T diag(e_vir) QC (in F), c x c
            st = blas_ok(rocblas_dgemm_strided_batched(h, T, N, c, c, V, &one, M, V, svc, S, V,
                                                       svc, &zero, F, c, scc, m),
                         "PNO Fock GEMM");
            if (st != FH_OK) return st;
            hipLaunchKernelGGL(pno_shift_kernel, dim3(m), dim3(kThreads), 0, stream, c, shift,
                               cls, kept + p0, F);
            st = blas_ok(rocsolver_dsyevd_strided_batched(h, rocblas_evect_original,
                                                          rocblas_fill_upper, c, F, c, scc, We,
                                                          c, Ex, c, info2, m),
                         "PNO semicanonical eigensolver");
            if (st != FH_OK) return st;
            // R = QC U (into S); X = g R (into M); g~ = R^T X (into F)
            st = blas_ok(rocblas_dgemm_strided_batched(h, N, N, V, c, c, &one, M, V, svc, F, c,
                                                       scc, &zero, S, V, svc, m),
                         "PNO rotation GEMM");
            if (st != FH_OK) return st;
            hipLaunchKernelGGL(pno_class_pointers_kernel, dim3((m + 63) / 64), dim3(64), 0,
                               stream, m, cls, G, (int64_t)mat, S, M, (int64_t)svc, pa, pb, pc);
            st = blas_ok(rocblas_dgemm_batched(h, N, N, V, c, V, &one, pa, V, pb, V, &zero, pc,
                                               V, m),
                         "PNO transform GEMM 1");
            if (st != FH_OK) return st;
            st = blas_ok(rocblas_dgemm_strided_batched(h, T, N, c, c, V, &one, S, V, svc, M, V,
                                                       svc, &zero, F, c, scc, m),
                         "PNO transform GEMM 2");
            if (st != FH_OK) return st;
            hipLaunchKernelGGL(pno_energy_kernel, dim3(m), dim3(kThreads), 0, stream, bp, cls,
                               deps, c, skip, We, F, out + 2 * p0);
            st = hip_check(hipGetLastError(), "PNO kernels");
            if (st != FH_OK) return st;
            s.pending.kernel_launches += 9;
            std::vector<int32_t> hi2(m);
            st = hip_check(hipMemcpyAsync(hi2.data(), info2, sizeof(int32_t) * m,
                                          hipMemcpyDeviceToHost, stream),
                           "D2H eigensolver info");
            if (st != FH_OK) return st;
            st = hip_check(hipStreamSynchronize(stream), "stream sync");
            if (st != FH_OK) return st;
            for (int k = 0; k < m; ++k)
                if (hi2[k] != 0)
                    return fail(FH_KERNEL,
                                "PNO semicanonical eigensolver failed for pair (%d, %d): info %d",
                                pairs[2 * (p0 + idx[k])], pairs[2 * (p0 + idx[k]) + 1], hi2[k]);
        }
        s.pending.kernel_launches += 8;
    }
    mark(s, 2, stream);
    std::vector<double> ho(3 * npairs);
    std::vector<int32_t> hk(npairs);
    st = hip_check(hipMemcpyAsync(ho.data(), out, sizeof(double) * ho.size(),
                                  hipMemcpyDeviceToHost, stream),
                   "D2H PNO energies");
    if (st != FH_OK) return st;
    st = hip_check(hipMemcpyAsync(hk.data(), kept, sizeof(int32_t) * npairs,
                                  hipMemcpyDeviceToHost, stream),
                   "D2H PNO counts");
    if (st != FH_OK) return st;
    st = hip_check(hipStreamSynchronize(stream), "stream sync");
    if (st != FH_OK) return st;
    mark(s, 3, stream);
    s.pending.d2h_bytes += sizeof(double) * ho.size() + sizeof(int32_t) * npairs;
    uint64_t total = 0;
    double worst = 0.0;
    for (size_t k = 0; k < npairs; ++k) {
        e_os[k] = ho[2 * k];
        e_ss[k] = ho[2 * k + 1];
        total += (uint64_t)hk[k];
        worst = std::max(worst, ho[2 * npairs + k]);
    }
    if (diag) {
        diag->kept_pnos = total;
        diag->max_discarded_weight = worst;
    }
    finish_timing(s, (uint64_t)npairs, 0);
    return FH_OK;
}

fh_status pno_checks(size_t naux, size_t nocc, size_t nvir, const double *eps, size_t n_eps,
                     size_t first_occ, size_t nocc_total, const int32_t *pairs, size_t npairs,
                     double t_cut_pno, const double *e_os, const double *e_ss, bool *empty) {
    if (!e_os || !e_ss) return fail(FH_LAYOUT, "null output pointer");
    if (t_cut_pno < 0.0)
        return fail(FH_LAYOUT, "t_cut_pno must be >= 0 (got %g); use 0.0 to disable truncation",
                    t_cut_pno);
    fh_status st = check_shapes(naux, nocc, nvir, n_eps, first_occ, nocc_total, empty);
    if (st != FH_OK) return st;
    if (npairs > 0 && !pairs) return fail(FH_LAYOUT, "null pair list");
    for (size_t k = 0; k < npairs; ++k) {
        const int32_t i = pairs[2 * k], j = pairs[2 * k + 1];
        if (i < 0 || i > j || (size_t)j >= nocc)
            return fail(FH_LAYOUT, "pair %zu is (%d, %d): need 0 <= i <= j < nocc = %zu", k, i,
                        j, nocc);
    }
    if (!*empty && npairs > 0 && !eps) return fail(FH_LAYOUT, "null eps");
    if (nvir > 46340) return fail(FH_LAYOUT, "nvir %zu too large for one PNO block", nvir);
    if (nocc * nvir > (size_t)INT32_MAX) return fail(FH_LAYOUT, "nocc*nvir beyond int32");
    return FH_OK;
}

}  // namespace
}  // namespace fh

using namespace fh;

extern "C" fh_status fh_dlpno_pno_pair_energies_dev(
    const double *b_dev, size_t naux, size_t nocc, size_t nvir, const double *eps, size_t n_eps,
    size_t first_occ, size_t nocc_total, const int32_t *pairs, size_t npairs, double t_cut_pno,
    void *stream, double *e_os, double *e_ss, fh_pno_diagnostics *diag) {
    State &s = state();
    std::lock_guard<std::mutex> lk(s.mu);
    clear_error();
    s.stats.calls++;
    s.pending = Pending{};
    bool empty = false;
    fh_status st = pno_checks(naux, nocc, nvir, eps, n_eps, first_occ, nocc_total, pairs, npairs,
                              t_cut_pno, e_os, e_ss, &empty);
    if (st != FH_OK) return st;
    if (diag) *diag = fh_pno_diagnostics{0, 0.0};
    if (empty || npairs == 0) {
        for (size_t k = 0; k < npairs; ++k) e_os[k] = e_ss[k] = 0.0;
        commit(s, Prec::F64, false, false);
        return FH_OK;
    }
    if (!b_dev) return fail(FH_LAYOUT, "null input pointer");
    st = ensure_device(s);
    if (st != FH_OK) return st;
    st = pno_core(s, b_dev, false, naux, nocc, nvir, eps, first_occ, nocc_total, pairs, npairs,
                  t_cut_pno, (hipStream_t)stream, e_os, e_ss, diag);
    if (st != FH_OK) return st;
    commit(s, Prec::F64, false, false);
    return FH_OK;
}

extern "C" fh_status fh_dlpno_pno_pair_energies(const double *b_ov, size_t naux, size_t nocc,
                                                size_t nvir, const double *eps, size_t n_eps,
                                                size_t first_occ, size_t nocc_total,
                                                const int32_t *pairs, size_t npairs,
                                                double t_cut_pno, double *e_os, double *e_ss,
                                                fh_pno_diagnostics *diag) {
    State &s = state();
    std::lock_guard<std::mutex> lk(s.mu);
    clear_error();
    s.stats.calls++;
    s.pending = Pending{};
    bool empty = false;
    fh_status st = pno_checks(naux, nocc, nvir, eps, n_eps, first_occ, nocc_total, pairs, npairs,
                              t_cut_pno, e_os, e_ss, &empty);
    if (st != FH_OK) return st;
    if (diag) *diag = fh_pno_diagnostics{0, 0.0};
    if (empty || npairs == 0) {
        for (size_t k = 0; k < npairs; ++k) e_os[k] = e_ss[k] = 0.0;
        commit(s, Prec::F64, false, false);
        return FH_OK;
    }
    if (!b_ov) return fail(FH_LAYOUT, "null input pointer");
    st = ensure_device(s);
    if (st != FH_OK) return st;
    const size_t bbytes = sizeof(double) * naux * nocc * nvir;
    st = reserve(s, {{kSlotStage, bbytes}});
    if (st != FH_OK) return st;
    st = hip_check(hipMemcpyAsync(s.slots[kSlotStage].ptr, b_ov, bbytes, hipMemcpyHostToDevice,
                                  nullptr),
                   "H2D b_ov");
    if (st != FH_OK) return st;
    s.pending.h2d_bytes += bbytes;
    st = pno_core(s, (const double *)s.slots[kSlotStage].ptr, false, naux, nocc, nvir, eps,
                  first_occ, nocc_total, pairs, npairs, t_cut_pno, nullptr, e_os, e_ss, diag);
    if (st != FH_OK) return st;
    commit(s, Prec::F64, false, false);
    return FH_OK;
}

extern "C" fh_status fh_dlpno_pno_pair_energies_from_g_dev(
    const double *g_dev, size_t nocc, size_t nvir, const double *eps, size_t n_eps,
    size_t first_occ, size_t nocc_total, const int32_t *pairs, size_t npairs, double t_cut_pno,
    void *stream, double *e_os, double *e_ss, fh_pno_diagnostics *diag) {
    State &s = state();
    std::lock_guard<std::mutex> lk(s.mu);
    clear_error();
    s.stats.calls++;
    s.pending = Pending{};
    bool empty = false;
    fh_status st = pno_checks(1, nocc, nvir, eps, n_eps, first_occ, nocc_total, pairs, npairs,
                              t_cut_pno, e_os, e_ss, &empty);
    if (st != FH_OK) return st;
    if (diag) *diag = fh_pno_diagnostics{0, 0.0};
    if (empty || npairs == 0) {
        for (size_t k = 0; k < npairs; ++k) e_os[k] = e_ss[k] = 0.0;
        commit(s, Prec::F64, false, false);
        return FH_OK;
    }
    if (!g_dev) return fail(FH_LAYOUT, "null input pointer");
    st = ensure_device(s);
    if (st != FH_OK) return st;
    st = pno_core(s, g_dev, true, 0, nocc, nvir, eps, first_occ, nocc_total, pairs, npairs,
                  t_cut_pno, (hipStream_t)stream, e_os, e_ss, diag);
    if (st != FH_OK) return st;
    commit(s, Prec::F64, false, false);
    return FH_OK;
}
