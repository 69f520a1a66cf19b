/*
 * ferric_hip.h -- C ABI of the MI355X (gfx950) RI-MP2 pair-energy library.
 *
 * Drop-in for the device boundary of ferric's closed-shell RI-MP2 stage:
 *
 *   try_spin_components_on_device(b_ov, eps, nocc, nvir, first_occ,
 *                                 nocc_total, kappa, mixed_ok)
 *       -> Option<SpinComponents>
 *   (crates/ferric-mp2/src/rimp2_gpu.rs:105-114; called by
 *    dispatch_spin_components, crates/ferric-mp2/src/rimp2.rs:1360-1375)
 *
 * Conventions kept from the reference:
 *   - b_ov is row-major (naux, nocc*nvir): b_ov[P*nocc*nvir + i*nvir + a] = B^P_ia,
 *     over ACTIVE occupied i (rimp2.rs:1445-1499).
 *   - eps is the full MO energy vector: occupied i at eps[first_occ + i],
 *     virtual a at eps[nocc_total + a]; first_occ is the frozen-core count
 *     (rimp2.rs:1399-1441, :99-106).
 *   - Shape checks first, in the reference's order (rimp2_gpu.rs:211-237):
 *     nocc*nvir overflow, eps too short -> FH_LAYOUT; nocc == 0 || nvir == 0
 *     -> zeros and FH_OK; naux == 0 -> FH_LAYOUT (the CPU path owns it).
 *   - kappa: NaN means None (plain MP2); otherwise it must be finite and > 0
 *     (rimp2.rs:1266-1272) and every term is damped by (1 - e^{kappa D})^2.
 *   - Never aborts. Every failure is a typed status mirroring GpuError
 *     (crates/ferric-core/src/gpu/device.rs:16-36); the caller runs its own
 *     CPU path on any status other than FH_OK. This library has no CPU path.
 *   - Calls are serialised per device (one mutex per device). The library
 *     allocates its device workspace itself against a pool ceiling
 *     (FERRIC_GPU_MEM_GB, default 80% of free device memory,
 *     crates/ferric-core/src/gpu/mod.rs:348-353) and refuses with
 *     FH_POOL_FULL before any transfer when a call would exceed it.
 *
 * Every pointer argument is a plain host or device pointer as stated per
 * entry point; no entry point takes or returns a framework type.
 */
#ifndef FERRIC_HIP_H
#define FERRIC_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FERRIC_HIP_ABI_VERSION 11

/* GpuError classes (crates/ferric-core/src/gpu/device.rs:16-36). */
typedef enum fh_status {
    FH_OK = 0,
    FH_LAYOUT = 1,      /* GpuError::Layout: shapes/offsets inconsistent */
    FH_POOL_FULL = 2,   /* GpuError::PoolFull: workspace exceeds the pool ceiling */
    FH_DEVICE = 3,      /* GpuError::Device: no usable gfx950 device */
    FH_HIP = 4,         /* GpuError::Cuda: a HIP runtime call failed */
    FH_KERNEL = 5,      /* GpuError::Kernel: a kernel could not be launched */
    FH_RANGE = 6,       /* GpuError::F32Range: operand outside the reduced format's range */
    FH_UNSUPPORTED = 7, /* option outside this library's surface (e.g. kappa in the adjoint) */
} fh_status;

/* Arithmetic of the (ia|jb) contraction. The orbital-energy denominators and
 * the pair-energy sums are FP64 (compensated) in every mode.
 *   FH_FP64     FP64 MFMA, FP64 accumulation (the plain method)
 *   FH_FP32     FP32 operands, FP32 MFMA, FP32 accumulation over the whole Q range
 *   FH_MIXED    FP32 operands (RN from the caller's values), FP32 MFMA over Q
 *               panels of at most 64 rows, each panel's sums added into FP64
 *               accumulators: the reference's Precision::Mixed arithmetic
 *               (f32 storage, b = 64 f32 panel products, f64 panel sum;
 *               crates/ferric-core/src/gpu/mixed.rs:70,122-180); refused
 *               (FH_RANGE) when max|B| > sqrt(FLT_MAX / 64), as the reference's
 *               check_f32_arithmetic_range with b = 64
 *   FH_BF16X3   BF16 hi/lo split, hi*hi + hi*lo + lo*hi, FP32 accumulation
 *   FH_FP8_E4M3 OCP E4M3 operands with one power-of-two scale per
 *               (64-row Q block, occupied i), FP32 accumulation
 *   FH_INT8X4   every element as four signed byte digits of Y = RN(x 2^(30-e)),
 *               |Y| < 2^31, e per (i, a) row from max_Q |B^Q_ia|; the 13 digit
 *               products of levels 2..6 accumulated exactly in int32 (INT8
 *               MFMA) and combined in FP64: only the rounding of Y (2^-32 of
 *               the row's power-of-two range) and the dropped levels depart
 *               from FP64. naux up to 32832 (FH_RANGE beyond).
 * The reference's Precision::{F64, Mixed} (crates/ferric-core/src/gpu/precision.rs)
 * map to FH_FP64 and FH_MIXED. */
typedef enum fh_precision {
    FH_FP64 = 0,
    FH_FP32 = 1,
    FH_BF16X3 = 2,
    FH_FP8_E4M3 = 3,
    FH_INT8X4 = 4,
    FH_MIXED = 5,
} fh_precision;

/* Element type of a device-resident b_ov. */
typedef enum fh_dtype {
    FH_DT_F64 = 0,
    FH_DT_F32 = 1,
} fh_dtype;

/* SpinComponents (crates/ferric-mp2/src/rimp2.rs:134-141). */
typedef struct fh_spin_components {
    double e_os;
    double e_ss;
    double e_total; /* e_os + e_ss */
} fh_spin_components;

/* Counters in the spirit of GpuStatsSnapshot
 * (crates/ferric-core/src/gpu/stats.rs:308-350). Committed only when a call
 * succeeds as a whole, as the reference's Tally (rimp2_gpu.rs:86-100). */
typedef struct fh_stats {
    uint64_t calls;            /* entry-point calls */
    uint64_t offloaded;        /* calls that returned FH_OK */
    uint64_t fp64_calls;
    uint64_t fp32_calls;
    uint64_t bf16x3_calls;
    uint64_t fp8_calls;
    uint64_t adjoint_calls;
    uint64_t batched_calls;
    uint64_t h2d_bytes;
    uint64_t d2h_bytes;
    uint64_t kernel_launches;
    uint64_t fail_layout;      /* failures per class (counted whether or not the call succeeded) */
    uint64_t fail_pool_full;
    uint64_t fail_hip;
    uint64_t fail_range;
    uint64_t fail_unsupported;
    uint64_t pool_peak_bytes;  /* high-water mark of the library's workspace */
    uint64_t int8x4_calls;
    uint64_t mixed_calls;
} fh_stats;

/* One local-MP2 pair domain for fh_rimp2_batched_dev: its B (row-major
 * (naux, nocc*nvir)) starts at element b_offset of the packed B buffer and its
 * orbital energies at eps_offset of the packed eps buffer, laid out
 * [nocc occupied, then nvir virtual] (no frozen core inside a domain). */
typedef struct fh_domain {
    uint64_t b_offset;
    uint64_t eps_offset;
    uint32_t nocc;
    uint32_t nvir;
    uint32_t naux;
    uint32_t reserved;
} fh_domain;

/* Phase times of the last successful call on this device, measured with HIP
 * events on the call's stream (the reference's FERRIC_TIMING phases,
 * crates/ferric-core/src/timing.rs:161). Recorded only while timing is on
 * (fh_set_timing(1) or FERRIC_TIMING=1 at first use); otherwise zeros. */
typedef struct fh_timing {
    double pack_ms;      /* operand packing / quantisation */
    double contract_ms;  /* fused contraction + pair-energy kernel(s) */
    double reduce_ms;    /* deterministic reductions */
    double total_ms;     /* first to last event */
    uint64_t items;      /* work items of the contraction launch */
    uint32_t tile;       /* virtual tile side of the contraction kernel */
    uint32_t reserved;
} fh_timing;

/* ---- Library state ---------------------------------------------------- */

/* ABI version (FERRIC_HIP_ABI_VERSION). */
int fh_abi_version(void);

/* Select and initialise a device (install-once settings,
 * crates/ferric-core/src/gpu/mod.rs:145-207). device < 0 takes
 * FERRIC_GPU_DEVICE or 0. Idempotent per device. */
fh_status fh_init(int device);

/* Message of the last failure on the calling thread ("" if none). */
const char *fh_last_error(void);

/* Pool ceiling in bytes for the library's device workspace (0 restores the
 * default: FERRIC_GPU_MEM_GB if set, else 80% of free device memory). */
fh_status fh_set_pool_limit(size_t bytes);
fh_status fh_get_pool_limit(size_t *bytes);

/* Free the cached device workspace (it is otherwise kept between calls). */
fh_status fh_release_workspace(void);

fh_status fh_get_stats(fh_stats *out);
fh_status fh_reset_stats(void);

/* Tile table (no counterpart in the reference: this library's own tuning).
 * Calls whose (precision, nvir, naux) fall inside an installed bucket use its
 * tile side (32, 64 or 128); the first matching entry in installation order
 * wins; FERRIC_HIP_TILE overrides every entry. ferric_amd/autotune.py measures
 * the table and the Python package installs it when the library loads. */
fh_status fh_tune_set(fh_precision precision, size_t nvir_max, size_t naux_max, int tile);
fh_status fh_tune_clear(void);

/* Phase timing on/off (see fh_timing) and the last call's phase times. */
fh_status fh_set_timing(int on);
fh_status fh_last_timing(fh_timing *out);

/* ---- Energy ------------------------------------------------------------ */

/* The drop-in for try_spin_components_on_device (rimp2_gpu.rs:105-170):
 * b_ov and eps are HOST pointers (borrowed, read-only); the library makes a
 * resident device copy for the call (Stage, rimp2_gpu.rs:312-373). */
fh_status fh_rimp2_energy(const double *b_ov, size_t naux, size_t nocc, size_t nvir,
                          const double *eps, size_t n_eps, size_t first_occ,
                          size_t nocc_total, double kappa, fh_precision precision,
                          fh_spin_components *out);

/* Same energy with b_ov already resident on the device (b_dev, element type
 * b_dtype, row-major (naux, nocc*nvir)); eps stays a host pointer. Work is
 * queued on `stream` (a hipStream_t, NULL = the null stream) and the call
 * returns after the result is on the host. row_partials (host, 2*nocc doubles,
 * may be NULL) receives the per-i partials (e_os_i, e_ss_i), each already
 * carrying the pair weights, whose ascending sum is the energy
 * (spin_components_from_blocks, rimp2.rs:1506-1544). */
fh_status fh_rimp2_energy_dev(const void *b_dev, fh_dtype b_dtype, size_t naux, size_t nocc,
                              size_t nvir, const double *eps, size_t n_eps, size_t first_occ,
                              size_t nocc_total, double kappa, fh_precision precision,
                              void *stream, fh_spin_components *out, double *row_partials);

/* Bytes of the fixed device workspace of fh_rimp2_energy_dev_noalloc
 * (128 MiB per device, held from the library's load until the process ends). */
#define FH_NOALLOC_SCRATCH_BYTES (128u * 1024u * 1024u)

/* fh_rimp2_energy_dev with no device allocation: the same arguments, kernels
 * and results (bit for bit), but the workspace (packed operand, per-item and
 * per-row partials, orbital energies, result) is carved from a fixed
 * FH_NOALLOC_SCRATCH_BYTES device array instead of the library's pool; a call
 * whose workspace does not fit is FH_POOL_FULL before any device work. No
 * counterpart in the reference: this entry exists for callers that cannot
 * allocate device memory (code under graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). All calls on one device share the array:
 * calls that overlap on one device race (the library serialises its own
 * calls; do not run this entry concurrently from several processes or
 * streams of one device). FP64 at 41/209/960 needs about 71 MB. */
fh_status fh_rimp2_energy_dev_noalloc(const void *b_dev, fh_dtype b_dtype, size_t naux,
                                      size_t nocc, size_t nvir, const double *eps, size_t n_eps,
                                      size_t first_occ, size_t nocc_total, double kappa,
                                      fh_precision precision, void *stream,
                                      fh_spin_components *out, double *row_partials);

/* ---- Adjoint ----------------------------------------------------------- */

/* E = c_os*E_os + c_ss*E_ss (SCS surface, crates/ferric-mp2/src/scs.rs:93-114)
 * and its adjoint:
 *   gamma[P*nov + i*nvir + a] = dE/dB^P_ia = 4 x[P, ia], with
 *     x[P, ia] = sum_jb T~_ij^ab B^P_jb, T~ = (c_os + c_ss) t - c_ss t^T
 *     (zvector.rs:651-697, rimp2.rs:1832-1840, :1987-2011);
 *   deps[first_occ + k] = dE/deps_k = 2 P_kk (active occupied),
 *   deps[nocc_total + c] = 2 P_cc (virtual), 0 elsewhere (frozen core)
 *     (oo_rimp2.rs:1192-1300, rimp2.rs:2026-2052).
 * FP64 and BF16X3 recompute T~ per occupied row and never store it for the whole
 * problem; INT8X4 holds the whole T~ in device memory as BF16 hi/lo pairs
 * ((nocc nvir_pad)^2 x 4 bytes) for one GEMM Gamma = 4 B T~.
 * kappa must be NaN (None): FH_UNSUPPORTED otherwise.
 * Host-pointer version: b_ov, eps, gamma (naux*nov) and deps (n_eps) on the host. */
fh_status fh_rimp2_energy_adjoint(const double *b_ov, size_t naux, size_t nocc, size_t nvir,
                                  const double *eps, size_t n_eps, size_t first_occ,
                                  size_t nocc_total, double c_os, double c_ss,
                                  fh_precision precision, fh_spin_components *out,
                                  double *gamma, double *deps);

/* Device-resident version: b_dev (b_dtype) and gamma_dev (FP64, naux*nov,
 * overwritten) on the device; eps and deps on the host. */
fh_status fh_rimp2_energy_adjoint_dev(const void *b_dev, fh_dtype b_dtype, size_t naux,
                                      size_t nocc, size_t nvir, const double *eps, size_t n_eps,
                                      size_t first_occ, size_t nocc_total, double c_os,
                                      double c_ss, fh_precision precision, void *stream,
                                      fh_spin_components *out, double *gamma_dev, double *deps);

/* ---- MP2 one-particle densities ----------------------------------------- */

/* The MP2 densities of E = c_os E_os + c_ss E_ss over the active orbitals:
 *   p_oo[i*nocc + j] = P_ij = - sum_k sum_ab t_ik^ab T~_jk^ab
 *   p_vv[a*nvir + b] = P_ab =   sum_ij sum_c t_ij^ac T~_ij^bc
 * with t = (ia|jb) / D and T~ = (c_os + c_ss) t - c_ss t^T: for (1, 1) the
 * reference's build_mp2_density (crates/ferric-mp2/src/oo_rimp2.rs:1192-1300,
 * T~ = 2t - t^T), otherwise its spin-scaled c_ss P_MP2 + (c_os - c_ss) P_OS
 * (apply_spin_scaling / opposite_spin_density, rimp2.rs:1917-2052); host
 * arrays, row-major, symmetric. diag(P) is what fh_rimp2_energy_adjoint
 * returns as deps / 2. out receives the plain (E_os, E_ss). FP64 only. The
 * amplitudes of every ordered occupied pair are held in the workspace, as the
 * reference holds t2: 8 nocc^2 nvir_pad^2 bytes (nvir_pad = nvir rounded up
 * to 16; 0.68 GB at 41/209, 66.5 GB at 100/900), refused with FH_POOL_FULL
 * beyond the pool ceiling. */
fh_status fh_rimp2_mp2_density(const double *b_ov, size_t naux, size_t nocc, size_t nvir,
                               const double *eps, size_t n_eps, size_t first_occ,
                               size_t nocc_total, double c_os, double c_ss,
                               fh_spin_components *out, double *p_oo, double *p_vv);
/* The same with a device-resident b_ov (FP64 or FP32) on the given stream. */
fh_status fh_rimp2_mp2_density_dev(const void *b_dev, fh_dtype b_dtype, size_t naux, size_t nocc,
                                   size_t nvir, const double *eps, size_t n_eps, size_t first_occ,
                                   size_t nocc_total, double c_os, double c_ss, void *stream,
                                   fh_spin_components *out, double *p_oo, double *p_vv);

/* ---- Batched local-MP2 domains ---------------------------------------- */

/* Energies of ndom independent pair domains in one persistent launch
 * (the serial pair/domain loop it replaces: crates/ferric-mp2/src/dlpno_mp2.rs:718-745).
 * domains is a HOST array; b_dev (FP64) and eps_dev (FP64) are the packed
 * device buffers the descriptors point into; e_os/e_ss are HOST arrays of
 * ndom. precision: FH_FP64 or FH_BF16X3. use_graph != 0 captures the launch
 * sequence in a hipGraph on first use for this descriptor set and replays it.
 * A domain must have nocc <= 4096, nvir <= 65536, naux <= 2^20 and
 * nocc nvir < 2^22 (FH_LAYOUT otherwise). */
fh_status fh_rimp2_batched_dev(const fh_domain *domains, size_t ndom, const double *b_dev,
                               size_t b_len, const double *eps_dev, size_t eps_len,
                               fh_precision precision, int use_graph, void *stream,
                               double *e_os, double *e_ss);

/* ---- Pair energies (the local-MP2 pair screen) --------------------------- */

/* estimate_pair_energies_from_b_ov (crates/ferric-mp2/src/dlpno_mp2.rs:527-594):
 * the MP2 energy of every occupied pair from one pass of the fused kernel, as
 * its opposite-spin part e_os[i][j] = fac sum_ab (ia|jb)^2 / D and same-spin
 * part e_ss[i][j] = fac sum_ab (ia|jb)((ia|jb) - (ib|ja)) / D (fac = 1 for
 * i == j, 2 otherwise; the reference's estimate e_ij is their sum). e_os and
 * e_ss are host nocc x nocc row-major, symmetric. A non-negative denominator
 * (2 max eps_occ >= 2 min eps_vir) is FH_LAYOUT, as the reference's refusal;
 * so is a NaN among the active orbital energies.
 * Any precision of fh_rimp2_energy. The ABI version is 5 since these entries
 * were added. */
fh_status fh_rimp2_pair_energies(const double *b_ov, size_t naux, size_t nocc, size_t nvir,
                                 const double *eps, size_t n_eps, size_t first_occ,
                                 size_t nocc_total, fh_precision precision, double *e_os,
                                 double *e_ss);
fh_status fh_rimp2_pair_energies_dev(const void *b_dev, fh_dtype b_dtype, size_t naux,
                                     size_t nocc, size_t nvir, const double *eps, size_t n_eps,
                                     size_t first_occ, size_t nocc_total, fh_precision precision,
                                     void *stream, double *e_os, double *e_ss);

/* PNO truncation (dlpno_mp2.rs:670-760 with t_cut_pno >= 0): for each listed
 * pair, pairs[2k] = i <= pairs[2k+1] = j < nocc (active indices), the pair
 * density D = sym(T T^T + T^T T) of the amplitudes T = g / D_ij
 * (local_pno.rs:132-200), its eigenvectors with |occupation| >= t_cut_pno
 * (never none: the largest is kept), the semicanonical rotation of those PNOs
 * and the pair energy in that basis (dlpno_mp2.rs:249-325): e_os[k], e_ss[k]
 * (host, npairs, the mirror factor applied). diag (may be null) receives the
 * total kept PNO count and the largest discarded weight, from which
 * virtual_retention = kept_pnos / (npairs * nvir) (dlpno_mp2.rs:464-477).
 * FP64 only (b_dev is FP64). t_cut_pno < 0, a malformed pair, or a
 * non-negative denominator in the listed pairs is FH_LAYOUT; an eigensolver
 * failure is FH_KERNEL. At t_cut_pno = 0 every pair energy equals the
 * canonical one of fh_rimp2_pair_energies to rounding. The ABI version is 6
 * since these entries were added. */
typedef struct fh_pno_diagnostics {
    uint64_t kept_pnos;
    double max_discarded_weight;
} fh_pno_diagnostics;

fh_status fh_dlpno_pno_pair_energies(const double *b_ov, size_t naux, size_t nocc, size_t nvir,
                                     const double *eps, size_t n_eps, size_t first_occ,
                                     size_t nocc_total, const int32_t *pairs, size_t npairs,
                                     double t_cut_pno, double *e_os, double *e_ss,
                                     fh_pno_diagnostics *diag);
fh_status fh_dlpno_pno_pair_energies_dev(const double *b_dev, size_t naux, size_t nocc,
                                         size_t nvir, const double *eps, size_t n_eps,
                                         size_t first_occ, size_t nocc_total,
                                         const int32_t *pairs, size_t npairs, double t_cut_pno,
                                         void *stream, double *e_os, double *e_ss,
                                         fh_pno_diagnostics *diag);

/* Domain-local RI fit blocks (domain_fit_pair, lmp2_amplitude.rs:1707-1758):
 * for each listed pair (pairs[2k] = i <= pairs[2k+1] = j < nocc),
 *   J_ij[a][b] = sum_{P,Q in D} A[P][i nvir + a] (V_DD^-1)[P][Q] A[Q][j nvir + b],
 * D = the aux functions P with |aux_xyz[P] - centers[i]| <= radius or
 * |aux_xyz[P] - centers[j]| <= radius (bohr). a: the unwhitened (P|ia),
 * (naux, nocc*nvir) row-major; v: the aux metric (naux, naux); aux_xyz
 * (naux, 3) and centers (nocc, 3) are host arrays. j_out receives npairs
 * row-major nvir x nvir blocks (device for _dev, host otherwise); max_domain
 * (may be null) the largest |D|. An empty domain or a malformed pair is
 * FH_LAYOUT; a V_DD that is not positive definite is FH_KERNEL. FP64. The ABI
 * version is 7 since these entries were added. */
fh_status fh_domain_fit_pair_blocks(const double *a, const double *v, size_t naux, size_t nocc,
                                    size_t nvir, const double *aux_xyz, const double *centers,
                                    double radius, const int32_t *pairs, size_t npairs,
                                    double *j_out, size_t *max_domain);
fh_status fh_domain_fit_pair_blocks_dev(const double *a_dev, const double *v_dev, size_t naux,
                                        size_t nocc, size_t nvir, const double *aux_xyz,
                                        const double *centers, double radius,
                                        const int32_t *pairs, size_t npairs, void *stream,
                                        double *j_dev, size_t *max_domain);

/* ---- Energies from explicit (ia|jb) blocks ------------------------------ */

/* spin_components_from_g (crates/ferric-mp2/src/rimp2.rs:984-1018): the MP2
 * components from a dense (nocc*nvir) x (nocc*nvir) row-major g, for integrals
 * that are not a Gram product (the robust fit's three-term G); pair (i, j)
 * reads (ia|jb) = g[i*nvir + a][j*nvir + b] and (ib|ja) = g[i*nvir + b][j*nvir
 * + a]. One HBM pass over the i <= j blocks, FP64. Like the reference, no
 * denominator is refused. The host entry uploads g (8 (nocc nvir)^2 bytes). */
fh_status fh_rimp2_energy_from_g(const double *g, size_t nocc, size_t nvir, const double *eps,
                                 size_t n_eps, size_t first_occ, size_t nocc_total,
                                 fh_spin_components *out);
fh_status fh_rimp2_energy_from_g_dev(const double *g_dev, size_t nocc, size_t nvir,
                                     const double *eps, size_t n_eps, size_t first_occ,
                                     size_t nocc_total, void *stream, fh_spin_components *out);

/* Per-pair energies from per-pair blocks (the layout fh_domain_fit_pair_blocks
 * writes): block k (nvir x nvir, row-major, device) holds (i_k a|j_k b) at
 * [a][b]; e_os[k], e_ss[k] (host) are the pair's terms of
 * spin_components_from_g, the mirror factor applied. No counterpart entry in
 * the reference: it is the energy of its domain-fit integrals
 * (lmp2_amplitude.rs:1666-1700) without the amplitude iterations. */
fh_status fh_pair_energies_from_blocks_dev(const double *blocks_dev, const int32_t *pairs,
                                           size_t npairs, size_t nocc, size_t nvir,
                                           const double *eps, size_t n_eps, size_t first_occ,
                                           size_t nocc_total, void *stream, double *e_os,
                                           double *e_ss);

/* estimate_pair_energies (crates/ferric-mp2/src/pair_energy_screen.rs:125-190)
 * on a dense device g: the per-pair terms of fh_rimp2_energy_from_g as host
 * nocc x nocc symmetric e_os, e_ss (the mirror factor applied). A
 * non-negative denominator is FH_LAYOUT, as the reference refuses it. */
fh_status fh_pair_energies_from_g_dev(const double *g_dev, size_t nocc, size_t nvir,
                                      const double *eps, size_t n_eps, size_t first_occ,
                                      size_t nocc_total, void *stream, double *e_os,
                                      double *e_ss);

/* The pair loop of dlpno_mp2_spin_components (dlpno_mp2.rs:354-440): PNO
 * truncation as fh_dlpno_pno_pair_energies with each pair's (ia|jb) block read
 * from a dense device g ((nocc*nvir)^2 row-major, not necessarily a Gram
 * product) instead of formed from B. The ABI version is 9 since this entry
 * and fh_pair_energies_from_g_dev were added. */
fh_status fh_dlpno_pno_pair_energies_from_g_dev(const double *g_dev, size_t nocc, size_t nvir,
                                                const double *eps, size_t n_eps,
                                                size_t first_occ, size_t nocc_total,
                                                const int32_t *pairs, size_t npairs,
                                                double t_cut_pno, void *stream, double *e_os,
                                                double *e_ss, fh_pno_diagnostics *diag);

/* ---- Dressing / MO transform upstream of the pair stage ---------------- */

/* stream_dressed_mo_band (crates/ferric-mp2/src/rimp2.rs:648-756) on the
 * device: the RI factor of the MO pairs (l, r) over an output band of aux rows,
 *   B[P, l*nright + r] = sum_Q m[P, Q] sum_{mu,nu} c_left[mu, l] (Q|mu nu) c_right[nu, r],
 * ready for fh_rimp2_energy_dev (c_left: active occupied, c_right: virtual
 * coefficients), so B never crosses PCIe. Device pointers, FP64, row-major:
 * j3 (naux, nao, nao), symmetric in (mu, nu); c_left (nao, nleft); c_right
 * (nao, nright); m (nband, naux), the band's rows of the dressing matrix
 * (V^-1/2 in the reference, or L^-1); b_out (nband, nleft*nright). The two
 * intermediates live in the library's workspace; the three GEMMs run in
 * rocBLAS on the given stream. */
fh_status fh_dress_b_ov_dev(const double *j3, size_t naux, size_t nao, const double *c_left,
                            size_t nleft, const double *c_right, size_t nright, const double *m,
                            size_t nband, void *stream, double *b_out);

/* ---- Unrestricted (open-shell) blocks and the signed factor ------------ */

/* One spin channel of an unrestricted reference, as SpinChannel
 * (crates/ferric-mp2/src/u_rimp2.rs:669-679): b is row-major
 * (naux, nocc*nvir) -- a host const double* for fh_u_same_spin /
 * fh_u_opposite_spin, a device pointer (FP64 or FP32) for the _dev entries;
 * eps is the channel's full orbital-energy vector on the host, occupied i at
 * eps[first_occ + i], virtual a at eps[nocc_total + a]. */
typedef struct fh_spin_channel {
    const void *b;
    size_t nocc;
    size_t nvir;
    const double *eps;
    size_t n_eps;
    size_t first_occ;
    size_t nocc_total;
} fh_spin_channel;

/* The drop-in for try_u_same_spin_on_device (rimp2_gpu.rs:884-889): the
 * same-spin energy 1/4 sum_{i<=j} f_ij sum_ab (g_ab - g_ba)^2 / D of one
 * channel (same_spin_block_energy, u_rimp2.rs:693-729). Any precision. */
fh_status fh_u_same_spin(const fh_spin_channel *ch, size_t naux, fh_precision precision,
                         double *energy);
fh_status fh_u_same_spin_dev(const fh_spin_channel *ch, fh_dtype b_dtype, size_t naux,
                             fh_precision precision, void *stream, double *energy);

/* The drop-in for try_u_opposite_spin_on_device (rimp2_gpu.rs:893-902): the
 * opposite-spin energy sum_{i,J} sum_{a,B} (ia|JB)^2 / D of an alpha and a beta
 * channel fitted in one aux basis (opposite_spin_block_energy,
 * u_rimp2.rs:731-760). precision: FH_FP64, FH_FP32 or FH_INT8X4 (the INT8X4
 * mode, the other precisions are FH_UNSUPPORTED; naux beyond its int32 range
 * is FH_RANGE). */
fh_status fh_u_opposite_spin(const fh_spin_channel *alpha, const fh_spin_channel *beta,
                             size_t naux, fh_precision precision, double *energy);
fh_status fh_u_opposite_spin_dev(const fh_spin_channel *alpha, const fh_spin_channel *beta,
                                 fh_dtype b_dtype, size_t naux, fh_precision precision,
                                 void *stream, double *energy);

/* spin_components_from_signed_b_ov (rimp2.rs:1559-1607) on the device: the
 * closed-shell spin components of K = pos^T pos - neg^T neg, pos
 * (naux_pos, nocc*nvir) and neg (naux_neg, nocc*nvir) row-major (the terf
 * kernel's Coulomb and terfc factors). The reference runs it on the CPU only;
 * no kappa. naux_neg = 0 gives the plain energy of pos. precision: FH_FP64 or
 * FH_FP32. The _dev entry takes device pointers of one dtype. */
fh_status fh_rimp2_energy_signed(const double *pos, size_t naux_pos, const double *neg,
                                 size_t naux_neg, size_t nocc, size_t nvir, const double *eps,
                                 size_t n_eps, size_t first_occ, size_t nocc_total,
                                 fh_precision precision, fh_spin_components *out);
fh_status fh_rimp2_energy_signed_dev(const void *pos, size_t naux_pos, const void *neg,
                                     size_t naux_neg, fh_dtype b_dtype, size_t nocc, size_t nvir,
                                     const double *eps, size_t n_eps, size_t first_occ,
                                     size_t nocc_total, fh_precision precision, void *stream,
                                     fh_spin_components *out);

#ifdef __cplusplus
}
#endif

#endif /* FERRIC_HIP_H */
