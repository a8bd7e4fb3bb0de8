/*
 * zxh_oracle.h -- CPU restatement of ZXH-Sim's CPU path (test infrastructure).
 *
 * THIS IS THE ORACLE, NOT THE PRODUCT.  Only tests, the smoke check and the
 * CPU-baseline leg of the benchmark may load it, and only as a checker or as
 * the timed CPU baseline.  The product path (zxh_adae_amd) never links it.
 *
 * It restates, in plain C, the semantics of (reference paths, relative to the
 * reference checkout):
 *   src/zxh/src/core/zxh.cpp            host scheduler (execute, U3 fusion,
 *                                        diag window, affine map, get_state)
 *   src/zxh/src/core/bitmat.cpp         GF(2) solve / mul / row_xor
 *   src/zxh/src/core/selector.cpp       selector parity evaluation
 *   src/zxh/src/sv/sv_common.cpp        choose_layout / reset / expand
 *   src/zxh/src/kernel/kernel_general.cpp, kernel_omp.cpp   CPU kernels
 *   src/zxh/src/measure/measure_common.cpp, measure_general.cpp,
 *   src/zxh/src/measure/measure_streaming_impl.h, measure_result_pack.h
 *                                        streaming multinomial sampler
 *
 * Pinning (see tests/test_oracle_*.py): golden fixtures produced by the
 * compiled reference (tests/golden/), the standard's mt19937_64 known answer,
 * closed forms (uniform superposition, product states), and brute-force dense
 * matrix-vector products on small qubit counts.
 */
#ifndef ZXH_ORACLE_H
#define ZXH_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Gate codes follow the reference's gate_type_t order (defs.h:65-78). */
enum {
    ZO_NONE = 0,
    ZO_BARRIER = 1,
    ZO_Z = 2,
    ZO_RZ = 3,
    ZO_P = 4,
    ZO_CRZ = 5,
    ZO_CP = 6,
    ZO_X = 7,
    ZO_CX = 8,
    ZO_H = 9,
    ZO_U3 = 10,
    ZO_RX = 11 /* recorded as U3(theta, pi/2, -pi/2), zxh.cpp:303-310 */
};

/* Error codes (the reference aborts the process instead). */
enum {
    ZO_OK = 0,
    ZO_ERR_RANGE = 1,     /* qubit out of range */
    ZO_ERR_STATE = 2,     /* execute() not called / no state */
    ZO_ERR_ZERO_PROB = 3, /* "state has zero total probability" */
    ZO_ERR_UNRESOLVED = 4,/* "measure failed to resolve all local samples" */
    ZO_ERR_ALLOC = 5,
    ZO_ERR_ARG = 6
};

typedef struct zo_sim zo_sim;

/* Mirror of the ZXH class (zxh.h:11-53). */
zo_sim *zo_create(size_t n, int disable_x, int eager_expand_all);
void zo_destroy(zo_sim *s);
void zo_set_seed(zo_sim *s, uint64_t seed);
void zo_clear_seed(zo_sim *s);
void zo_clear_gates(zo_sim *s);
/* Records one gate with the reference's recording rules (CRz lowering,
 * Rx mapping, range checks). */
int zo_add_gate(zo_sim *s, int type, size_t cq, size_t q, double theta, double lambda, double phi);
int zo_execute(zo_sim *s);
int zo_measure(zo_sim *s, size_t cnt);
size_t zo_measured_count(const zo_sim *s);
/* packed virtual results, ceil(N/64) words per shot (measure_result_pack.h) */
int zo_get_results(const zo_sim *s, uint64_t *packed, size_t cnt);
/* real (physical) sampled addresses, ascending */
int zo_get_real_results(const zo_sim *s, uint64_t *real, size_t cnt);
/* full 2^N state, bit-reversed ("cudaq") order, times the global phase */
int zo_get_state(const zo_sim *s, double *out_interleaved);
/* raw physical state: 2^m complex values; *m_out receives m */
int zo_get_phys_state(const zo_sim *s, double *out_interleaved, size_t cap_elems, size_t *m_out);
size_t zo_required_M(const zo_sim *s);
size_t zo_num_qubits(const zo_sim *s);
void zo_global_phase(const zo_sim *s, double *re, double *im);
/* ZXHSIM_SEG_MIN_I analogue (sv_common.cpp:62-66, 87-100); default 18 */
void zo_set_seg_min_i(zo_sim *s, size_t min_i);
/* affine map export: rows[q] (bit c = A[q][c]), b[q] */
int zo_get_affine(const zo_sim *s, uint64_t *rows, uint8_t *b, size_t *m_out);

/* ---- standalone pieces (per-kernel and sampler pins) ---- */
uint64_t zo_mix_seed(uint64_t base, uint64_t salt);
/* n-th (1-based) output of mt19937_64 seeded with seed */
uint64_t zo_mt64_nth(uint64_t seed, size_t n);
/* sorted threshold stream (measure_common.cpp:14-71) */
int zo_thresholds(uint64_t stream_seed, size_t cnt, double begin, double end, double *out);
/* streaming multinomial sampler over a physical state of 2^m amplitudes with
 * segments of 2^seg_bits (I = min(m, seg_bits)); stream_seed is already mixed.
 * Writes cnt real addresses (ascending). */
int zo_sample_state(const double *psi, size_t m, size_t seg_bits, uint64_t stream_seed, size_t cnt,
                    uint64_t *out_real);
/* the same sampler on caller-given sorted thresholds */
int zo_sample_thresholds(const double *psi, size_t m, size_t seg_bits, const double *thr, size_t cnt, uint64_t *out);
double zo_prob_total(const double *psi, size_t m);

/* kernels over a raw block of 2^local_bits amplitudes (kernel_general.cpp) */
typedef struct {
    uint64_t packed_sel; /* alpha | beta<<63 */
    uint64_t delta;
    double u[8];         /* u00 re,im, u01, u10, u11 */
} zo_u3desc;
void zo_k_u3_batch(double *psi, size_t local_bits, const zo_u3desc *descs, size_t ndesc);
void zo_k_x(double *psi, size_t local_bits, uint64_t delta, uint64_t packed_sel);
void zo_k_cx(double *psi, size_t local_bits, uint64_t delta, uint64_t packed_sel_cq, uint64_t packed_sel_q);
void zo_k_cp(double *psi, size_t local_bits, uint64_t packed_sel_cq, uint64_t packed_sel_q, double theta);
/* diag window: ngates entries; sel1[i] == ZO_NO_SEL marks a P gate.  Gates are
 * bucketed exactly as diag_pending_push_p / try_push_cp; returns the number of
 * CP gates rejected (cross selectors), which are applied with zo_k_cp after
 * the window, as the reference's execute() does after a flush. */
#define ZO_NO_SEL (~(uint64_t)0)
int zo_k_diag(double *psi, size_t local_bits, const uint64_t *sel0, const uint64_t *sel1, const double *theta,
              size_t ngates);
/* U3 matrix of the reference (zxh.cpp:101-113) */
void zo_u3_matrix(double theta, double lambda, double phi, double out[8]);

#ifdef __cplusplus
}
#endif

#endif
