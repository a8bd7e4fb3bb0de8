/*
 * zxh_amd.h -- C ABI of the MI355X state-vector engine (libzxh_amd.so).
 *
 * Two layers, both plain C (no C++ or torch types in any signature):
 *
 *  1. A handle-based mirror of ZXH-Sim's Python core module, i.e. the class
 *     the reference's pybind11 module exposes (reference
 *     src/zxh/src/pybind/bindings.cpp:79-118, class declared in
 *     src/zxh/include/zxhsim/zxh.h:11-53).  Each zxh_* call below names the
 *     method it replaces.  A Python ctypes binding, a cgo or a JNI stub binds
 *     these directly.
 *
 *  2. Kernel-level entry points over caller-owned device buffers, mirroring
 *     the reference's internal backend API (src/zxh/include/zxhsim/kernels.h:11-46
 *     and the sampler behind src/zxh/include/zxhsim/measure.h:12-51).
 *
 * Every function returns 0 on success or a ZXH_E* code; zxh_last_error()
 * gives the message of the last failure on the calling thread.  The reference
 * aborts the process on these conditions instead (runtime_general.cpp:79-89).
 *
 * Amplitude layout: interleaved complex, element x at index x (little-endian
 * qubit order of the physical address), complex128 (2 x f64) or complex64
 * (2 x f32).
 */
#ifndef ZXH_AMD_H
#define ZXH_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---- */
#define ZXH_OK 0
#define ZXH_EINVAL 1      /* bad argument (qubit out of range, bad size, ...) */
#define ZXH_ESTATE 2      /* call out of order (e.g. get_state before execute) */
#define ZXH_EZEROPROB 3   /* "state has zero total probability" */
#define ZXH_EUNRESOLVED 4 /* "measure failed to resolve all local samples" */
#define ZXH_ENOMEM 5      /* device or host allocation failed */
#define ZXH_EHIP 6        /* HIP runtime error (launch, copy, sync) */
#define ZXH_ENODEV 7      /* no usable GPU */

/* ---- precision ---- */
#define ZXH_C128 0 /* complex<double>, the reference's only type (types.h:11-12) */
#define ZXH_C64 1  /* complex<float> state, f32 arithmetic, f64 reductions */

/* ---- gate codes: the reference's gate_type_t order (defs.h:65-78) ---- */
#define ZXH_G_BARRIER 1
#define ZXH_G_Z 2
#define ZXH_G_RZ 3
#define ZXH_G_P 4
#define ZXH_G_CRZ 5
#define ZXH_G_CP 6
#define ZXH_G_X 7
#define ZXH_G_CX 8
#define ZXH_G_H 9
#define ZXH_G_U3 10
#define ZXH_G_RX 11 /* recorded as U3(theta, pi/2, -pi/2) (zxh.cpp:303-310) */

/* ---- batching policy ----
 * Both policies apply the same operations the reference's execute() applies
 * (same U3 clusters, same identity-skip decisions, same diagonal windows);
 * they differ only in how many HBM sweeps carry them. */
#define ZXH_POLICY_REFERENCE 0 /* one pass per flushed batch, as the reference (zxh.cpp:437-548) */
#define ZXH_POLICY_FUSED 1     /* default: consecutive batches share a tile pass when they commute */

typedef struct zxh_sim zxh_sim;

const char *zxh_last_error(void);

/* runtime: replaces init/finalize/active/rank/nprocs (bindings.cpp:86-90,
 * runtime_general.cpp:22-89).  Single GPU, rank 0 of 1. */
int zxh_init(void);
int zxh_finalize(void);
int zxh_active(void);
int zxh_rank(void);
int zxh_nprocs(void);

/* ZXH(n, disable_x=False, eager_expand_all=False)  (bindings.cpp:92-94);
 * precision is ZXH_C128 for the reference's behaviour. */
int zxh_create(size_t n, int disable_x, int eager_expand_all, int precision, zxh_sim **out);
int zxh_destroy(zxh_sim *s);

int zxh_set_seed(zxh_sim *s, uint64_t seed); /* ZXH::set_seed  (zxh.cpp:216-220) */
int zxh_clear_seed(zxh_sim *s);              /* ZXH::clear_seed (zxh.cpp:222-226) */
int zxh_clear_gates(zxh_sim *s);             /* ZXH::clear_gates (zxh.cpp:228-232) */

/* gate recording: ZXH::Barrier/Rz/CRz/CP/P/Z/X/CX/H/U3/Rx (zxh.cpp:234-310).
 * U3 argument order is (theta, lambda, phi), as in the reference. */
int zxh_barrier(zxh_sim *s);
int zxh_rz(zxh_sim *s, size_t q, double theta);
int zxh_crz(zxh_sim *s, size_t cq, size_t q, double theta);
int zxh_cp(zxh_sim *s, size_t cq, size_t q, double theta);
int zxh_p(zxh_sim *s, size_t q, double theta);
int zxh_z(zxh_sim *s, size_t q);
int zxh_x(zxh_sim *s, size_t q);
int zxh_cx(zxh_sim *s, size_t cq, size_t q);
int zxh_h(zxh_sim *s, size_t q);
int zxh_u3(zxh_sim *s, size_t q, double theta, double lambda, double phi);
int zxh_rx(zxh_sim *s, size_t q, double theta);
/* one call per gate code (ZXH_G_*), for bulk loaders */
int zxh_gate(zxh_sim *s, int code, size_t cq, size_t q, double theta, double lambda, double phi);

/* ZXH::execute (zxh.cpp:312-553); synchronous.  Starts from the reset state
 * |0...0>, as the reference does; under the fused policy the first tile pass
 * forms that state in registers instead of a separate zero fill (no family-6
 * launch then; the pass's algorithmic bytes are one write of the state). */
int zxh_execute(zxh_sim *s);

int zxh_measure(zxh_sim *s, size_t shots);                        /* ZXH::measure (zxh.cpp:555-559) */
int zxh_measured_count(const zxh_sim *s, size_t *out);            /* ZXH::measured_count */
/* ZXH::get_results: ceil(N/64) little-endian words per shot, bit q = qubit q
 * (measure_result_pack.h:24-52) */
int zxh_get_results(const zxh_sim *s, uint64_t *packed, size_t shots);
/* physical (real) addresses of the last measure, ascending */
int zxh_get_real_results(const zxh_sim *s, uint64_t *real, size_t shots);
/* ZXH::get_state (zxh.cpp:579-636): 2^N complex128 values (interleaved
 * doubles, out_len = 2^(N+1) doubles) in bit-reversed ("cudaq") order,
 * multiplied by the global phase.  Works for both precisions. */
int zxh_get_state(const zxh_sim *s, double *out, size_t out_len);
int zxh_required_M(const zxh_sim *s, size_t *out); /* ZXH::required_M (zxh.cpp:638-641) */
int zxh_num_qubits(const zxh_sim *s, size_t *out); /* ZXH::num_qubits */

/* ---- extensions over the reference's class (no reference counterpart) ---- */
int zxh_set_policy(zxh_sim *s, int policy); /* ZXH_POLICY_*; default ZXH_POLICY_FUSED */
/* sampling segment bits I (segments of 2^min(M, I) amplitudes,
 * sv_common.cpp:87-100).  zxh_create reads the default from the environment
 * variable ZXHSIM_SEG_MIN_I with the reference's parse rules
 * (parse_env_size_t, sv_common.cpp:49-66: decimal, the whole string; unset,
 * empty, trailing text or 0 give 18); this call overrides it for one
 * simulator (0 restores 18). */
int zxh_set_seg_bits(zxh_sim *s, size_t bits);
/* raw physical state: 2^m amplitudes of the simulator's precision (host copy) */
int zxh_get_phys_state(const zxh_sim *s, void *out, size_t out_bytes, size_t *m_out);
/* replace the physical state after execute() (same width m, host data of the
 * simulator's precision); used to sample a caller-given probability vector */
int zxh_set_phys_state(zxh_sim *s, const void *in, size_t in_bytes);
/* device pointer of the physical state and its width m (valid until the next
 * execute()) */
int zxh_state_device_ptr(const zxh_sim *s, void **dptr, size_t *m_out);
/* per-kernel-family timing with HIP events (0 = off) */
int zxh_set_timing(zxh_sim *s, int on);
/* family: 0 diag, 1 u3 tile pass, 2 u3 pair (general), 3 perm, 4 cp,
 * 5 sample, 6 reset, 7 copy, 8 scale, 9 fused tile pass (u3 + diagonal
 * window), 10 observable, 11 adjoint, 12 dense block.  Fills launches, total
 * milliseconds and algorithmic bytes.  Family 13 only counts the launches of
 * families 1 and 9 that ran on the LDS-free upper kernel (time and bytes 0;
 * they are in those families).  Family 14 is the sampler's probability tree
 * (all passes of one tree build as one entry; bytes = one read of the state
 * plus the stored tree levels); family 5 is the sampler's search. */
int zxh_get_timing(zxh_sim *s, int family, uint64_t *launches, double *total_ms, uint64_t *bytes);
int zxh_reset_timing(zxh_sim *s);

/* ---- kernel level: caller-owned device buffers (kernels.h:11-46) ---- */

/* u3_batch_desc_t (defs.h:85-98): selector packed as alpha | beta<<63 */
typedef struct {
    uint64_t packed_sel;
    uint64_t delta;
    double u[8]; /* u00 re,im  u01 re,im  u10 re,im  u11 re,im */
} zxh_u3_desc;

/* zeroed device allocation (worker_alloc, mem.h:9-19) */
int zxh_dev_alloc(size_t bytes, void **dptr);
int zxh_dev_free(void *dptr);
int zxh_dev_upload(void *dptr, const void *host, size_t bytes);
int zxh_dev_download(void *host, const void *dptr, size_t bytes);
/* worker_realloc (mem.h:9-19, mem_host.cpp:152-169): a zeroed allocation of
 * new_bytes holding the first min(old, new) bytes of old_ptr, which is freed;
 * new_bytes == 0 frees old_ptr and returns NULL */
int zxh_dev_realloc(void *old_ptr, size_t old_bytes, size_t new_bytes, void **new_ptr);
/* worker_set_zero: amplitudes [begin, end) of a state of the given precision */
int zxh_dev_set_zero(void *dptr, int precision, size_t begin, size_t end);
/* worker_mem_set: amplitude 0 = re + i im (the reset's |0...0>) */
int zxh_dev_mem_set(void *dptr, int precision, double re, double im);
int zxh_dev_sync(void);

/* u3_block_batch_kernel (kernels.h:26): applies the descriptors (distinct
 * target qubits, commuting) to 2^local_bits amplitudes in place. */
int zxh_k_u3_batch(void *state, int precision, size_t local_bits, const zxh_u3_desc *descs, size_t ndesc);
/* one fused tile pass: the U3 descriptors (single-bit pair offsets with
 * alpha == delta, at most 4 targets below bit 4 and 8 above, distinct), then
 * the diagonal window, in one sweep.  The window is a phase polynomial
 * (selectors with at most one alpha bit, or one two-bit selector) or any
 * selectors whose alpha parts span at most 10 dimensions over GF(2) (applied
 * as a table over those forms); other windows are rejected (no reference
 * counterpart: the reference runs these as separate sweeps) */
int zxh_k_fused_pass(void *state, int precision, size_t local_bits, const zxh_u3_desc *descs, size_t ndesc,
                     const uint64_t *sel0, const uint64_t *sel1, const double *theta, size_t ndiag);
/* The table a fused tile pass applies for a diagonal window beyond the
 * phase-polynomial form (host only; needs no device and no init).  The
 * window's selector forms alpha (sel & ~(1 << 63)) are reduced to a basis
 * a_0 .. a_{r-1} over GF(2), r <= 10, written to basis[0 .. 9] (zero past r)
 * and *rank; with y(x) = sum_k parity(a_k & x) << k the window's phase at
 * amplitude address x is table[2 y] + i table[2 y + 1] (2^r entries; the
 * caller provides 2048 doubles).  ZXH_EINVAL when the forms span more than
 * 10 dimensions or reach above nbits.  No reference counterpart. */
int zxh_window_table(size_t nbits, const uint64_t *sel0, const uint64_t *sel1, const double *theta, size_t ngates,
                     uint64_t *basis, int *rank, double *table);
/* How a batch of single-qubit gates on distinct amplitude address bits
 * (targets[0 .. ntargets-1]) is split into tile passes when it needs more
 * than one and the wide tile is on (knob tile.wide.<precision>), under tile
 * geometry 3..6 (anything else: the knob tile.seg.<precision>): pass_of[i]
 * is the pass of targets[i], pass 0 a 4096-unit tile pass, the others wide
 * passes (every pass wide when no target lies below the wide tile's 16-unit
 * run and that takes fewer passes); *npasses = 0 when the batch fits one
 * 4096-unit tile pass or no such split applies (pass_of untouched).  The engine and zxh_k_u3_batch split batches
 * this way.  Host only; needs no device.  No reference counterpart. */
int zxh_tile_layer_plan(int precision, size_t nbits, int geometry, const uint32_t *targets, size_t ntargets,
                        uint32_t *pass_of, size_t *npasses);
/* diag_pending_push_p / try_push_cp / flush (kernels.h:36-41) as one call:
 * ngates entries, sel1[i] == ~0 marks a P gate (Z/Rz already lowered to P). */
int zxh_k_diag(void *state, int precision, size_t local_bits, const uint64_t *sel0, const uint64_t *sel1,
               const double *theta, size_t ngates);
/* zxh_k_diag that allocates no device memory.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * per-chunk coefficients of the sweep go to a 32 MiB workspace that the
 * library holds per GPU until the process ends (a device array, looked up by
 * zxh_init on the current device); zxh_init must have been called.  A degree-2
 * polynomial window (every plain-method P/CP/Rz/Z window) is applied to states
 * of up to 30 qubits (complex128) or 31 (complex64): where the autotuned flat
 * sweep's coefficients (128 bytes per 2^9 / 2^10 amplitudes) do not fit, the
 * polynomial sweep (128 bytes per 2^12 / 2^13) runs instead, whose results
 * agree within the same rounding bound; wider states return ZXH_EINVAL.
 * Other windows run the per-gate sweep, which needs no workspace.  Every
 * call on a GPU shares that GPU's workspace (rewritten in full before it is
 * read, so nothing needs clearing): calls that overlap on one GPU race. */
int zxh_k_diag_noalloc(void *state, int precision, size_t local_bits, const uint64_t *sel0, const uint64_t *sel1,
                       const double *theta, size_t ngates);
/* x_block_kernel / cx_block_kernel / cp_kernel (kernels.h:18-19, 47) */
int zxh_k_x(void *state, int precision, size_t local_bits, uint64_t delta, uint64_t packed_sel);
int zxh_k_cx(void *state, int precision, size_t local_bits, uint64_t delta, uint64_t packed_sel_cq,
             uint64_t packed_sel_q);
int zxh_k_cp(void *state, int precision, size_t local_bits, uint64_t packed_sel_cq, uint64_t packed_sel_q,
             double theta);
/* multi-controlled X on physical bits: swaps amplitude pairs differing in
 * bit target where every bit in ctrl_mask is set (no reference counterpart:
 * the reference has no MCX; exact index work). */
int zxh_k_mcx(void *state, int precision, size_t local_bits, uint64_t ctrl_mask, size_t target);
/* streaming multinomial sampler over 2^m device amplitudes (measure_general.cpp,
 * measure_streaming_impl.h): sorted thresholds generated from stream_seed as
 * the reference does, segments of 2^min(m, seg_bits); writes shots physical
 * addresses (ascending) to host memory. */
int zxh_k_sample(const void *state, int precision, size_t m, size_t seg_bits, uint64_t stream_seed, size_t shots,
                 uint64_t *out_real);
/* sampler on a caller-provided sorted threshold array (host, length shots) */
/* measure(sv, A, b, cache, cnt, use_seed, seed) + get_results (measure.h:
 * 12-51, measure_common.cpp:73-92) on a state of 2^m amplitudes: the
 * reference's threshold stream (mix_seed(use_seed ? seed : random_device,
 * 0)), segments of 2^min(m, seg_bits), and the affine map of nq virtual qubits
 * (rows[q] bit c = A[q][c], b[q]): packed gets ceil(nq / 64) words per shot
 * (bit q = parity(real & rows[q]) ^ b[q], measure_result_pack.h:24-44),
 * out_real (nullable) the real addresses, ascending. */
int zxh_k_measure(const void *state, int precision, size_t m, size_t seg_bits, size_t nq, const uint64_t *rows,
                  const uint8_t *b, int use_seed, uint64_t seed, size_t shots, uint64_t *packed, uint64_t *out_real);
int zxh_k_sample_thresholds(const void *state, int precision, size_t m, size_t seg_bits, const double *thresholds,
                            size_t shots, uint64_t *out_real);
/* sum of |amplitude|^2 in the reference's pairwise order (f64) */
int zxh_k_prob_total(const void *state, int precision, size_t m, double *out);
/* out (host, count amplitudes of the state's precision) = the amplitudes at
 * physical addresses indices[0..count): a partial get_state (zxh.cpp:579-636
 * reads all 2^m) for probes of states too large to copy whole */
int zxh_k_gather(const void *state, int precision, size_t local_bits, const uint64_t *indices, size_t count,
                 void *out);
/* device copy of bytes (a multiple of 16); the bandwidth and HBM-counter
 * calibration reference (no reference counterpart) */
int zxh_k_copy(void *dst, const void *src, size_t bytes);
/* state *= (re + i im) in place: renormalisation or a global phase (no
 * reference counterpart) */
int zxh_k_scale(void *state, int precision, size_t local_bits, double re, double im);
/* Expectation of a diagonal observable H = sum_k coef[k] Z_qi[k] Z_qj[k]
 * (qj[k] < 0 for a single Z) on a plain-addressed state (qubit q = bit q):
 * out[0] = <psi|H|psi>, out[1] = <psi|psi>, accumulated in FP64 with Kahan
 * compensation and a fixed reduction order (deterministic).  No reference
 * counterpart: the reference only samples. */
int zxh_k_expect_z(const void *state, int precision, size_t local_bits, const int32_t *qi, const int32_t *qj,
                   const double *coef, size_t nterms, double *out);
/* dst = H src for the same observable (dst may not alias src); out (nullable)
 * receives the same two sums as zxh_k_expect_z.  The adjoint method's first
 * step.  No reference counterpart. */
int zxh_k_apply_z(const void *src, void *dst, int precision, size_t local_bits, const int32_t *qi,
                  const int32_t *qj, const double *coef, size_t nterms, double *out);
/* Two states of 2^local_bits amplitudes (each complex128 or complex64),
 * read once each: out5 = { Re <a|b>, Im <a|b>, |a|^2, |b|^2 (FP64 with
 * compensated, fixed-order sums), max_i |a_i - b_i| }.  Gives a complex64
 * result's infidelity 1 - |<a|b>| and maximum amplitude error against the
 * complex128 run of the same circuit.  No reference counterpart. */
int zxh_k_compare(const void *a, int precision_a, const void *b, int precision_b, size_t local_bits, double *out5);
/* |0...0> in place: worker_set_zero + worker_mem_set (mem.h:9-19) */
int zxh_k_reset(void *state, int precision, size_t local_bits);
/* Dense 5-qubit unitary on physical bits targets[0..5) (bit b of the
 * combination index is targets[b]): u is 32x32 complex, row-major (re, im)
 * pairs.  One sweep on the matrix cores (FP32 accumulation for complex64,
 * FP64 for complex128).  No reference counterpart. */
int zxh_k_block5(void *state, int precision, size_t local_bits, const uint32_t *targets, const double *u);

/* ---- parameterized circuits: expectation values and adjoint gradients ----
 * No reference counterpart (the reference has no gradient path).  Operations
 * act on qubit q = address bit q; param >= 0 selects params[param] as the
 * angle, param < 0 uses value.  RX/RY/RZ(theta) = exp(-i theta G / 2) with
 * G = X/Y/Z; CZ and CP(theta) are diagonal. */
enum
{
    ZXH_A_RX = 0,
    ZXH_A_RY = 1,
    ZXH_A_RZ = 2,
    ZXH_A_H = 3,
    ZXH_A_CZ = 4,
    ZXH_A_CP = 5
};
typedef struct
{
    int32_t kind;
    int32_t q0;
    int32_t q1;
    int32_t param;
    double value;
} zxh_ansatz_op;
typedef struct zxh_ansatz zxh_ansatz;
int zxh_ansatz_create(size_t n, int precision, const zxh_ansatz_op *ops, size_t nops, size_t nparams,
                      zxh_ansatz **out);
int zxh_ansatz_destroy(zxh_ansatz *a);
/* observable: sum_k coef[k] Z_qi[k] Z_qj[k] (qj[k] < 0: one Z) */
int zxh_ansatz_set_observable(zxh_ansatz *a, const int32_t *qi, const int32_t *qj, const double *coef,
                              size_t nterms);
/* out[0] = <psi(params)|H|psi(params)>, out[1] = <psi|psi> (FP64 sums); the
 * final state stays on the device (zxh_ansatz_state) */
int zxh_ansatz_forward(zxh_ansatz *a, const double *params, double *out);
/* grads[nparams] = dE/dparams by the adjoint method (no stored intermediate
 * states); re-runs the forward pass unless the last forward used the same
 * params; out (nullable) as in forward.  Uncomputes the device state. */
int zxh_ansatz_backward(zxh_ansatz *a, const double *params, double *grads, double *out);
/* device address of the state and its width; tile passes per forward */
int zxh_ansatz_state(zxh_ansatz *a, void **state, size_t *nbits);
int zxh_ansatz_passes(zxh_ansatz *a, size_t *tile_passes);
/* HIP-event timing of the ansatz's own stream (families as zxh_get_timing) */
int zxh_ansatz_set_timing(zxh_ansatz *a, int on);
int zxh_ansatz_get_timing(zxh_ansatz *a, int family, uint64_t *launches, double *total_ms, uint64_t *bytes);
int zxh_ansatz_reset_timing(zxh_ansatz *a);

/* ---- batches of parameterized circuits (different widths and depths) ----
 * One arena for every set's states; each forward pass index, the observables
 * and each backward step run as one persistent launch over all sets, the
 * whole sequence replayed from a hipGraph.  No reference counterpart. */
typedef struct zxh_batch zxh_batch;
int zxh_batch_create(int precision, zxh_batch **out);
int zxh_batch_destroy(zxh_batch *b);
int zxh_batch_add(zxh_batch *b, size_t n, const zxh_ansatz_op *ops, size_t nops, size_t nparams,
                  const int32_t *qi, const int32_t *qj, const double *coef, size_t nterms);
/* allocates the arena, prepares the launches, captures the graphs */
/* Tile geometry (3..6 contiguous unit bits; 0 = the fewest-passes rule) of
 * the sets of n qubits added after this call: the batch autotuner times each
 * width class under every geometry and sets its choice here. */
int zxh_batch_set_geometry(zxh_batch *b, size_t n, int seg);
int zxh_batch_finalize(zxh_batch *b);
/* params: all sets' parameters concatenated in set order; energies[2 * sets]
 * (<H>, <psi|psi>); grads (nullable: forward only) concatenated like params */
int zxh_batch_run(zxh_batch *b, const double *params, double *energies, double *grads);
int zxh_batch_info(zxh_batch *b, size_t *sets, size_t *total_params, size_t *launches);
/* one set's width and number of tile passes per forward (after finalize) */
int zxh_batch_set_info(zxh_batch *b, size_t set, unsigned *nbits, size_t *passes);

/* tuning knobs (process-wide), set by the autotuner (zxh_adae_amd.autotune):
 *   "diag.ept.c128", "diag.ept.c64": 0 = persistent diagonal sweep,
 *                                    1, 2, 4 = flat grid, 16-byte units per thread
 *   "diag.nt.c128", "diag.nt.c64": nontemporal loads/stores in the flat
 *                                  sweep (0/1)
 *   "tile.seg.c128", "tile.seg.c64": 3, 4, 5, 6 contiguous 16-byte-unit address
 *                                    bits per tile pass (targets per pass: 12
 *                                    minus this above them)
 *   "tile.bufs.c128", "tile.bufs.c64": 2 or 3 register buffers per tile-pass
 *                                      workgroup (tiles in flight)
 *   "tile.upper.c128", "tile.upper.c64": a pass whose targets all sit on tile
 *       unit bits 6..11 runs on the LDS-free upper kernel when its highest
 *       target unit-address bit is at least this (0..64, 64 = never;
 *       default 18); results are bit-identical either way
 *   "block.variant.c128", "block.variant.c64": dense-block kernel variant,
 *       bit 0 prefetch the next wave tile, bit 1 matrix re-read from LDS,
 *       bit 2 cached (not nontemporal) loads and stores (0..7)
 *   "block.tiles.c128", "block.tiles.c64": wave tiles per wave of a block
 *       sweep (1..4096), or 0 for a persistent grid (default 0) */
int zxh_tune_set(const char *name, int value);
int zxh_tune_get(const char *name, int *value);
/* HIP-event timing of the kernel-level calls (same families as zxh_get_timing,
 * plus 7 copy, 8 scale, 9 fused pass, 10 observable, 11 adjoint pass,
 * 12 dense block) */
int zxh_k_set_timing(int on);
int zxh_k_get_timing(int family, uint64_t *launches, double *total_ms, uint64_t *bytes);
int zxh_k_reset_timing(void);

#ifdef __cplusplus
}
#endif

#endif
