// dev.h -- shared host/device definitions of the MI355X engine.
//
// One Ctx per simulator (plus one process-wide Ctx for the kernel-level C
// entry points) owns a HIP stream of default priority, grow-only device
// scratch buffers and a ring of parameter slots for per-launch tables.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace zxa
{

enum Prec
{
    C128 = 0,
    C64 = 1
};

inline size_t elem_bytes(int prec)
{
    return prec == C128 ? 16 : 8;
}

// Status codes (mirror include/zxh_amd.h).
enum Status
{
    OK = 0,
    EINVAL_ = 1,
    ESTATE = 2,
    EZEROPROB = 3,
    EUNRESOLVED = 4,
    ENOMEM_ = 5,
    EHIP = 6,
    ENODEV_ = 7
};

// Thread-local error message, read through zxh_last_error().
void set_error(const std::string &msg);
const char *last_error();

// Returns EHIP with the message recorded when e != hipSuccess.
int hip_check(hipError_t e, const char *what);
#define ZXA_HIP(call)                                                                                                  \
    do                                                                                                                 \
    {                                                                                                                  \
        int rc_ = ::zxa::hip_check((call), #call);                                                                     \
        if (rc_)                                                                                                       \
            return rc_;                                                                                                \
    } while (0)
#define ZXA_TRY(call)                                                                                                  \
    do                                                                                                                 \
    {                                                                                                                  \
        int rc_ = (call);                                                                                              \
        if (rc_)                                                                                                       \
            return rc_;                                                                                                \
    } while (0)

// Grow-only device buffer.
struct DevBuf
{
    void *ptr = nullptr;
    size_t bytes = 0;
    int ensure(size_t need);
    void release();
    ~DevBuf()
    {
        release();
    }
};

// Grow-only pinned host buffer (hipHostMalloc).
struct PinBuf
{
    void *ptr = nullptr;
    size_t bytes = 0;
    int ensure(size_t need);
    void release();
    PinBuf() = default;
    PinBuf(const PinBuf &) = delete;
    PinBuf &operator=(const PinBuf &) = delete;
    ~PinBuf()
    {
        release();
    }
};

// Kernel families for event timing.
enum Family
{
    F_DIAG = 0,
    F_U3TILE = 1,
    F_U3PAIR = 2,
    F_PERM = 3,
    F_CP = 4,
    F_SAMPLE = 5,
    F_RESET = 6,
    F_COPY = 7,
    F_SCALE = 8,
    F_FUSED = 9,
    F_OBS = 10,
    F_ADJ = 11,
    F_BLOCK = 12,
    F_UPPER = 13, // count only: the F_U3TILE / F_FUSED launches that ran upper_kernel
    F_TREE = 14,  // the sampler's probability tree (all its passes as one entry)
    F_COUNT = 15
};

struct FamilyTiming
{
    uint64_t launches = 0;
    double ms = 0.0;
    uint64_t bytes = 0; // algorithmic HBM bytes of the timed launches
};

// Ring of parameter slots: mapped host staging + device copy + event.  A slot
// is reused only after the event recorded behind its last consumer fired.
class ParamRing
{
  public:
    static constexpr int kSlots = 16;
    static constexpr size_t kSlotBytes = 256 * 1024;
    int init();
    void release();
    // Copies bytes from src into the next slot (ring_copy on stream) and
    // returns the device address.  Call done() after the consuming launches.
    int stage(const void *src, size_t bytes, hipStream_t stream, void **dev);
    int done(hipStream_t stream);
    // Batched staging: between begin_batch() and flush(), stage() packs the
    // blocks into one slot at 256-byte aligned offsets and issues no copy (a
    // block that no longer fits copies the packed ones and opens the next
    // slot); flush() copies the packed bytes with one ring_copy and ends the
    // batch.  Launches that read a block come after flush().  end_batch()
    // drops an unflushed batch (error returns).
    int begin_batch();
    int flush(hipStream_t stream);
    void end_batch();

  private:
    int open_slot(int *slot);
    void *host_ = nullptr;
    void *hdev_ = nullptr; // host_ as the device addresses it (mapped, coherent)
    void *dev_ = nullptr;
    hipEvent_t ev_[kSlots] = {};
    bool used_[kSlots] = {};
    int next_ = 0;
    std::vector<int> open_; // staged since the last done()
    int batch_ = -1;        // slot being packed, or -1
    size_t boff_ = 0;       // bytes packed in it
};

// Copies a staged block from the ring's mapped host slot (src: its device
// address) to its device slot with a copy kernel on stream; bytes is rounded
// up to 16, which the slots' 256-byte offsets and size keep inside the slot.
int ring_copy(hipStream_t stream, void *dst, const void *src, size_t bytes);

struct Ctx
{
    int device = 0;
    int num_cus = 0;
    hipStream_t stream = nullptr;
    ParamRing ring;
    DevBuf scratch_a; // diag per-chunk coefficients
    DevBuf scratch_b; // sampler tree levels
    DevBuf scratch_c; // sampler thresholds / results
    DevBuf scratch_d; // sampler per-segment tails
    DevBuf scratch_e; // observable per-workgroup partial sums
    DevBuf scratch_f; // adjoint per-wave transition accumulators
    DevBuf scratch_g; // tile-pass per-tile fan factors
    bool timing = false;
    FamilyTiming fam[F_COUNT];
    // sampler state between prepare and search (tree in scratch_b, tails in scratch_d)
    alignas(8) unsigned char sample_levels[64 * 8] = {};
    unsigned sample_m = 0, sample_I = 0;
    bool sample_ready = false;
    PinBuf sample_total_pin; // the tree root, copied back asynchronously
    PinBuf sample_thr_pin;   // thresholds of the kernel-level measure
    template <typename T>
    T &sample_levels_as()
    {
        static_assert(sizeof(T) <= sizeof(sample_levels), "levels table too large");
        return *reinterpret_cast<T *>(sample_levels);
    }
    int init();
    void release();
    ~Ctx()
    {
        release();
    }
    // Timing brackets around one launch (no-ops when timing is off).  Events
    // are recorded on the launch stream without synchronising; collect()
    // waits for them and accumulates per-family durations.
    int tic();
    int toc(int family, uint64_t bytes);
    int collect();

  private:
    struct Pending
    {
        int family;
        uint64_t bytes;
        hipEvent_t a, b;
    };
    std::vector<Pending> pending_;
    std::vector<hipEvent_t> pool_;
    hipEvent_t cur_ = nullptr;
    int take(hipEvent_t *e);
};

// Tuning knobs (process-wide): variant choices the autotuner sweeps.
//   diag.ept.c128 / diag.ept.c64  0 = persistent sweep, 1/2/4 = flat grid with
//                                 that many 16-byte units per thread
//   diag.nt.c128 / diag.nt.c64     nontemporal loads/stores in the flat sweep
//   tile.seg.c128 / tile.seg.c64   contiguous unit-address bits of a tile (3..6)
//   tile.bufs.c128 / tile.bufs.c64 register buffers per tile-pass workgroup (2, 3)
//   tile.wide.c128 / tile.wide.c64 1: passes of up to nine targets above a
//                                  16-unit run may use the 8192-unit tile
//   tile.upper.c128 / .c64         a pass with every target on unit bits 6..11
//                                  whose highest target unit-address bit is at
//                                  least this runs on the LDS-free upper kernel
//                                  (64: never)
//   block.variant.c128 / .c64      dense-block kernel variant (bit 0 prefetch,
//                                  bit 1 matrix from LDS, bit 2 cached access)
//   block.tiles.c128 / .c64        wave tiles per wave of a block sweep (0 =
//                                  persistent grid)
int tune_get(const char *name);
int tune_set(const char *name, int value);

// Process-wide context for the kernel-level entry points.
Ctx *default_ctx(int *rc);

// ---- launchers (implemented in the .hip files) ----

struct DiagGate
{
    uint64_t s0;
    uint64_t s1; // ~0 for a P gate
    double theta;
};
constexpr uint64_t kNoSel = ~uint64_t(0);

// One diagonal sweep applying every gate of the window.
int diag_apply(Ctx &ctx, void *state, int prec, unsigned nbits, const DiagGate *gates, size_t ngates);
// diag_apply with a caller-given coefficient workspace (nullable; see
// kernels_diag.hip): no device allocation when ws is given.
int diag_apply_ws(Ctx &ctx, void *state, int prec, unsigned nbits, const DiagGate *gates, size_t ngates, double *ws,
                  size_t ws_bytes);
// Device address and size of the no-allocation entry point's workspace on
// the current device (looked up once per device, allocates nothing).
int noalloc_workspace(void **ptr, size_t *bytes);
// The process-wide context if it exists (nullptr otherwise; creates nothing).
Ctx *default_ctx_if_ready();

struct U3Gate
{
    uint64_t packed_sel; // alpha | beta << 63
    uint64_t delta;
    double u[8];
};

// Tile geometry (tuning knob tile.seg.<prec> = 3, 4, 5 or 6 contiguous
// 16-byte-unit address bits): a tile pass reaches every target below tile_low_bits(prec)
// and up to tile_upper_slots(prec) more; it needs a state of at least
// tile_min_bits(prec) bits (12 for complex128, 13 for complex64).
// geometry: 3..6 for that geometry, anything else for the knob's.
unsigned tile_seg(int prec, int geometry = -1);
unsigned tile_low_bits(int prec, int geometry = -1);
unsigned tile_upper_slots(int prec, int geometry = -1);
unsigned tile_min_bits(int prec);
// True when single-bit targets (amplitude address bits) fit one wide tile
// pass (knob tile.wide.<prec> = 1): at most nine, distinct, all above the
// 16-unit contiguous run, on a state of at least 13 unit bits.  Such a pass
// moves 256-byte runs where the 4096-unit tile would move 128-byte ones.
bool tile_wide_fits(int prec, unsigned n, const std::vector<unsigned> &targets);
// The passes of a batch of single-bit gates on distinct targets that needs
// more than one 4096-unit tile pass under geometry, with the wide tile on:
// passes[0] for a 4096-unit tile pass (every target below tile_low_bits and
// at most tile_upper_slots others), then as few wide passes as the batch
// allows, each fitting tile_wide_fits; when no target lies below the wide
// tile's run and wide passes alone take fewer passes, wide passes only.  The slots are spread over address
// bits, not taken in runs of consecutive bits (which move the state slower).
// False when the batch fits one pass or the wide tile does not apply.
bool tile_layer_plan(int prec, unsigned n, int geometry, const std::vector<unsigned> &targets,
                     std::vector<std::vector<unsigned>> &passes);
// True when single-bit targets in this (circuit) order fit one staged tile
// pass of a 2^n complex128 state: the order a pass with fans keeps.
bool tile_staged_fits(int prec, unsigned n, const std::vector<unsigned> &targets, int geometry = -1);
// A diagonal window whose selector forms span at most kWinRankMax
// dimensions over GF(2) is a table over those forms: a tile pass can apply
// it as its epilogue whatever its selectors (several bits, after CX
// transport) and whether or not the pass has per-pair selectors.
constexpr int kWinRankMax = 10;
bool window_tabulable(unsigned nbits, const DiagGate *gates, size_t ngates);
// The table a tile pass would use (host only): the basis forms (rank of
// them, over amplitude addresses) and 2^rank (cos, sin) entries.  False when
// the window is not tabulable.
bool window_table_host(unsigned nbits, const DiagGate *gates, size_t ngates, uint64_t *basis, int *rank,
                       std::vector<double> &table);
// Applies commuting gates on distinct targets (a fused U3 cluster), then the
// diagonal window win (may be empty): as the last tile pass's epilogue when
// the window is tabulable (or a phase polynomial on a pass without per-pair
// selectors), else in a sweep of its own.
int u3_apply(Ctx &ctx, void *state, int prec, unsigned nbits, const U3Gate *gates, size_t ngates,
             const DiagGate *win = nullptr, size_t nwin = 0);
// A phase applied just before a gate of a tile pass to the amplitudes whose
// target bit is one: exp(i (c + sum_k theta[k] x_k)), theta over the state's
// address bits (theta[target] is ignored).  A diagonal window that lies
// between two single-bit gates and involves the later one's target becomes
// that gate's fan, so the gates can share a pass.
struct GateFan
{
    bool on = false;
    double c = 0.0;
    std::vector<double> theta;
};
// One tile pass applying single-bit U3 gates (alpha == delta, any number of
// targets below tile_low_bits and at most tile_upper_slots at or above),
// then a diagonal window (a phase polynomial, or tabulable and without
// fans), in one sweep.  Without fans the gates must commute; with fans (complex128, every
// target at or above tile_low_bits) they are applied in the given order,
// each after its fan.  fresh: the state is |0...0> of 2^nbits amplitudes but
// not yet written (a reset left pending); the pass then forms its input
// instead of reading it where it can, and resets the state first where not.
int fused_pass_apply(Ctx &ctx, void *state, int prec, unsigned nbits, const U3Gate *gates, size_t ngates,
                     const DiagGate *diag, size_t ndiag, int geometry = -1, const GateFan *fans = nullptr,
                     bool fresh = false);
// One target of an adjoint pass: the uncompute matrix m (= U^dagger of the
// forward pass's fused single-qubit gate on this qubit, (re, im) x 4), up to
// two transition operators a[k] (2x2 traceless Hermitian; zero when unused)
// whose values Im <lam| a[k] |psi> are added to grads[param[k]] (param < 0:
// none; param >= nparams is rejected).
struct AdjTarget
{
    unsigned target;
    double m[8];
    double a[2][8];
    int param[2];
};
// The reverse of one forward tile pass on psi and lam together (plain
// addressing, single-bit targets): transition values, then m on both states,
// then the diagonal window (given with the angles to apply) on both.
// dgrads: device array of parameter gradients, accumulated in stream order.
int adjoint_pass_apply(Ctx &ctx, void *psi, void *lam, int prec, unsigned nbits, const AdjTarget *targets,
                       size_t ntargets, const DiagGate *diag, size_t ndiag, double *dgrads, size_t nparams,
                       int geometry = -1);
// Dense 5-qubit block: y[t] = sum_t' u[t][t'] x[t'] over every group of 32
// amplitudes differing in the target bits (bit b of t = targets[b]); u is
// 32x32 complex, row-major (re, im) pairs.  One sweep, matrix cores.
int block5_apply(Ctx &ctx, void *state, int prec, unsigned nbits, const unsigned *targets, const double *u);
// ---- batched tile passes (many independent states per launch) ----
// One state's pass in a batched launch.  Forward launches use gates (the
// U3 gates of the pass); adjoint launches use targets (AdjTarget); window
// is the epilogue with the angles to apply.  The tables for the matrices are
// written per run (batch_set_*), the rest is fixed at prepare time.
struct BatchPassSpec
{
    void *psi = nullptr;
    void *lam = nullptr; // adjoint launches
    unsigned n = 0;
    std::vector<unsigned> targets;
    std::vector<DiagGate> window;
    double *grads = nullptr; // adjoint: the state's gradient array (device)
    size_t nparams = 0;      // adjoint: its length
    int geometry = -1;       // tile geometry of the state's passes (-1: the knob's)
};
struct BatchLaunch
{
    int prec = 0;
    bool adj = false;
    bool l1 = false, l2 = false; // some item uses register layout 1 / 2
    size_t nitems = 0;
    uint64_t total_tiles = 0;
    int grid = 0;
    void *d_items = nullptr, *d_tabs = nullptr, *d_adjs = nullptr, *d_partial = nullptr, *d_collect = nullptr,
         *d_epi = nullptr;
    void *h_tables = nullptr; // pinned mirror of d_tabs | d_adjs (d_adjs points into d_tabs)
    size_t tab_bytes = 0, adj_bytes = 0;
    std::vector<std::vector<int>> slots; // per item: slot of each target
    std::vector<size_t> nparams;         // adjoint: per item, length of its gradient array
    ~BatchLaunch();
};
int batch_prepare(Ctx &ctx, int prec, bool adj, const std::vector<BatchPassSpec> &specs, BatchLaunch &out);
// Matrices of item i for this run (into the pinned mirror).
void batch_set_gates(BatchLaunch &b, size_t item, const U3Gate *gates, size_t ngates);
int batch_set_adjoint(BatchLaunch &b, size_t item, const AdjTarget *targets, size_t ntargets);
// Copies the mirror to the device tables (async on ctx.stream).
int batch_upload(Ctx &ctx, BatchLaunch &b);
// The batched kernel (and, for adjoint launches, the gradient collection).
int batch_launch(Ctx &ctx, BatchLaunch &b);
// Batched observables: per state E, <psi|psi> (into energies[2 i], device)
// and, when dst is set, dst = H psi.  Block coefficients fixed at prepare.
struct ObsSpec
{
    const void *src = nullptr;
    void *dst = nullptr;
    unsigned n = 0;
    std::vector<int32_t> qi, qj;
    std::vector<double> coef;
};
struct ObsBatch
{
    int prec = 0;
    size_t nitems = 0;
    uint64_t total_blocks = 0;
    int grid = 0;
    void *d_items = nullptr, *d_tables = nullptr, *d_partial = nullptr;
    std::vector<void *> bufs;
    ~ObsBatch();
};
int obs_batch_prepare(Ctx &ctx, int prec, const std::vector<ObsSpec> &specs, double *d_energies, ObsBatch &out);
int obs_batch_launch(Ctx &ctx, ObsBatch &b);
// Graph-capturable fills: zero bytes (multiple of 16) and amplitude 0 = 1 of
// each state in the device array dstates.
int zero_fill(Ctx &ctx, void *p, size_t bytes);
int set_ones(Ctx &ctx, const uint64_t *dstates, int count, int prec);
// out (host) = the amplitudes at idx[0..count) of a 2^n state (synchronous).
int gather(Ctx &ctx, const void *state, int prec, unsigned n, const uint64_t *idx, size_t count, void *out);
// One general pair sweep (any delta / selector).
int u3_pair_apply(Ctx &ctx, void *state, int prec, unsigned nbits, const U3Gate &g);
// The U3 of a qubit expansion: delta is the state's top bit (nbits - 1) and
// that half of the state is zero (it reads the lower half only).
int u3_expand_apply(Ctx &ctx, void *state, int prec, unsigned nbits, const U3Gate &g);
// count (at most four) expansions in a row from a state of base bits: gate
// i expands bit base + i (delta = 1 << (base + i), selector below bit
// base + i + 1); the region above bit base is zero.  Same products as count
// u3_expand_apply calls; one read of the old state, one write of the new.
int u3_expand_group_apply(Ctx &ctx, void *state, int prec, unsigned base, const U3Gate *gates, size_t count);

int x_apply(Ctx &ctx, void *state, int prec, unsigned nbits, uint64_t delta, uint64_t packed_sel);
int cx_apply(Ctx &ctx, void *state, int prec, unsigned nbits, uint64_t delta, uint64_t sel_cq, uint64_t sel_q);
int cp_apply(Ctx &ctx, void *state, int prec, unsigned nbits, uint64_t sel_cq, uint64_t sel_q, double theta);
int mcx_apply(Ctx &ctx, void *state, int prec, unsigned nbits, uint64_t ctrl_mask, unsigned target);

// State helpers.
int state_reset(Ctx &ctx, void *state, int prec, size_t alloc_elems);
// Streaming helpers: device copy (bandwidth calibration) and in-place scale.
int copy_apply(Ctx &ctx, void *dst, const void *src, size_t bytes);
int scale_apply(Ctx &ctx, void *state, int prec, unsigned nbits, double re, double im);

// Sampling in two steps: prepare builds the probability tree of 2^m
// amplitudes and returns the total (the reference's local_prob); search
// resolves sorted thresholds (host array) against it.
int sample_prepare(Ctx &ctx, const void *state, int prec, unsigned m, unsigned seg_bits, double *total);
// sample_prepare in two halves, so that host work can overlap the tree build:
// launch enqueues it (and the copy of the root to pinned memory); total waits.
int sample_launch(Ctx &ctx, const void *state, int prec, unsigned m, unsigned seg_bits);
int sample_total(Ctx &ctx, double *total);
// Optional packing of the samples into virtual result words on the device:
// bit q = parity(real & rows[q]) ^ b[q]; out (host) gets ceil(nq / 64) words
// per shot.
struct PackSpec
{
    unsigned nq = 0;
    const uint64_t *rows = nullptr;
    const uint8_t *b = nullptr;
    uint64_t *out = nullptr;
};
int sample_search(Ctx &ctx, const void *state, int prec, const double *thresholds, size_t shots, uint64_t *out_real,
                  const PackSpec *pack = nullptr);
// measure (measure_common.cpp:73-92) on a caller's state of 2^m amplitudes
// and affine map (rows[q] bit c = A[q][c], b[q]; nq virtual qubits):
// thresholds from mix_seed(use_seed ? seed : random_device, 0), segments of
// 2^min(m, seg_bits); packed gets ceil(nq / 64) words per shot, out_real
// (nullable) the real addresses, ascending.
int measure_affine(Ctx &ctx, const void *state, int prec, unsigned m, unsigned seg_bits, unsigned nq,
                   const uint64_t *rows, const uint8_t *b, bool use_seed, uint64_t seed, size_t shots,
                   uint64_t *packed, uint64_t *out_real);
// Both steps with caller-given thresholds.
int sample_apply(Ctx &ctx, const void *state, int prec, unsigned m, unsigned seg_bits, const double *thresholds,
                 size_t shots, uint64_t *out_real);
// Root of the pairwise |psi|^2 tree (the reference's total probability).
int prob_total(Ctx &ctx, const void *state, int prec, unsigned m, double *out);

// Diagonal Z / ZZ observable H = sum_k coef[k] Z_qi[k] Z_qj[k] (qj[k] < 0:
// one Z) on a state of n address bits (qubit q = address bit q).  out2
// (host, nullable) receives <psi|H|psi> and <psi|psi>, both accumulated in
// FP64 with compensation.  With dst non-null also writes dst = H psi.
int observable_apply(Ctx &ctx, const void *src, void *dst, int prec, unsigned nbits, const int32_t *qi,
                     const int32_t *qj, const double *coef, size_t nterms, double *out2);

// Compare two states of 2^n amplitudes (either precision each): out5 (host)
// receives Re <a|b>, Im <a|b>, |a|^2, |b|^2 (FP64, compensated, fixed order)
// and max_i |a_i - b_i|.
int compare_states(Ctx &ctx, const void *a, int prec_a, const void *b, int prec_b, unsigned n, double *out5);

// Host threshold stream (measure_common.cpp:14-71) from an already mixed seed.
void threshold_stream(uint64_t stream_seed, size_t cnt, double begin, double end, double *out);
// The same stream in three steps.  draws: the cnt uniforms u_k (mt19937_64,
// uniform_real_distribution, clamped below by DBL_MIN), in order.  factors:
// in place, f_k = 1 - u_k^(1 / (cnt - k)), which does not depend on the
// interval, on host threads.  apply: in place, the recurrence
// head_k = lo + (end - lo) * f_k, lo = head_k, from lo = begin.
void threshold_draws(uint64_t stream_seed, size_t cnt, double *u);
void threshold_factors(double *uf, size_t cnt);
void threshold_apply(double *fh, size_t cnt, double begin, double end);
uint64_t mix_seed(uint64_t base, uint64_t salt);

} // namespace zxa
