// capi.cpp -- the extern "C" boundary declared in include/zxh_amd.h.
#include "../../include/zxh_amd.h"

#include "engine.h"
#include "ansatz.h"
#include "batch.h"

#include <algorithm>
#include <cstring>
#include <new>

using zxa::Sim;

struct zxh_sim
{
    Sim *sim;
};

struct zxh_ansatz
{
    zxa::Ansatz *a;
};

struct zxh_batch
{
    zxa::AnsatzBatch *b;
};

namespace
{
int fail(int code, const char *msg)
{
    zxa::set_error(msg);
    return code;
}
#define CHECK_SIM(s)                                                                                                   \
    do                                                                                                                 \
    {                                                                                                                  \
        if (!(s) || !(s)->sim)                                                                                         \
            return fail(ZXH_EINVAL, "null simulator handle");                                                         \
    } while (0)
} // namespace

extern "C" {

const char *zxh_last_error(void)
{
    return zxa::last_error();
}

int zxh_init(void)
{
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0)
        return fail(ZXH_ENODEV, "no HIP device available");
    zxa::runtime_set_active(true);
    // the kernel-level context and, on the current device, the workspace of
    // the no-allocation entry point
    int rc = 0;
    if (!zxa::default_ctx(&rc))
        return rc;
    void *ws = nullptr;
    size_t ws_bytes = 0;
    return zxa::noalloc_workspace(&ws, &ws_bytes);
}

int zxh_finalize(void)
{
    zxa::runtime_set_active(false);
    return ZXH_OK;
}

int zxh_active(void)
{
    return zxa::runtime_active() ? 1 : 0;
}

int zxh_rank(void)
{
    return 0;
}

int zxh_nprocs(void)
{
    return 1;
}

int zxh_create(size_t n, int disable_x, int eager_expand_all, int precision, zxh_sim **out)
{
    if (!out)
        return fail(ZXH_EINVAL, "null output handle");
    *out = nullptr;
    if (!zxa::runtime_active())
        return fail(ZXH_ESTATE, "init must be called before constructing or using ZXH objects");
    if (precision != ZXH_C128 && precision != ZXH_C64)
        return fail(ZXH_EINVAL, "precision must be ZXH_C128 or ZXH_C64");
    Sim *sim = new (std::nothrow) Sim(n, disable_x != 0, eager_expand_all != 0, precision);
    if (!sim)
        return fail(ZXH_ENOMEM, "host allocation failed");
    int rc = sim->init();
    if (rc)
    {
        delete sim;
        return rc;
    }
    zxh_sim *h = new (std::nothrow) zxh_sim{sim};
    if (!h)
    {
        delete sim;
        return fail(ZXH_ENOMEM, "host allocation failed");
    }
    *out = h;
    return ZXH_OK;
}

int zxh_destroy(zxh_sim *s)
{
    if (!s)
        return ZXH_OK;
    delete s->sim;
    delete s;
    return ZXH_OK;
}

int zxh_set_seed(zxh_sim *s, uint64_t seed)
{
    CHECK_SIM(s);
    s->sim->set_seed(seed);
    return ZXH_OK;
}

int zxh_clear_seed(zxh_sim *s)
{
    CHECK_SIM(s);
    s->sim->clear_seed();
    return ZXH_OK;
}

int zxh_clear_gates(zxh_sim *s)
{
    CHECK_SIM(s);
    s->sim->clear_gates();
    return ZXH_OK;
}

int zxh_gate(zxh_sim *s, int code, size_t cq, size_t q, double theta, double lambda, double phi)
{
    CHECK_SIM(s);
    return s->sim->add_gate(code, cq, q, theta, lambda, phi);
}

int zxh_barrier(zxh_sim *s)
{
    return zxh_gate(s, ZXH_G_BARRIER, 0, 0, 0.0, 0.0, 0.0);
}
int zxh_rz(zxh_sim *s, size_t q, double theta)
{
    return zxh_gate(s, ZXH_G_RZ, 0, q, theta, 0.0, 0.0);
}
int zxh_crz(zxh_sim *s, size_t cq, size_t q, double theta)
{
    return zxh_gate(s, ZXH_G_CRZ, cq, q, theta, 0.0, 0.0);
}
int zxh_cp(zxh_sim *s, size_t cq, size_t q, double theta)
{
    return zxh_gate(s, ZXH_G_CP, cq, q, theta, 0.0, 0.0);
}
int zxh_p(zxh_sim *s, size_t q, double theta)
{
    return zxh_gate(s, ZXH_G_P, 0, q, theta, 0.0, 0.0);
}
int zxh_z(zxh_sim *s, size_t q)
{
    return zxh_gate(s, ZXH_G_Z, 0, q, 0.0, 0.0, 0.0);
}
int zxh_x(zxh_sim *s, size_t q)
{
    return zxh_gate(s, ZXH_G_X, 0, q, 0.0, 0.0, 0.0);
}
int zxh_cx(zxh_sim *s, size_t cq, size_t q)
{
    return zxh_gate(s, ZXH_G_CX, cq, q, 0.0, 0.0, 0.0);
}
int zxh_h(zxh_sim *s, size_t q)
{
    return zxh_gate(s, ZXH_G_H, 0, q, 0.0, 0.0, 0.0);
}
int zxh_u3(zxh_sim *s, size_t q, double theta, double lambda, double phi)
{
    return zxh_gate(s, ZXH_G_U3, 0, q, theta, lambda, phi);
}
int zxh_rx(zxh_sim *s, size_t q, double theta)
{
    return zxh_gate(s, ZXH_G_RX, 0, q, theta, 0.0, 0.0);
}

int zxh_execute(zxh_sim *s)
{
    CHECK_SIM(s);
    return s->sim->execute();
}

int zxh_measure(zxh_sim *s, size_t shots)
{
    CHECK_SIM(s);
    return s->sim->measure(shots);
}

int zxh_measured_count(const zxh_sim *s, size_t *out)
{
    CHECK_SIM(s);
    *out = s->sim->measured_count();
    return ZXH_OK;
}

int zxh_get_results(const zxh_sim *s, uint64_t *packed, size_t shots)
{
    CHECK_SIM(s);
    return s->sim->get_results(packed, shots);
}

int zxh_get_real_results(const zxh_sim *s, uint64_t *real, size_t shots)
{
    CHECK_SIM(s);
    return s->sim->get_real_results(real, shots);
}

int zxh_get_state(const zxh_sim *s, double *out, size_t out_len)
{
    CHECK_SIM(s);
    return s->sim->get_state(out, out_len);
}

int zxh_required_M(const zxh_sim *s, size_t *out)
{
    CHECK_SIM(s);
    *out = s->sim->required_M();
    return ZXH_OK;
}

int zxh_num_qubits(const zxh_sim *s, size_t *out)
{
    CHECK_SIM(s);
    *out = s->sim->num_qubits();
    return ZXH_OK;
}

int zxh_set_policy(zxh_sim *s, int policy)
{
    CHECK_SIM(s);
    if (policy != ZXH_POLICY_REFERENCE && policy != ZXH_POLICY_FUSED)
        return fail(ZXH_EINVAL, "unknown batching policy");
    s->sim->set_policy(policy);
    return ZXH_OK;
}

int zxh_set_seg_bits(zxh_sim *s, size_t bits)
{
    CHECK_SIM(s);
    if (bits > 40)
        return fail(ZXH_EINVAL, "segment bits out of range");
    s->sim->set_seg_bits(bits);
    return ZXH_OK;
}

int zxh_get_phys_state(const zxh_sim *s, void *out, size_t out_bytes, size_t *m_out)
{
    CHECK_SIM(s);
    return s->sim->get_phys_state(out, out_bytes, m_out);
}

int zxh_set_phys_state(zxh_sim *s, const void *in, size_t in_bytes)
{
    CHECK_SIM(s);
    return s->sim->set_phys_state(in, in_bytes);
}

int zxh_state_device_ptr(const zxh_sim *s, void **dptr, size_t *m_out)
{
    CHECK_SIM(s);
    return s->sim->device_ptr(dptr, m_out);
}

int zxh_set_timing(zxh_sim *s, int on)
{
    CHECK_SIM(s);
    s->sim->ctx().timing = on != 0;
    return ZXH_OK;
}

int zxh_get_timing(zxh_sim *s, int family, uint64_t *launches, double *total_ms, uint64_t *bytes)
{
    CHECK_SIM(s);
    if (family < 0 || family >= zxa::F_COUNT)
        return fail(ZXH_EINVAL, "timing family out of range");
    int rc = s->sim->ctx().collect();
    if (rc)
        return rc;
    const zxa::FamilyTiming &f = s->sim->ctx().fam[family];
    *launches = f.launches;
    *total_ms = f.ms;
    *bytes = f.bytes;
    return ZXH_OK;
}

int zxh_reset_timing(zxh_sim *s)
{
    CHECK_SIM(s);
    int rc = s->sim->ctx().collect();
    for (auto &f : s->sim->ctx().fam)
        f = zxa::FamilyTiming();
    return rc;
}

// ---- kernel level ----

int zxh_dev_alloc(size_t bytes, void **dptr)
{
    if (!dptr)
        return fail(ZXH_EINVAL, "null output pointer");
    hipError_t e = hipMalloc(dptr, bytes);
    if (e != hipSuccess)
    {
        *dptr = nullptr;
        return fail(ZXH_ENOMEM, hipGetErrorString(e));
    }
    // zeroed, like worker_alloc (mem_cuda.cu:28-40)
    e = hipMemset(*dptr, 0, bytes);
    if (e != hipSuccess)
    {
        (void)hipFree(*dptr);
        *dptr = nullptr;
        return fail(ZXH_EHIP, hipGetErrorString(e));
    }
    return ZXH_OK;
}

int zxh_dev_free(void *dptr)
{
    return zxa::hip_check(hipFree(dptr), "hipFree");
}

int zxh_dev_upload(void *dptr, const void *host, size_t bytes)
{
    return zxa::hip_check(hipMemcpy(dptr, host, bytes, hipMemcpyHostToDevice), "hipMemcpy H2D");
}

int zxh_dev_download(void *host, const void *dptr, size_t bytes)
{
    return zxa::hip_check(hipMemcpy(host, dptr, bytes, hipMemcpyDeviceToHost), "hipMemcpy D2H");
}

int zxh_dev_realloc(void *old_ptr, size_t old_bytes, size_t new_bytes, void **new_ptr)
{
    if (!new_ptr)
        return fail(ZXH_EINVAL, "null output pointer");
    *new_ptr = nullptr;
    if (new_bytes == 0)
        return old_ptr ? zxa::hip_check(hipFree(old_ptr), "hipFree") : ZXH_OK;
    void *p = nullptr;
    int rc = zxh_dev_alloc(new_bytes, &p);
    if (rc)
        return rc;
    if (old_ptr)
    {
        const size_t keep = old_bytes < new_bytes ? old_bytes : new_bytes;
        rc = zxa::hip_check(hipMemcpy(p, old_ptr, keep, hipMemcpyDeviceToDevice), "hipMemcpy D2D");
        if (rc)
        {
            (void)hipFree(p);
            return rc;
        }
        rc = zxa::hip_check(hipFree(old_ptr), "hipFree");
        if (rc)
        {
            (void)hipFree(p);
            return rc;
        }
    }
    *new_ptr = p;
    return ZXH_OK;
}

int zxh_dev_set_zero(void *dptr, int precision, size_t begin, size_t end)
{
    if (!dptr)
        return fail(ZXH_EINVAL, "null pointer");
    if (precision != ZXH_C128 && precision != ZXH_C64)
        return fail(ZXH_EINVAL, "unknown precision");
    if (end < begin)
        return fail(ZXH_EINVAL, "set_zero: end before begin");
    const size_t esz = precision == ZXH_C128 ? 16 : 8;
    return zxa::hip_check(hipMemset(static_cast<char *>(dptr) + begin * esz, 0, (end - begin) * esz), "hipMemset");
}

int zxh_dev_mem_set(void *dptr, int precision, double re, double im)
{
    if (!dptr)
        return fail(ZXH_EINVAL, "null pointer");
    if (precision == ZXH_C128)
    {
        const double v[2] = {re, im};
        return zxa::hip_check(hipMemcpy(dptr, v, sizeof v, hipMemcpyHostToDevice), "hipMemcpy H2D");
    }
    if (precision == ZXH_C64)
    {
        const float v[2] = {static_cast<float>(re), static_cast<float>(im)};
        return zxa::hip_check(hipMemcpy(dptr, v, sizeof v, hipMemcpyHostToDevice), "hipMemcpy H2D");
    }
    return fail(ZXH_EINVAL, "unknown precision");
}

int zxh_dev_sync(void)
{
    int rc = 0;
    zxa::Ctx *c = zxa::default_ctx(&rc);
    if (!c)
        return rc;
    rc = zxa::hip_check(hipStreamSynchronize(c->stream), "stream sync");
    if (rc)
        return rc;
    return zxa::hip_check(hipDeviceSynchronize(), "device sync");
}

#define WITH_CTX(c)                                                                                                    \
    int rc_ctx_ = 0;                                                                                                   \
    zxa::Ctx *c = zxa::default_ctx(&rc_ctx_);                                                                          \
    if (!c)                                                                                                            \
        return rc_ctx_;

static int finish(zxa::Ctx *c, int rc)
{
    if (rc)
        return rc;
    return zxa::hip_check(hipStreamSynchronize(c->stream), "stream sync");
}

static int check_kernel_args(const void *state, int precision, size_t local_bits)
{
    if (!state)
        return fail(ZXH_EINVAL, "null state pointer");
    if (precision != ZXH_C128 && precision != ZXH_C64)
        return fail(ZXH_EINVAL, "precision must be ZXH_C128 or ZXH_C64");
    if (local_bits > 40)
        return fail(ZXH_EINVAL, "local_bits out of range");
    return ZXH_OK;
}

int zxh_k_u3_batch(void *state, int precision, size_t local_bits, const zxh_u3_desc *descs, size_t ndesc)
{
    int rc = check_kernel_args(state, precision, local_bits);
    if (rc)
        return rc;
    WITH_CTX(c);
    static_assert(sizeof(zxh_u3_desc) == sizeof(zxa::U3Gate), "descriptor layout");
    return finish(c, zxa::u3_apply(*c, state, precision, unsigned(local_bits),
                                   reinterpret_cast<const zxa::U3Gate *>(descs), ndesc));
}

int zxh_k_fused_pass(void *state, int precision, size_t local_bits, const zxh_u3_desc *descs, size_t ndesc,
                     const uint64_t *sel0, const uint64_t *sel1, const double *theta, size_t ndiag)
{
    int rc = check_kernel_args(state, precision, local_bits);
    if (rc)
        return rc;
    WITH_CTX(c);
    std::vector<zxa::DiagGate> g(ndiag);
    for (size_t i = 0; i < ndiag; i++)
        g[i] = {sel0[i], sel1[i], theta[i]};
    return finish(c, zxa::fused_pass_apply(*c, state, precision, unsigned(local_bits),
                                           reinterpret_cast<const zxa::U3Gate *>(descs), ndesc, g.data(), ndiag));
}

int zxh_window_table(size_t nbits, const uint64_t *sel0, const uint64_t *sel1, const double *theta, size_t ngates,
                     uint64_t *basis, int *rank, double *table)
{
    if ((ngates && (!sel0 || !sel1 || !theta)) || !basis || !rank || !table)
        return fail(ZXH_EINVAL, "null argument");
    if (nbits > 40)
        return fail(ZXH_EINVAL, "window table: state wider than 40 bits");
    std::vector<zxa::DiagGate> g(ngates);
    for (size_t i = 0; i < ngates; i++)
        g[i] = {sel0[i], sel1[i], theta[i]};
    std::vector<double> tab;
    if (!zxa::window_table_host(unsigned(nbits), g.data(), ngates, basis, rank, tab))
        return fail(ZXH_EINVAL, "window table: the selector forms span more than 10 dimensions or reach above the "
                                "state");
    std::memcpy(table, tab.data(), tab.size() * sizeof(double));
    return ZXH_OK;
}

int zxh_tile_layer_plan(int precision, size_t nbits, int geometry, const uint32_t *targets, size_t ntargets,
                        uint32_t *pass_of, size_t *npasses)
{
    if ((ntargets && (!targets || !pass_of)) || !npasses)
        return fail(ZXH_EINVAL, "null argument");
    if (precision != ZXH_C128 && precision != ZXH_C64)
        return fail(ZXH_EINVAL, "precision must be ZXH_C128 or ZXH_C64");
    if (nbits > 40)
        return fail(ZXH_EINVAL, "layer plan: state wider than 40 bits");
    *npasses = 0;
    const std::vector<unsigned> ts(targets, targets + ntargets);
    std::vector<std::vector<unsigned>> plan;
    if (!zxa::tile_layer_plan(precision, unsigned(nbits), geometry, ts, plan))
        return ZXH_OK;
    for (size_t k = 0; k < plan.size(); k++)
        for (unsigned t : plan[k])
            pass_of[size_t(std::find(ts.begin(), ts.end(), t) - ts.begin())] = uint32_t(k);
    *npasses = plan.size();
    return ZXH_OK;
}

int zxh_k_diag(void *state, int precision, size_t local_bits, const uint64_t *sel0, const uint64_t *sel1,
               const double *theta, size_t ngates)
{
    int rc = check_kernel_args(state, precision, local_bits);
    if (rc)
        return rc;
    WITH_CTX(c);
    std::vector<zxa::DiagGate> g(ngates);
    for (size_t i = 0; i < ngates; i++)
        g[i] = {sel0[i], sel1[i], theta[i]};
    return finish(c, zxa::diag_apply(*c, state, precision, unsigned(local_bits), g.data(), ngates));
}

int zxh_k_diag_noalloc(void *state, int precision, size_t local_bits, const uint64_t *sel0, const uint64_t *sel1,
                       const double *theta, size_t ngates)
{
    int rc = check_kernel_args(state, precision, local_bits);
    if (rc)
        return rc;
    if (ngates && (!sel0 || !sel1 || !theta))
        return fail(ZXH_EINVAL, "null gate arrays");
    zxa::Ctx *c = zxa::default_ctx_if_ready();
    if (!c)
        return fail(ZXH_ESTATE, "zxh_k_diag_noalloc: call zxh_init() first (it creates the device context)");
    void *ws = nullptr;
    size_t ws_bytes = 0;
    rc = zxa::noalloc_workspace(&ws, &ws_bytes);
    if (rc)
        return rc;
    std::vector<zxa::DiagGate> g(ngates);
    for (size_t i = 0; i < ngates; i++)
        g[i] = {sel0[i], sel1[i], theta[i]};
    return finish(c, zxa::diag_apply_ws(*c, state, precision, unsigned(local_bits), g.data(), ngates,
                                        static_cast<double *>(ws), ws_bytes));
}

int zxh_k_x(void *state, int precision, size_t local_bits, uint64_t delta, uint64_t packed_sel)
{
    int rc = check_kernel_args(state, precision, local_bits);
    if (rc)
        return rc;
    WITH_CTX(c);
    return finish(c, zxa::x_apply(*c, state, precision, unsigned(local_bits), delta, packed_sel));
}

int zxh_k_cx(void *state, int precision, size_t local_bits, uint64_t delta, uint64_t packed_sel_cq,
             uint64_t packed_sel_q)
{
    int rc = check_kernel_args(state, precision, local_bits);
    if (rc)
        return rc;
    WITH_CTX(c);
    return finish(c, zxa::cx_apply(*c, state, precision, unsigned(local_bits), delta, packed_sel_cq, packed_sel_q));
}

int zxh_k_cp(void *state, int precision, size_t local_bits, uint64_t packed_sel_cq, uint64_t packed_sel_q,
             double theta)
{
    int rc = check_kernel_args(state, precision, local_bits);
    if (rc)
        return rc;
    WITH_CTX(c);
    return finish(c, zxa::cp_apply(*c, state, precision, unsigned(local_bits), packed_sel_cq, packed_sel_q, theta));
}

int zxh_k_mcx(void *state, int precision, size_t local_bits, uint64_t ctrl_mask, size_t target)
{
    int rc = check_kernel_args(state, precision, local_bits);
    if (rc)
        return rc;
    WITH_CTX(c);
    return finish(c, zxa::mcx_apply(*c, state, precision, unsigned(local_bits), ctrl_mask, unsigned(target)));
}

int zxh_k_sample(const void *state, int precision, size_t m, size_t seg_bits, uint64_t stream_seed, size_t shots,
                 uint64_t *out_real)
{
    int rc = check_kernel_args(state, precision, m);
    if (rc)
        return rc;
    if (shots == 0)
        return ZXH_OK;
    WITH_CTX(c);
    double total = 0.0;
    rc = zxa::sample_prepare(*c, state, precision, unsigned(m), unsigned(seg_bits), &total);
    if (rc)
        return rc;
    if (!(total > 0.0))
        return fail(ZXH_EZEROPROB, "state has zero total probability");
    std::vector<double> thr(shots);
    zxa::threshold_stream(stream_seed, shots, 0.0, total, thr.data());
    return zxa::sample_search(*c, state, precision, thr.data(), shots, out_real);
}

int zxh_k_measure(const void *state, int precision, size_t m, size_t seg_bits, size_t nq, const uint64_t *rows,
                  const uint8_t *b, int use_seed, uint64_t seed, size_t shots, uint64_t *packed, uint64_t *out_real)
{
    int rc = check_kernel_args(state, precision, m);
    if (rc)
        return rc;
    if (nq == 0 || (nq && (!rows || !b)))
        return fail(ZXH_EINVAL, "measure: null affine map");
    if (shots && !packed)
        return fail(ZXH_EINVAL, "null output");
    for (size_t q = 0; q < nq; q++)
        if (m < 64 && (rows[q] >> m) != 0)
            return fail(ZXH_EINVAL, "measure: affine row reaches above the state's width");
    if (shots == 0)
        return ZXH_OK;
    WITH_CTX(c);
    return zxa::measure_affine(*c, state, precision, unsigned(m), unsigned(seg_bits), unsigned(nq), rows, b,
                               use_seed != 0, seed, shots, packed, out_real);
}

int zxh_k_sample_thresholds(const void *state, int precision, size_t m, size_t seg_bits, const double *thresholds,
                            size_t shots, uint64_t *out_real)
{
    int rc = check_kernel_args(state, precision, m);
    if (rc)
        return rc;
    WITH_CTX(c);
    return zxa::sample_apply(*c, state, precision, unsigned(m), unsigned(seg_bits), thresholds, shots, out_real);
}

int zxh_k_prob_total(const void *state, int precision, size_t m, double *out)
{
    int rc = check_kernel_args(state, precision, m);
    if (rc)
        return rc;
    WITH_CTX(c);
    return zxa::prob_total(*c, state, precision, unsigned(m), out);
}

int zxh_k_gather(const void *state, int precision, size_t local_bits, const uint64_t *indices, size_t count,
                 void *out)
{
    int rc = check_kernel_args(state, precision, local_bits);
    if (rc)
        return rc;
    if (count && (!indices || !out))
        return fail(ZXH_EINVAL, "null pointer");
    WITH_CTX(c);
    return zxa::gather(*c, state, precision, unsigned(local_bits), indices, count, out);
}

int zxh_k_copy(void *dst, const void *src, size_t bytes)
{
    if (!dst || !src)
        return fail(ZXH_EINVAL, "null pointer");
    WITH_CTX(c);
    return finish(c, zxa::copy_apply(*c, dst, src, bytes));
}

int zxh_k_scale(void *state, int precision, size_t local_bits, double re, double im)
{
    int rc = check_kernel_args(state, precision, local_bits);
    if (rc)
        return rc;
    WITH_CTX(c);
    return finish(c, zxa::scale_apply(*c, state, precision, unsigned(local_bits), re, im));
}

static int observable_call(const void *src, void *dst, int precision, size_t local_bits, const int32_t *qi,
                           const int32_t *qj, const double *coef, size_t nterms, double *out)
{
    int rc = check_kernel_args(src, precision, local_bits);
    if (rc)
        return rc;
    if (nterms > 0 && (!qi || !qj || !coef))
        return fail(ZXH_EINVAL, "null observable arrays");
    if (dst && dst == src)
        return fail(ZXH_EINVAL, "observable: dst aliases src");
    WITH_CTX(c);
    return finish(c, zxa::observable_apply(*c, src, dst, precision, unsigned(local_bits), qi, qj, coef, nterms, out));
}

int zxh_k_expect_z(const void *state, int precision, size_t local_bits, const int32_t *qi, const int32_t *qj,
                   const double *coef, size_t nterms, double *out)
{
    if (!out)
        return fail(ZXH_EINVAL, "null output");
    return observable_call(state, nullptr, precision, local_bits, qi, qj, coef, nterms, out);
}

int zxh_k_apply_z(const void *src, void *dst, int precision, size_t local_bits, const int32_t *qi,
                  const int32_t *qj, const double *coef, size_t nterms, double *out)
{
    if (!dst)
        return fail(ZXH_EINVAL, "null destination");
    return observable_call(src, dst, precision, local_bits, qi, qj, coef, nterms, out);
}

int zxh_k_compare(const void *a, int precision_a, const void *b, int precision_b, size_t local_bits, double *out5)
{
    int rc = check_kernel_args(a, precision_a, local_bits);
    if (rc)
        return rc;
    rc = check_kernel_args(b, precision_b, local_bits);
    if (rc)
        return rc;
    if (!out5)
        return fail(ZXH_EINVAL, "null output");
    if (local_bits < 1)
        return fail(ZXH_EINVAL, "compare: state width must be at least 1 bit");
    WITH_CTX(c);
    return finish(c, zxa::compare_states(*c, a, precision_a, b, precision_b, unsigned(local_bits), out5));
}

int zxh_k_reset(void *state, int precision, size_t local_bits)
{
    int rc = check_kernel_args(state, precision, local_bits);
    if (rc)
        return rc;
    WITH_CTX(c);
    return finish(c, zxa::state_reset(*c, state, precision, size_t(1) << local_bits));
}

int zxh_k_block5(void *state, int precision, size_t local_bits, const uint32_t *targets, const double *u)
{
    int rc = check_kernel_args(state, precision, local_bits);
    if (rc)
        return rc;
    if (!targets || !u)
        return fail(ZXH_EINVAL, "null targets or matrix");
    unsigned t[5];
    for (int b = 0; b < 5; b++)
        t[b] = targets[b];
    WITH_CTX(c);
    return finish(c, zxa::block5_apply(*c, state, precision, unsigned(local_bits), t, u));
}

// ---- parameterized circuits ----

#define CHECK_ANS(x)                                                                                                   \
    if (!(x) || !(x)->a)                                                                                               \
        return fail(ZXH_EINVAL, "null ansatz handle");

int zxh_ansatz_create(size_t n, int precision, const zxh_ansatz_op *ops, size_t nops, size_t nparams,
                      zxh_ansatz **out)
{
    if (!out)
        return fail(ZXH_EINVAL, "null output handle");
    *out = nullptr;
    if (!zxa::runtime_active())
        return fail(ZXH_ESTATE, "init must be called before constructing or using ansatz objects");
    if (precision != ZXH_C128 && precision != ZXH_C64)
        return fail(ZXH_EINVAL, "precision must be ZXH_C128 or ZXH_C64");
    if (nops > 0 && !ops)
        return fail(ZXH_EINVAL, "null operation array");
    static_assert(sizeof(zxh_ansatz_op) == sizeof(zxa::AnsatzOp), "operation layout");
    zxa::Ansatz *a = new zxa::Ansatz(unsigned(std::min<size_t>(n, 64)), precision, nparams);
    int rc = a->init(reinterpret_cast<const zxa::AnsatzOp *>(ops), nops);
    if (rc)
    {
        delete a;
        return rc;
    }
    *out = new zxh_ansatz{a};
    return ZXH_OK;
}

int zxh_ansatz_destroy(zxh_ansatz *a)
{
    if (!a)
        return ZXH_OK;
    delete a->a;
    delete a;
    return ZXH_OK;
}

int zxh_ansatz_set_observable(zxh_ansatz *a, const int32_t *qi, const int32_t *qj, const double *coef,
                              size_t nterms)
{
    CHECK_ANS(a);
    if (nterms > 0 && (!qi || !qj || !coef))
        return fail(ZXH_EINVAL, "null observable arrays");
    return a->a->set_observable(qi, qj, coef, nterms);
}

int zxh_ansatz_forward(zxh_ansatz *a, const double *params, double *out)
{
    CHECK_ANS(a);
    if (!params || !out)
        return fail(ZXH_EINVAL, "null params or output");
    return a->a->forward(params, out);
}

int zxh_ansatz_backward(zxh_ansatz *a, const double *params, double *grads, double *out)
{
    CHECK_ANS(a);
    if (!params || !grads)
        return fail(ZXH_EINVAL, "null params or gradient output");
    return a->a->backward(params, grads, out);
}

int zxh_ansatz_state(zxh_ansatz *a, void **state, size_t *nbits)
{
    CHECK_ANS(a);
    if (state)
        *state = a->a->state();
    if (nbits)
        *nbits = a->a->nbits();
    return ZXH_OK;
}

int zxh_ansatz_passes(zxh_ansatz *a, size_t *tile_passes)
{
    CHECK_ANS(a);
    if (!tile_passes)
        return fail(ZXH_EINVAL, "null output");
    *tile_passes = a->a->tile_passes();
    return ZXH_OK;
}

int zxh_ansatz_set_timing(zxh_ansatz *a, int on)
{
    CHECK_ANS(a);
    a->a->ctx().timing = on != 0;
    return ZXH_OK;
}

int zxh_ansatz_get_timing(zxh_ansatz *a, int family, uint64_t *launches, double *total_ms, uint64_t *bytes)
{
    CHECK_ANS(a);
    if (family < 0 || family >= zxa::F_COUNT)
        return fail(ZXH_EINVAL, "timing family out of range");
    if (!launches || !total_ms || !bytes)
        return fail(ZXH_EINVAL, "null output");
    int rc = a->a->ctx().collect();
    if (rc)
        return rc;
    const zxa::FamilyTiming &f = a->a->ctx().fam[family];
    *launches = f.launches;
    *total_ms = f.ms;
    *bytes = f.bytes;
    return ZXH_OK;
}

int zxh_ansatz_reset_timing(zxh_ansatz *a)
{
    CHECK_ANS(a);
    int rc = a->a->ctx().collect();
    for (auto &f : a->a->ctx().fam)
        f = zxa::FamilyTiming();
    return rc;
}

#define CHECK_BATCH(x)                                                                                                 \
    if (!(x) || !(x)->b)                                                                                               \
        return fail(ZXH_EINVAL, "null batch handle");

int zxh_batch_create(int precision, zxh_batch **out)
{
    if (!out)
        return fail(ZXH_EINVAL, "null output handle");
    *out = nullptr;
    if (!zxa::runtime_active())
        return fail(ZXH_ESTATE, "init must be called before constructing or using batch objects");
    if (precision != ZXH_C128 && precision != ZXH_C64)
        return fail(ZXH_EINVAL, "precision must be ZXH_C128 or ZXH_C64");
    *out = new zxh_batch{new zxa::AnsatzBatch(precision)};
    return ZXH_OK;
}

int zxh_batch_destroy(zxh_batch *b)
{
    if (!b)
        return ZXH_OK;
    delete b->b;
    delete b;
    return ZXH_OK;
}

int zxh_batch_add(zxh_batch *b, size_t n, const zxh_ansatz_op *ops, size_t nops, size_t nparams,
                  const int32_t *qi, const int32_t *qj, const double *coef, size_t nterms)
{
    CHECK_BATCH(b);
    if ((nops > 0 && !ops) || (nterms > 0 && (!qi || !qj || !coef)))
        return fail(ZXH_EINVAL, "null arrays");
    return b->b->add(unsigned(std::min<size_t>(n, 64)), reinterpret_cast<const zxa::AnsatzOp *>(ops), nops,
                     nparams, qi, qj, coef, nterms);
}

int zxh_batch_set_geometry(zxh_batch *b, size_t n, int seg)
{
    CHECK_BATCH(b);
    if (seg < 0)
        return fail(ZXH_EINVAL, "batch: negative geometry");
    return b->b->set_geometry(unsigned(std::min<size_t>(n, 65)), unsigned(seg));
}

int zxh_batch_finalize(zxh_batch *b)
{
    CHECK_BATCH(b);
    return b->b->finalize();
}

int zxh_batch_run(zxh_batch *b, const double *params, double *energies, double *grads)
{
    CHECK_BATCH(b);
    if (!params && b->b->total_params())
        return fail(ZXH_EINVAL, "null params");
    return b->b->run(params, energies, grads);
}

int zxh_batch_info(zxh_batch *b, size_t *sets, size_t *total_params, size_t *launches)
{
    CHECK_BATCH(b);
    if (sets)
        *sets = b->b->sets();
    if (total_params)
        *total_params = b->b->total_params();
    if (launches)
        *launches = b->b->launches();
    return ZXH_OK;
}

int zxh_batch_set_info(zxh_batch *b, size_t set, unsigned *nbits, size_t *passes)
{
    CHECK_BATCH(b);
    if (set >= b->b->sets())
        return fail(ZXH_EINVAL, "batch: set index out of range");
    if (nbits)
        *nbits = b->b->set_nbits(set);
    if (passes)
        *passes = b->b->set_passes(set);
    return ZXH_OK;
}

int zxh_tune_set(const char *name, int value)
{
    if (!name)
        return fail(ZXH_EINVAL, "null knob name");
    return zxa::tune_set(name, value);
}

int zxh_tune_get(const char *name, int *value)
{
    if (!name || !value)
        return fail(ZXH_EINVAL, "null argument");
    const int v = zxa::tune_get(name);
    if (v < 0)
        return fail(ZXH_EINVAL, "unknown tuning knob");
    *value = v;
    return ZXH_OK;
}

int zxh_k_set_timing(int on)
{
    WITH_CTX(c);
    c->timing = on != 0;
    return ZXH_OK;
}

int zxh_k_get_timing(int family, uint64_t *launches, double *total_ms, uint64_t *bytes)
{
    WITH_CTX(c);
    if (family < 0 || family >= zxa::F_COUNT)
        return fail(ZXH_EINVAL, "timing family out of range");
    int rc = c->collect();
    if (rc)
        return rc;
    *launches = c->fam[family].launches;
    *total_ms = c->fam[family].ms;
    *bytes = c->fam[family].bytes;
    return ZXH_OK;
}

int zxh_k_reset_timing(void)
{
    WITH_CTX(c);
    int rc = c->collect();
    for (auto &f : c->fam)
        f = zxa::FamilyTiming();
    return rc;
}

} // extern "C"
