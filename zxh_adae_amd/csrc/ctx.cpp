// ctx.cpp -- device context, scratch buffers, parameter ring, errors, and the
// host threshold stream of the sampler.
#include "dev.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <mutex>
#include <random>
#include <thread>

// The threshold recurrence must round every product and sum separately, as
// the reference's x86-64 build does.
#pragma clang fp contract(off)

namespace zxa
{

namespace
{
thread_local std::string g_err;
}

void set_error(const std::string &msg)
{
    g_err = msg;
}

const char *last_error()
{
    return g_err.c_str();
}

int hip_check(hipError_t e, const char *what)
{
    if (e == hipSuccess)
        return OK;
    set_error(std::string(what) + ": " + hipGetErrorString(e));
    return EHIP;
}

int DevBuf::ensure(size_t need)
{
    if (need <= bytes && ptr)
        return OK;
    release();
    size_t cap = std::max<size_t>(need, 4096);
    hipError_t e = hipMalloc(&ptr, cap);
    if (e != hipSuccess)
    {
        ptr = nullptr;
        bytes = 0;
        set_error(std::string("hipMalloc scratch failed: ") + hipGetErrorString(e));
        return ENOMEM_;
    }
    bytes = cap;
    return OK;
}

void DevBuf::release()
{
    if (ptr)
        (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
}

int ParamRing::init()
{
    if (host_)
        return OK;
    // mapped and coherent: the copy kernel reads the slots where the host
    // wrote them, uncached, so a reused slot is never read stale
    ZXA_HIP(hipHostMalloc(&host_, size_t(kSlots) * kSlotBytes, hipHostMallocMapped | hipHostMallocCoherent));
    ZXA_HIP(hipHostGetDevicePointer(&hdev_, host_, 0));
    ZXA_HIP(hipMalloc(&dev_, size_t(kSlots) * kSlotBytes));
    for (int i = 0; i < kSlots; i++)
        ZXA_HIP(hipEventCreateWithFlags(&ev_[i], hipEventDisableTiming));
    return OK;
}

int PinBuf::ensure(size_t need)
{
    if (need <= bytes && ptr)
        return OK;
    release();
    const size_t cap = std::max<size_t>(need, 4096);
    hipError_t e = hipHostMalloc(&ptr, cap, hipHostMallocDefault);
    if (e != hipSuccess)
    {
        ptr = nullptr;
        bytes = 0;
        set_error(std::string("hipHostMalloc failed: ") + hipGetErrorString(e));
        return ENOMEM_;
    }
    bytes = cap;
    return OK;
}

void PinBuf::release()
{
    if (ptr)
        (void)hipHostFree(ptr);
    ptr = nullptr;
    bytes = 0;
}

void ParamRing::release()
{
    for (int i = 0; i < kSlots; i++)
    {
        if (ev_[i])
        {
            if (used_[i])
                (void)hipEventSynchronize(ev_[i]);
            (void)hipEventDestroy(ev_[i]);
        }
        ev_[i] = nullptr;
        used_[i] = false;
    }
    if (dev_)
        (void)hipFree(dev_);
    if (host_)
        (void)hipHostFree(host_);
    dev_ = host_ = hdev_ = nullptr;
    open_.clear();
    end_batch();
}

// The next slot, once its last consumer has finished; open until done().
int ParamRing::open_slot(int *slot)
{
    const int s = next_;
    for (int o : open_)
        if (o == s)
        {
            set_error("parameter ring exhausted before done()");
            return EINVAL_;
        }
    next_ = (next_ + 1) % kSlots;
    if (used_[s])
        ZXA_HIP(hipEventSynchronize(ev_[s]));
    open_.push_back(s);
    *slot = s;
    return OK;
}

int ParamRing::stage(const void *src, size_t bytes, hipStream_t stream, void **dev)
{
    if (bytes > kSlotBytes)
    {
        set_error("parameter block exceeds ring slot size");
        return EINVAL_;
    }
    if (batch_ >= 0)
    {
        size_t off = (boff_ + 255) & ~size_t(255);
        if (off + bytes > kSlotBytes)
        {
            ZXA_TRY(flush(stream));
            ZXA_TRY(begin_batch());
            off = 0;
        }
        const size_t at = size_t(batch_) * kSlotBytes + off;
        std::memcpy(static_cast<char *>(host_) + at, src, bytes);
        boff_ = off + bytes;
        *dev = static_cast<char *>(dev_) + at;
        return OK;
    }
    int s = 0;
    ZXA_TRY(open_slot(&s));
    char *h = static_cast<char *>(host_) + size_t(s) * kSlotBytes;
    char *d = static_cast<char *>(dev_) + size_t(s) * kSlotBytes;
    std::memcpy(h, src, bytes);
    ZXA_TRY(ring_copy(stream, d, static_cast<char *>(hdev_) + size_t(s) * kSlotBytes, bytes));
    *dev = d;
    return OK;
}

int ParamRing::begin_batch()
{
    end_batch();
    int s = 0;
    ZXA_TRY(open_slot(&s));
    batch_ = s;
    return OK;
}

int ParamRing::flush(hipStream_t stream)
{
    const int s = batch_;
    const size_t bytes = boff_;
    end_batch();
    if (s < 0 || bytes == 0)
        return OK;
    const size_t at = size_t(s) * kSlotBytes;
    return ring_copy(stream, static_cast<char *>(dev_) + at, static_cast<char *>(hdev_) + at, bytes);
}

void ParamRing::end_batch()
{
    batch_ = -1;
    boff_ = 0;
}

int ParamRing::done(hipStream_t stream)
{
    for (int s : open_)
    {
        ZXA_HIP(hipEventRecord(ev_[s], stream));
        used_[s] = true;
    }
    open_.clear();
    return OK;
}

int Ctx::init()
{
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count <= 0)
    {
        set_error("no HIP device available");
        return ENODEV_;
    }
    ZXA_HIP(hipGetDevice(&device));
    hipDeviceProp_t prop;
    ZXA_HIP(hipGetDeviceProperties(&prop, device));
    num_cus = prop.multiProcessorCount;
    ZXA_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    return ring.init();
}

void Ctx::release()
{
    if (stream)
        (void)hipStreamSynchronize(stream);
    ring.release();
    scratch_a.release();
    scratch_b.release();
    scratch_c.release();
    scratch_d.release();
    scratch_e.release();
    scratch_f.release();
    scratch_g.release();
    (void)collect();
    for (hipEvent_t e : pool_)
        (void)hipEventDestroy(e);
    pool_.clear();
    if (stream)
        (void)hipStreamDestroy(stream);
    stream = nullptr;
}

int Ctx::take(hipEvent_t *e)
{
    if (!pool_.empty())
    {
        *e = pool_.back();
        pool_.pop_back();
        return OK;
    }
    return hip_check(hipEventCreate(e), "hipEventCreate");
}

int Ctx::tic()
{
    if (!timing)
        return OK;
    ZXA_TRY(take(&cur_));
    ZXA_HIP(hipEventRecord(cur_, stream));
    return OK;
}

int Ctx::toc(int family, uint64_t bytes)
{
    if (!timing || !cur_)
        return OK;
    hipEvent_t b = nullptr;
    ZXA_TRY(take(&b));
    ZXA_HIP(hipEventRecord(b, stream));
    pending_.push_back({family, bytes, cur_, b});
    cur_ = nullptr;
    if (pending_.size() >= 4096)
        return collect();
    return OK;
}

int Ctx::collect()
{
    for (const Pending &p : pending_)
    {
        ZXA_HIP(hipEventSynchronize(p.b));
        float ms = 0.f;
        ZXA_HIP(hipEventElapsedTime(&ms, p.a, p.b));
        fam[p.family].launches++;
        fam[p.family].ms += ms;
        fam[p.family].bytes += p.bytes;
        pool_.push_back(p.a);
        pool_.push_back(p.b);
    }
    pending_.clear();
    return OK;
}

namespace
{
std::mutex g_ctx_mu;
Ctx *g_ctx = nullptr;
} // namespace

Ctx *default_ctx_if_ready()
{
    std::lock_guard<std::mutex> lk(g_ctx_mu);
    return g_ctx;
}

Ctx *default_ctx(int *rc)
{
    std::lock_guard<std::mutex> lk(g_ctx_mu);
    Ctx *&ctx = g_ctx;
    if (!ctx)
    {
        Ctx *c = new Ctx();
        int r = c->init();
        if (r)
        {
            delete c;
            *rc = r;
            return nullptr;
        }
        ctx = c;
    }
    *rc = OK;
    return ctx;
}

namespace
{
struct Knob
{
    const char *name;
    int value;
    int lo, hi;
};
Knob g_knobs[] = {
    {"diag.ept.c128", 2, 0, 4},
    {"diag.ept.c64", 2, 0, 4},
    {"diag.nt.c128", 1, 0, 1},
    {"diag.nt.c64", 1, 0, 1},
    {"tile.seg.c128", 4, 3, 6},
    {"tile.seg.c64", 4, 3, 6},
    {"tile.bufs.c128", 2, 2, 3},
    {"tile.bufs.c64", 2, 2, 3},
    {"tile.wide.c128", 1, 0, 1},
    {"tile.wide.c64", 1, 0, 1},
    {"tile.upper.c128", 18, 0, 64},
    {"tile.upper.c64", 18, 0, 64},
    {"block.variant.c128", 4, 0, 7},
    {"block.variant.c64", 7, 0, 7},
    {"block.tiles.c128", 0, 0, 4096},
    {"block.tiles.c64", 0, 0, 4096},
};
std::mutex g_knob_mu;
} // namespace

int tune_get(const char *name)
{
    std::lock_guard<std::mutex> lk(g_knob_mu);
    for (const Knob &k : g_knobs)
        if (std::strcmp(k.name, name) == 0)
            return k.value;
    return -1;
}

int tune_set(const char *name, int value)
{
    std::lock_guard<std::mutex> lk(g_knob_mu);
    for (Knob &k : g_knobs)
        if (std::strcmp(k.name, name) == 0)
        {
            const bool ept_knob = std::strncmp(name, "diag.ept.", 9) == 0;
            if (value < k.lo || value > k.hi || (ept_knob && value == 3))
            {
                set_error(std::string("tuning value out of range for ") + name);
                return EINVAL_;
            }
            k.value = value;
            return OK;
        }
    set_error(std::string("unknown tuning knob ") + name);
    return EINVAL_;
}

// SplitMix64 finaliser (measure_general.cpp:16-22).
uint64_t mix_seed(uint64_t base, uint64_t salt)
{
    uint64_t z = base + 0x9e3779b97f4a7c15ULL + salt;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

// Sorted order statistics of cnt uniforms on [begin, end): each new head is
// begin + (end - begin) * (1 - u^(1/rest)) with u from mt19937_64 through
// uniform_real_distribution<double>(0, 1), clamped below by DBL_MIN
// (threshold_stream_t, measure_common.cpp:14-71).  The engine and the
// distribution are the standard library's, as in the reference.
void threshold_draws(uint64_t stream_seed, size_t cnt, double *u)
{
    std::mt19937_64 rng(stream_seed);
    std::uniform_real_distribution<double> dist(0.0, 1.0);
    for (size_t k = 0; k < cnt; k++)
        u[k] = std::max(dist(rng), std::numeric_limits<double>::min());
}

void threshold_factors(double *uf, size_t cnt)
{
    const auto run = [uf, cnt](size_t b, size_t e) {
        for (size_t k = b; k < e; k++)
            uf[k] = 1.0 - std::pow(uf[k], 1.0 / static_cast<double>(cnt - k));
    };
    const unsigned hw = std::thread::hardware_concurrency();
    const size_t nth = std::min<size_t>({hw ? hw : 1u, 16u, cnt / 16384 + 1});
    if (nth <= 1)
    {
        run(0, cnt);
        return;
    }
    std::vector<std::thread> pool;
    const size_t per = (cnt + nth - 1) / nth;
    for (size_t t = 1; t < nth; t++)
        pool.emplace_back(run, std::min(cnt, t * per), std::min(cnt, (t + 1) * per));
    run(0, std::min(cnt, per));
    for (auto &th : pool)
        th.join();
}

void threshold_apply(double *fh, size_t cnt, double begin, double end)
{
    double lo = begin;
    for (size_t k = 0; k < cnt; k++)
    {
        const double span = end - lo;
        const double head = lo + span * fh[k];
        fh[k] = head;
        lo = head;
    }
}

void threshold_stream(uint64_t stream_seed, size_t cnt, double begin, double end, double *out)
{
    threshold_draws(stream_seed, cnt, out);
    threshold_factors(out, cnt);
    threshold_apply(out, cnt, begin, end);
}

} // namespace zxa
