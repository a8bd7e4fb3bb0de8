// ref_harness.cpp -- C entry points over the compiled reference (checker only).
//
// Linked together with the reference's own CPU sources (single/omp backend,
// see oracle/Makefile) into oracle/_ref/libzxhref.so.  Used only in this
// container to write golden fixtures (tests/golden/make_golden.py) and by CPU
// tests that compare the oracle with the reference directly.  The product
// never links it.
//
// Drives the reference through its public class (include/zxhsim/zxh.h) and,
// for the standalone sampler pin, through its internal measurement API
// (include/zxhsim/measure.h, sv.h) on a state written by the caller.
#include "zxhsim/defs.h"
#include "zxhsim/measure.h"
#include "zxhsim/runtime.h"
#include "zxhsim/sv.h"
#include "zxhsim/zxh.h"

#include <cstdint>
#include <cstring>
#include <vector>

using namespace ZXHSim;

namespace
{
void ensure_init()
{
    if (!active())
        init(nullptr, nullptr);
}
} // namespace

extern "C" {

void *zr_create(size_t n, int disable_x, int eager)
{
    ensure_init();
    return new ZXH(n, disable_x != 0, eager != 0);
}

void zr_destroy(void *h)
{
    delete static_cast<ZXH *>(h);
}

void zr_set_seed(void *h, uint64_t seed)
{
    static_cast<ZXH *>(h)->set_seed(seed);
}

// Same gate codes as oracle/zxh_oracle.h.
int zr_gate(void *h, int type, size_t cq, size_t q, double theta, double lambda, double phi)
{
    ZXH *s = static_cast<ZXH *>(h);
    switch (type)
    {
    case 1: s->Barrier(); return 0;
    case 2: s->Z(q); return 0;
    case 3: s->Rz(q, theta); return 0;
    case 4: s->P(q, theta); return 0;
    case 5: s->CRz(cq, q, theta); return 0;
    case 6: s->CP(cq, q, theta); return 0;
    case 7: s->X(q); return 0;
    case 8: s->CX(cq, q); return 0;
    case 9: s->H(q); return 0;
    case 10: s->U3(q, theta, lambda, phi); return 0;
    case 11: s->Rx(q, theta); return 0;
    default: return 6;
    }
}

void zr_execute(void *h)
{
    static_cast<ZXH *>(h)->execute();
}

size_t zr_required_M(void *h)
{
    return static_cast<ZXH *>(h)->required_M();
}

// 2^N complex values (interleaved), bit-reversed order, times global phase.
void zr_get_state(void *h, double *out)
{
    const std::vector<val_t> st = static_cast<ZXH *>(h)->get_state();
    std::memcpy(out, st.data(), st.size() * sizeof(val_t));
}

// shots x N bits (one byte per bit).
void zr_sampling(void *h, size_t shots, uint8_t *out)
{
    ZXH *s = static_cast<ZXH *>(h);
    const std::vector<res_t> res = s->Sampling(shots);
    const size_t n = s->num_qubits();
    for (size_t i = 0; i < res.size(); i++)
        for (size_t q = 0; q < n; q++)
            out[i * n + q] = res[i].get_bit(q) ? 1 : 0;
}

// Sampler on a caller-written physical state of 2^m amplitudes with the
// identity affine map (real address == virtual address).  Segment size comes
// from the reference's layout rule (ZXHSIM_SEG_MIN_I is read from the
// environment of this process).  Writes cnt addresses.
void zr_sample_state(const double *psi, size_t m, uint64_t seed, size_t cnt, uint64_t *out)
{
    ensure_init();
    sv_t sv;
    sv.resize(m);
    sv.reset();
    while (sv.used_bits() < m)
        sv.expand();
    val_t *dst = sv.segment_ptr(sv.block_start);
    const size_t n = size_t(1) << m;
    for (size_t i = 0; i < n; i++)
        dst[i] = val_t(psi[2 * i], psi[2 * i + 1]);
    bitmat_t A(m);
    for (size_t q = 0; q < m; q++)
        A.append_col(bitvec_t::e_i(m, q));
    bitvec_t b(m, false);
    measure_cache_t cache(m);
    measure(sv, A, b, cache, cnt, true, seed);
    std::vector<res_t> res(cnt, res_t(m, false));
    get_results(cache, res.data(), cnt);
    for (size_t i = 0; i < cnt; i++)
    {
        uint64_t v = 0;
        for (size_t q = 0; q < m; q++)
            if (res[i].get_bit(q))
                v |= uint64_t(1) << q;
        out[i] = v;
    }
}

} // extern "C"
