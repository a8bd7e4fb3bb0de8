/*
 * zo_selector.c -- selector evaluation in a translation unit of its own, for
 * the CPU-baseline build of the oracle (liboracle_base.so) only.
 *
 * TEST INFRASTRUCTURE, NOT THE PRODUCT.  The reference defines
 * selector_t::eval / eval_u64 out of line in selector.cpp:26-38 and builds
 * without interprocedural optimisation (CMakeLists.txt:483), so its diagonal,
 * X, CX and CP kernels pay a call per evaluation.  Keeping the call here makes
 * the baseline's timing that of the reference's code structure; the result is
 * the same S(x) = parity(x & alpha) xor beta.
 */
#include <stdint.h>

int zo_sel_eval_call(uint64_t packed, uint64_t x);

int zo_sel_eval_call(uint64_t packed, uint64_t x)
{
    const uint64_t alpha = packed & ((UINT64_C(1) << 63) - 1);
    const int beta = (int)((packed >> 63) & 1u);
    const int parity = __builtin_popcountll(x & alpha) & 1;
    return parity != beta;
}
