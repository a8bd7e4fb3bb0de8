// poly.h -- a diagonal window as a phase polynomial over the address bits.
//
// When every selector of a two-selector gate has at most one alpha bit and
// every one-selector gate's selector at most two,
//     theta(x) = sum_g theta_g [S0_g(x)] [S1_g(x)]
//              = K + sum_j L_j x_j + sum_{j<k} Q_jk x_j x_k ,
// using [x_j ^ beta] = beta + (1 - 2 beta) x_j and
// [x_u ^ x_v] = x_u + x_v - 2 x_u x_v.  Both the diagonal sweep and
// the diagonal epilogue of the fused tile pass evaluate theta through this
// form, split into per-lane / per-row / per-block factors.
#pragma once

#include "dev.h"

#include <vector>

namespace zxa
{

struct Poly
{
    unsigned n = 0;
    double K = 0.0;
    std::vector<double> L; // n
    std::vector<double> Q; // n*n, upper triangle (j < k) used
    double q(int j, int k) const
    {
        return j < k ? Q[size_t(j) * n + k] : Q[size_t(k) * n + j];
    }
};

// True when the polynomial form exists: the selectors have at most one
// alpha bit (two for a gate with one selector), all inside the n address
// bits.
bool window_is_polynomial(unsigned n, const DiagGate *gates, size_t ngates);
// Builds the polynomial; requires window_is_polynomial.
void window_polynomial(unsigned n, const DiagGate *gates, size_t ngates, Poly &out);

// Per-block coefficients for a split of the address bits into nco "inner"
// bits (positions inner[0..nco)) and nh "outer" bits (positions outer[0..nh),
// ascending; block index bit i = address bit outer[i]):
//     out[b*16 + j] = sum_i Q(inner[j], outer[i]) b_i          (j < nco)
//     out[b*16 + nco] = K + sum_i b_i (L(outer[i]) + sum_{l>i} Q(outer[i], outer[l]) b_l)
// Launched on ctx.stream; the coefficient tables are staged through the ring
// (the caller calls ctx.ring.done() after its consumers).
int launch_block_coef(Ctx &ctx, const Poly &p, const int *inner, int nco, const int *outer, int nh, double *out);

} // namespace zxa
