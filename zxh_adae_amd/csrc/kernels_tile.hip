// kernels_tile.hip -- tile passes: fused single-qubit batches, optionally
// followed by a diagonal window, in one read and one write of the state.
//
// The reference applies a fused U3 cluster as one full-state sweep per
// descriptor (u3_block_batch_kernel_cuda, src/zxh/src/kernel/kernel_cuda.cu:390-423;
// CPU kernel_general.cpp:173-202) and a diagonal window as a further sweep
// (diag_pending_flush, diag_cuda.cu:260-281).  Here a pass works on tiles.
//
// A tile is 4096 16-byte units (complex128: one amplitude; complex64: two
// amplitudes adjacent in address bit 0) whose unit addresses differ in 12
// chosen "unit bits": the seg lowest unit-address bits (the tile geometry,
// seg = 3..6: 128 B to 1 KiB contiguous per wave access), plus up to 12 - seg
// target bits; fewer contiguous bits leave room for more targets per pass.
// A complex64 tile therefore spans address bits 0..seg and 12 - seg more.  Each
// workgroup streams through tiles; each tile is loaded once, every gate whose
// target lies in the tile is applied, the diagonal epilogue multiplies each
// amplitude by its phase, and the tile is written back.
//
// Inside a tile each of the 256 threads holds 16 units.  Gates act on
// register pairs; tile bits that are thread bits in the load layout are
// reached by transposing the tile through LDS (XOR-swizzled: conflict-free
// for the three layouts used) into a layout whose register bits they are.
//
// Streaming: the grid is persistent (two workgroups per CU, LDS-bound); a
// workgroup visits tiles blockIdx, blockIdx + grid, ... so the tiles in
// flight at any moment are consecutive, and it issues the loads of its next
// tile before working on the current one, which keeps HBM busy while the
// gates, transposes and phases run.  Loads and stores are nontemporal.
#include "dev.h"
#include "poly.h"

#include <algorithm>
#include <climits>
#include <cmath>
#include <complex>
#include <cstring>
#include <type_traits>
#include <vector>

namespace zxa
{
namespace
{

constexpr int kThreads = 256;
constexpr int kUnitBits = 12;
constexpr int kUnits = 1 << kUnitBits;
constexpr int kRegs = 16;
constexpr int kIntra = kUnitBits; // gate slot of the complex64 in-unit bit

typedef double dvec2 __attribute__((ext_vector_type(2)));
typedef float fvec4 __attribute__((ext_vector_type(4)));
typedef float fvec2 __attribute__((ext_vector_type(2)));

// Read-only tables are read through the constant address space: the loads
// become scalar (s_load, counted by lgkmcnt) instead of uniform vector loads,
// which would share vmcnt with the in-flight tile loads and force waits on
// them.
#define ZXA_CONST const __attribute__((address_space(4)))

template <bool kC64>
struct UnitT;
template <>
struct UnitT<false>
{
    using T = dvec2;
};
template <>
struct UnitT<true>
{
    using T = fvec4;
};

struct TGate
{
    double u[8];    // m00 m01 m10 m11 as (re, im)
    float uf[8];    // the same in float for complex64 states
    uint64_t amask; // selector bits besides the target, unit-address space
    uint32_t a0;    // complex64: selector bit 0 (differs between a unit's two amplitudes)
    uint32_t beta;
    uint32_t rpar;  // bit r: parity of amask over register r's offset in the gate's layout
    uint32_t fan;   // 1: the pair's second element takes a phase first (see TFan)
};

// A gate's fan (complex128 passes): before the gate, every amplitude whose
// target bit is one is multiplied by exp(i phi(x)), phi linear in the other
// address bits.  exp(i phi) = F(tile) * nib[t & 15] * nib[16 + (t >> 4)] *
// reg[r]: F per tile (fan_phase_kernel), nib over the low and high four
// thread bits and reg over the register offsets of the gate's layout.  The
// kernel copies nib to LDS (a thread's entries differ from its neighbours').
struct TFan
{
    double nib[32][2];
    double reg[kRegs][2];
};
constexpr int kFanStride = 2 * (kUnitBits + 1); // doubles of per-tile factors per tile

// Kernel arguments of a pass.  Unit bits 0..seg-1 are unit-address bits
// 0..seg-1 (seg = 3..6, the tile geometry); the host assigns the others
// freely (targets first into the register bits of layout 0, then of layout
// 1, then unit bit 3), so ub is not sorted; sorted holds the same positions
// ascending.
struct TPass
{
    uint32_t ub[kUnitBits];     // unit-address bit of each unit bit (its pivot)
    uint32_t sorted[kUnitBits]; // the same, ascending
    uint64_t vec[kUnitBits];    // unit-address offset of each unit bit
    uint64_t ntiles;
    uint64_t tmask; // unit-address mask of the 12 pivots
    uint64_t gstep; // the grid size deposited into the non-tile bits
    const double *fphase; // fan passes: per-tile factors, kFanStride doubles per tile
};

// Per-pass tables in device memory.  The kernel re-reads them through the
// scalar cache at every tile instead of holding them in SGPRs across the
// tile loop (12 matrices and 16 offsets would not fit).  Every slot of a
// layout the pass uses holds a matrix (the identity where there is no gate):
// the kernel applies all of them without branches.
// A diagonal window as a table (any selectors whose linear forms span at
// most kWinBits dimensions): with a_0 .. a_{r-1} a basis of the window's
// selector forms over the amplitude address x, y(x) = sum_k (a_k . x) 2^k
// and the window's phase is a function of y alone, exp(i theta) = T[y].  y
// is linear over XOR, so an element's y is y(tile base) ^ y(thread offset)
// ^ y(register offset) (^ yin for a complex64 unit's second amplitude).
constexpr int kWinBits = kWinRankMax;
struct TWin
{
    uint64_t a[kWinBits]; // basis forms over unit addresses (zero past the rank)
    uint32_t yr[kRegs];   // y of each layout-0 register offset
    uint32_t yin;         // complex64: y of the in-unit address bit
    uint32_t rank;
};

struct TTab
{
    uint64_t roff[kRegs]; // layout-0 register offsets (unit addresses)
    TGate g[kUnitBits + 1];
    TFan fan[kUnitBits + 1];
    TWin win;             // table epilogue (tile passes launched with kTab)
};

// Register bits of the three layouts (unit bits); the thread bits are the
// remaining eight in ascending order.  Layout 0 is the load/store layout: its
// thread bits start with unit bits 0..5, of which the lowest seg are the
// contiguous ones.
__host__ __device__ constexpr int reg_bit(int L, int i)
{
    return L == 0 ? 8 + i : (L == 1 ? 4 + i : i);
}
__host__ __device__ constexpr int thr_bit(int L, int j)
{
    return L == 0 ? j : (L == 1 ? (j < 4 ? j : j + 4) : j + 4);
}
int layout_of_unit_bit(int k)
{
    return k >= 8 ? 0 : (k >= 4 ? 1 : 2);
}

template <int L>
__device__ __forceinline__ int thr_tau(int t)
{
    int tau = 0;
#pragma unroll
    for (int j = 0; j < 8; j++)
        tau |= ((t >> j) & 1) << thr_bit(L, j);
    return tau;
}
template <int L>
__device__ __forceinline__ constexpr int reg_tau(int r)
{
    int tau = 0;
    for (int i = 0; i < 4; i++)
        tau |= ((r >> i) & 1) << reg_bit(L, i);
    return tau;
}
template <int L>
__device__ __forceinline__ uint64_t thr_off(int t, const TPass &P)
{
    uint64_t o = 0;
#pragma unroll
    for (int j = 0; j < 8; j++)
        o ^= ((t >> j) & 1) ? P.vec[thr_bit(L, j)] : 0;
    return o;
}
// Unit address of the first unit of a tile: the tile index deposited into
// the non-tile bits.
__host__ __device__ __forceinline__ uint64_t deposit(uint64_t tile, const uint32_t *ub)
{
    for (int k = 0; k < kUnitBits; k++)
    {
        const unsigned p = ub[k];
        tile = ((tile >> p) << (p + 1)) | (tile & ((uint64_t(1) << p) - 1));
    }
    return tile;
}

// deposit(tile + grid) from deposit(tile): an addition inside the non-tile
// bits, the tile bits set to one so that carries run through them.
__device__ __forceinline__ uint64_t next_base(uint64_t base, const TPass &P)
{
    return ((base | P.tmask) + P.gstep) & ~P.tmask;
}

// LDS swizzle: within every group of 16 lanes the 16 units of an access land
// in 16 distinct 16-byte bank groups, for all three layouts.  It is linear
// over XOR: swz(a | b) = swz(a) ^ swz(b) for disjoint a, b.
__host__ __device__ constexpr int swz(int tau)
{
    return tau ^ ((tau >> 4) & 15);
}

__device__ __forceinline__ uint32_t par(uint64_t x)
{
    return __popcll(x) & 1u;
}

template <typename R>
struct Pair
{
    R ax, ay, bx, by;
};

// (a, b) -> (m00 a + m01 b, m10 a + m11 b); with flip the roles of a and b
// swap (a is then the odd side of the pair).  m: the matrix, loaded once
// per gate.
template <typename R>
__device__ __forceinline__ Pair<R> pair2(R ax, R ay, R bx, R by, const R (&m)[8], bool flip)
{
    const R m00x = flip ? m[6] : m[0], m00y = flip ? m[7] : m[1];
    const R m01x = flip ? m[4] : m[2], m01y = flip ? m[5] : m[3];
    const R m10x = flip ? m[2] : m[4], m10y = flip ? m[3] : m[5];
    const R m11x = flip ? m[0] : m[6], m11y = flip ? m[1] : m[7];
    Pair<R> o;
    o.ax = m00x * ax - m00y * ay + m01x * bx - m01y * by;
    o.ay = m00x * ay + m00y * ax + m01x * by + m01y * bx;
    o.bx = m10x * ax - m10y * ay + m11x * bx - m11y * by;
    o.by = m10x * ay + m10y * ax + m11x * by + m11y * bx;
    return o;
}

template <bool kC64>
using Real = typename std::conditional<kC64, float, double>::type;

// Complex64 gates on the matrix cores.  A 2x2 complex gate on the pair
// (a, b) is the real 4x4 product (a', b') = M4 (ax, ay, bx, by) with rows
//   [m00x -m00y m01x -m01y], [m00y m00x m01y m01x],
//   [m10x -m10y m11x -m11y], [m10y m10x m11y m11x].
// v_mfma_f32_4x4x1f32 (16 blocks of four lanes) gives lane 4b+q, register i,
// the sum over k of A_k(lane 4b+i) * B_k(lane 4b+q): when lane q of every
// block supplies row q of M4 as A (the same gate in all lanes) and the
// components of its own pair as B, four steps (k = 0..3) leave each lane
// with its own pair transformed.  The matrix cores take the gate arithmetic
// off the VALU, which the complex64 passes otherwise saturate; the whole
// wave must be active (every caller is uniform over the workgroup).
typedef float v4f __attribute__((ext_vector_type(4)));
struct MRow
{
    float a0, a1, a2, a3;
};

__device__ __forceinline__ MRow mfma_row(const float (&m)[8])
{
    const int q = threadIdx.x & 3;
    const bool lo = q < 2, im = (q & 1) != 0;
    const float c0 = lo ? m[0] : m[4], s0 = lo ? m[1] : m[5];
    const float c1 = lo ? m[2] : m[6], s1 = lo ? m[3] : m[7];
    MRow r;
    r.a0 = im ? s0 : c0;
    r.a1 = im ? c0 : -s0;
    r.a2 = im ? s1 : c1;
    r.a3 = im ? c1 : -s1;
    return r;
}

__device__ __forceinline__ v4f mfma_pair(const MRow &r, float ax, float ay, float bx, float by)
{
    v4f acc = {0.f, 0.f, 0.f, 0.f};
    acc = __builtin_amdgcn_mfma_f32_4x4x1f32(r.a0, ax, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_4x4x1f32(r.a1, ay, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_4x4x1f32(r.a2, bx, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_4x4x1f32(r.a3, by, acc, 0, 0, 0);
    return acc;
}

// Units a, b of a complex64 register pair: both amplitude halves.
__device__ __forceinline__ void mfma_units(const MRow &r, fvec4 &a, fvec4 &b)
{
    const v4f lo = mfma_pair(r, a.x, a.y, b.x, b.y);
    const v4f hi = mfma_pair(r, a.z, a.w, b.z, b.w);
    a = fvec4{lo.x, lo.y, hi.x, hi.y};
    b = fvec4{lo.z, lo.w, hi.z, hi.w};
}

// The in-unit gate of a complex64 unit (its two amplitudes are the pair).
__device__ __forceinline__ void mfma_intra(const MRow &r, fvec4 &v)
{
    const v4f o = mfma_pair(r, v.x, v.y, v.z, v.w);
    v = fvec4{o.x, o.y, o.z, o.w};
}

template <bool kC64>
__device__ __forceinline__ void load_matrix(Real<kC64> (&m)[8], ZXA_CONST TGate &g)
{
#pragma unroll
    for (int i = 0; i < 8; i++)
    {
        if constexpr (kC64)
            m[i] = g.uf[i];
        else
            m[i] = g.u[i];
    }
}

template <bool kC64, bool kFlip>
__device__ __forceinline__ void unit_pair(typename UnitT<kC64>::T &a, typename UnitT<kC64>::T &b,
                                          const Real<kC64> (&m)[8], bool flip, bool a0)
{
    if constexpr (kC64)
    {
        const Pair<float> lo = pair2<float>(a.x, a.y, b.x, b.y, m, kFlip && flip);
        const Pair<float> hi = pair2<float>(a.z, a.w, b.z, b.w, m, kFlip && (flip ^ a0));
        a = fvec4{lo.ax, lo.ay, hi.ax, hi.ay};
        b = fvec4{lo.bx, lo.by, hi.bx, hi.by};
    }
    else
    {
        const Pair<double> o = pair2<double>(a.x, a.y, b.x, b.y, m, kFlip && flip);
        a = dvec2{o.ax, o.ay};
        b = dvec2{o.bx, o.by};
    }
}

// (a, b) -> M (a, b) with the matrix ml for the first amplitude of each unit
// and mh for the second (complex64; complex128 units hold one amplitude).
template <bool kC64>
__device__ __forceinline__ void unit_pair2(typename UnitT<kC64>::T &a, typename UnitT<kC64>::T &b,
                                           const Real<kC64> (&ml)[8], const Real<kC64> (&mh)[8])
{
    if constexpr (kC64)
    {
        const Pair<float> lo = pair2<float>(a.x, a.y, b.x, b.y, ml, false);
        const Pair<float> hi = pair2<float>(a.z, a.w, b.z, b.w, mh, false);
        a = fvec4{lo.ax, lo.ay, hi.ax, hi.ay};
        b = fvec4{lo.bx, lo.by, hi.bx, hi.by};
    }
    else
    {
        const Pair<double> o = pair2<double>(a.x, a.y, b.x, b.y, ml, false);
        a = dvec2{o.ax, o.ay};
        b = dvec2{o.bx, o.by};
    }
}

// Gate g on register bit i of the thread's units v (base ^ toff: the unit
// address of register 0; g.rpar: the selector's parity over the register
// offsets of the layout in use).
template <bool kC64, bool kFlip, bool kMfma = true>
__device__ __forceinline__ void gate_on_reg_bit(typename UnitT<kC64>::T (&v)[kRegs], int i, ZXA_CONST TGate &g,
                                                uint64_t base, uint64_t toff)
{
    Real<kC64> m[8];
    load_matrix<kC64>(m, g);
    const uint64_t amask = kFlip ? g.amask : 0;
    const uint32_t pt = kFlip ? (par((base ^ toff) & amask) ^ g.beta) : 0u;
    const bool a0 = kFlip && g.a0 != 0;
    if (kFlip && g.rpar == 0u)
    {
        // the side does not change across the thread's registers (rpar is
        // uniform): select the matrix once, for both unit halves (the
        // operands are laundered: the compiler would otherwise turn the
        // select into a run-time index into m)
        Real<kC64> ml[8], mh[8];
#pragma unroll
        for (int e = 0; e < 8; e++)
        {
            Real<kC64> keep = m[e], swap = m[e ^ 6];
            asm volatile("" : "+v"(keep), "+v"(swap));
            ml[e] = pt ? swap : keep;
            mh[e] = (pt ^ uint32_t(a0)) ? swap : keep;
        }
#pragma unroll
        for (int r = 0; r < kRegs; r++)
        {
            if ((r >> i) & 1)
                continue;
            unit_pair2<kC64>(v[r], v[r | (1 << i)], ml, mh);
        }
        return;
    }
    if constexpr (kC64 && !kFlip && kMfma)
    {
        const MRow mr = mfma_row(m);
#pragma unroll
        for (int r = 0; r < kRegs; r++)
        {
            if ((r >> i) & 1)
                continue;
            mfma_units(mr, v[r], v[r | (1 << i)]);
        }
        return;
    }
#pragma unroll
    for (int r = 0; r < kRegs; r++)
    {
        if ((r >> i) & 1)
            continue;
        const bool flip = kFlip && ((pt ^ (g.rpar >> r)) & 1u) != 0u;
        unit_pair<kC64, kFlip>(v[r], v[r | (1 << i)], m, flip, a0);
    }
}

// Gates on the register bits of layout L (and, in layout 0, the complex64
// in-unit gate).  base ^ toff is the unit address of the thread's register 0.
template <bool kC64, bool kFlip, int L, bool kFan = false, bool kMfma = true>
__device__ __forceinline__ void apply_layout(typename UnitT<kC64>::T (&v)[kRegs], const TPass &P, ZXA_CONST TTab *G,
                                             uint64_t base, uint64_t toff, int t = 0, ZXA_CONST double *fp = nullptr,
                                             const dvec2 *fl = nullptr)
{
#pragma unroll
    for (int i = 0; i < 4; i++)
    {
        const int k = reg_bit(L, i);
        ZXA_CONST TGate &g = G->g[k];
        if constexpr (kFan && !kC64)
        {
            if (g.fan != 0u)
            {
                // the gate's fan: exp(i phi) on the second element of each pair
                ZXA_CONST TFan &f = G->fan[k];
                const dvec2 lo = fl[k * 32 + (t & 15)], hi = fl[k * 32 + 16 + (t >> 4)];
                const double tx = lo.x * hi.x - lo.y * hi.y, ty = lo.x * hi.y + lo.y * hi.x;
                const double ux = fp[2 * k], uy = fp[2 * k + 1];
                const double fx = ux * tx - uy * ty, fy = ux * ty + uy * tx;
#pragma unroll
                for (int r = 0; r < kRegs; r++)
                {
                    if ((r >> i) & 1)
                        continue;
                    const double px = fx * f.reg[r][0] - fy * f.reg[r][1];
                    const double py = fx * f.reg[r][1] + fy * f.reg[r][0];
                    dvec2 &b = v[r | (1 << i)];
                    const double bx = b.x * px - b.y * py;
                    const double by = b.x * py + b.y * px;
                    b = dvec2{bx, by};
                }
            }
        }
        gate_on_reg_bit<kC64, kFlip, kMfma>(v, i, g, base, toff);
    }
    if constexpr (kC64 && L == 0)
    {
        ZXA_CONST TGate &g = G->g[kIntra];
        float m[8];
        load_matrix<true>(m, g);
        if constexpr (!kFlip && kMfma)
        {
            const MRow mr = mfma_row(m);
#pragma unroll
            for (int r = 0; r < kRegs; r++)
                mfma_intra(mr, v[r]);
            return;
        }
        const uint64_t amask = kFlip ? g.amask : 0;
        const uint32_t pt = kFlip ? (par((base ^ toff) & amask) ^ g.beta) : 0u;
#pragma unroll
        for (int r = 0; r < kRegs; r++)
        {
            const bool flip = kFlip && ((pt ^ (g.rpar >> r)) & 1u) != 0u;
            const Pair<float> o = pair2<float>(v[r].x, v[r].y, v[r].z, v[r].w, m, flip);
            v[r] = fvec4{o.ax, o.ay, o.bx, o.by};
        }
    }
}

template <bool kC64, int LW, int LR>
__device__ __forceinline__ void transpose(typename UnitT<kC64>::T (&v)[kRegs], typename UnitT<kC64>::T *lds, int t)
{
    // The LDS addresses are recomputed at every transpose (one XOR each)
    // rather than hoisted out of the tile loop, which would pin 32 VGPRs.
    int tw = swz(thr_tau<LW>(t));
    int tr = swz(thr_tau<LR>(t));
    asm volatile("" : "+v"(tw), "+v"(tr));
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kRegs; r++)
        lds[tw ^ swz(reg_tau<LW>(r))] = v[r];
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kRegs; r++)
        v[r] = lds[tr ^ swz(reg_tau<LR>(r))];
}

struct cd
{
    double x, y;
};
__device__ __forceinline__ cd cmul(cd a, cd b)
{
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}

// Diagonal epilogue (layout 0).  For the amplitude of thread t at register
// element rho (complex128: rho = r; complex64: rho = 2 r + in-unit bit) the
// window's phase is
//     Phi = ST(t) * prod_{j in t} E_j * D[rho] * prod_{b in rho} G_b(t)
// with ST, G tile-independent (per-thread table) and E, D per tile (written
// by tile_phase_kernel).  rho is visited in Gray-code order so that the last
// product costs one complex multiplication per amplitude.
template <bool kC64, int kThr = 8>
struct Epi
{
    static constexpr int kNB = kC64 ? 5 : 4;         // register element bits
    static constexpr int kNR = 1 << kNB;             // register elements
    static constexpr int kStride = 2 * (kThr + kNR); // doubles per tile in tphase
    static constexpr int kThrStride = 2 * (1 + kNB); // doubles per thread in ttab
    cd st;
    cd g[kNB];

    __device__ __forceinline__ void init(const double *__restrict__ ttab, int t)
    {
        const double *a = ttab + t * kThrStride;
        st = {a[0], a[1]};
#pragma unroll
        for (int b = 0; b < kNB; b++)
            g[b] = {a[2 + 2 * b], a[3 + 2 * b]};
    }

    // Multiplies element rho of v by ph.
    static __device__ __forceinline__ void mul(typename UnitT<kC64>::T (&v)[kRegs], int rho, cd ph)
    {
        if constexpr (kC64)
        {
            const float px = static_cast<float>(ph.x), py = static_cast<float>(ph.y);
            fvec4 &u = v[rho >> 1];
            if (rho & 1)
            {
                const float zx = u.z * px - u.w * py, zy = u.z * py + u.w * px;
                u.z = zx;
                u.w = zy;
            }
            else
            {
                const float xx = u.x * px - u.y * py, xy = u.x * py + u.y * px;
                u.x = xx;
                u.y = xy;
            }
        }
        else
        {
            dvec2 &u = v[rho];
            const double nx = u.x * ph.x - u.y * ph.y, ny = u.x * ph.y + u.y * ph.x;
            u.x = nx;
            u.y = ny;
        }
    }

    // The window's phases on v and, when kTwo, the same phases on w.
    template <bool kTwo>
    __device__ __forceinline__ void apply_n(typename UnitT<kC64>::T (&v)[kRegs], typename UnitT<kC64>::T (&w2)[kRegs],
                                            ZXA_CONST double *tp, int t) const
    {
        cd w = st;
#pragma unroll
        for (int j = 0; j < kThr; j++)
        {
            const cd e = {tp[2 * j], tp[2 * j + 1]};
            const cd we = cmul(w, e);
            const bool on = (t >> j) & 1;
            w.x = on ? we.x : w.x;
            w.y = on ? we.y : w.y;
        }
        ZXA_CONST double *D = tp + 2 * kThr;
#pragma unroll
        for (int k = 0; k < kNR; k++)
        {
            const int rho = k ^ (k >> 1);
            if (k > 0)
            {
                const int b = __builtin_ctz(k);
                const cd gb = ((rho >> b) & 1) ? g[b] : cd{g[b].x, -g[b].y};
                w = cmul(w, gb);
            }
            const cd ph = cmul(w, {D[2 * rho], D[2 * rho + 1]});
            mul(v, rho, ph);
            if constexpr (kTwo)
                mul(w2, rho, ph);
        }
    }

    __device__ __forceinline__ void apply(typename UnitT<kC64>::T (&v)[kRegs], ZXA_CONST double *tp, int t) const
    {
        apply_n<false>(v, v, tp, t);
    }
};

template <bool kC64>
__device__ __forceinline__ void load_tile(typename UnitT<kC64>::T (&v)[kRegs], const typename UnitT<kC64>::T *p,
                                          uint64_t at, ZXA_CONST TTab *G)
{
#pragma unroll
    for (int r = 0; r < kRegs; r++)
        v[r] = __builtin_nontemporal_load(p + (at ^ G->roff[r]));
}

// The store recomputes its addresses (at is laundered): keeping the loads'
// 64-bit addresses alive across a tile's work costs two registers per unit.
template <bool kC64>
__device__ __forceinline__ void store_tile(const typename UnitT<kC64>::T (&v)[kRegs], typename UnitT<kC64>::T *p,
                                           uint64_t at, ZXA_CONST TTab *G)
{
    asm volatile("" : "+v"(at));
#pragma unroll
    for (int r = 0; r < kRegs; r++)
        __builtin_nontemporal_store(v[r], p + (at ^ G->roff[r]));
}

// y of a unit address (see TWin).
__device__ __forceinline__ uint32_t win_y(ZXA_CONST TWin &W, uint64_t u)
{
    uint32_t y = 0;
#pragma unroll
    for (int k = 0; k < kWinBits; k++)
        y |= uint32_t(__popcll(u & W.a[k]) & 1) << k;
    return y;
}

// Table entries as the kernels keep them in LDS: complex128 in double,
// complex64 in float (the precision of the amplitudes they multiply).
template <bool kC64>
using WinT = std::conditional_t<kC64, fvec2, dvec2>;

// Copies the window's table (2^rank entries of (cos, sin) in double) into
// LDS; the caller synchronizes before the first use.
template <bool kC64>
__device__ __forceinline__ void win_load(WinT<kC64> *wt, const double *__restrict__ tab, ZXA_CONST TWin &W, int t)
{
    const int nent = 1 << W.rank;
    for (int e = t; e < nent; e += kThreads)
    {
        if constexpr (kC64)
            wt[e] = fvec2{float(tab[2 * e]), float(tab[2 * e + 1])};
        else
            wt[e] = dvec2{tab[2 * e], tab[2 * e + 1]};
    }
}

// The window's phases on a tile in layout 0; y0 = y(base) ^ y(thread).
template <bool kC64>
__device__ __forceinline__ void win_apply(typename UnitT<kC64>::T (&v)[kRegs], const WinT<kC64> *wt,
                                          ZXA_CONST TWin &W, uint32_t y0)
{
#pragma unroll
    for (int r = 0; r < kRegs; r++)
    {
        const uint32_t y = y0 ^ W.yr[r];
        if constexpr (kC64)
        {
            const fvec2 p0 = wt[y], p1 = wt[y ^ W.yin];
            fvec4 &u = v[r];
            const float ax = u.x * p0.x - u.y * p0.y, ay = u.x * p0.y + u.y * p0.x;
            const float bx = u.z * p1.x - u.w * p1.y, by = u.z * p1.y + u.w * p1.x;
            u.x = ax;
            u.y = ay;
            u.z = bx;
            u.w = by;
        }
        else
        {
            const dvec2 p = wt[y];
            dvec2 &u = v[r];
            const double nx = u.x * p.x - u.y * p.y, ny = u.x * p.y + u.y * p.x;
            u.x = nx;
            u.y = ny;
        }
    }
}

// Gates, transposes and the diagonal epilogue of one tile held in v
// (kDiag: the phase-polynomial epilogue; kTab: the window table wt, with
// yt the thread offset's y).
template <bool kC64, bool kL1, bool kL2, bool kDiag, bool kFlip, bool kFan = false, bool kTab = false>
__device__ __forceinline__ void tile_work(typename UnitT<kC64>::T (&v)[kRegs], typename UnitT<kC64>::T *lds,
                                          const TPass &P, ZXA_CONST TTab *G, const double *tphase,
                                          const Epi<kC64> &epi, int t, uint64_t toff, uint64_t tile, uint64_t base,
                                          const dvec2 *fl = nullptr, const WinT<kC64> *wt = nullptr,
                                          uint32_t yt = 0)
{
    ZXA_CONST double *fp = nullptr;
    if constexpr (kFan)
        fp = (ZXA_CONST double *)(P.fphase + tile * kFanStride);
    apply_layout<kC64, kFlip, 0, kFan>(v, P, G, base, toff, t, fp, fl);
    if constexpr (kL1)
    {
        transpose<kC64, 0, 1>(v, lds, t);
        apply_layout<kC64, kFlip, 1, kFan>(v, P, G, base, thr_off<1>(t, P), t, fp, fl);
        if constexpr (kL2)
        {
            transpose<kC64, 1, 2>(v, lds, t);
            apply_layout<kC64, kFlip, 2, kFan>(v, P, G, base, thr_off<2>(t, P), t, fp, fl);
            transpose<kC64, 2, 0>(v, lds, t);
        }
        else
        {
            transpose<kC64, 1, 0>(v, lds, t);
        }
    }
    else if constexpr (kL2)
    {
        transpose<kC64, 0, 2>(v, lds, t);
        apply_layout<kC64, kFlip, 2, kFan>(v, P, G, base, thr_off<2>(t, P), t, fp, fl);
        transpose<kC64, 2, 0>(v, lds, t);
    }
    if constexpr (kDiag)
        epi.apply(v, (ZXA_CONST double *)(tphase + tile * Epi<kC64>::kStride), t);
    if constexpr (kTab)
        win_apply<kC64>(v, wt, G->win, win_y(G->win, base) ^ yt);
}

// The tables are re-read at every tile (see TTab): the pointer is laundered
// so that the loads are not hoisted out of the tile loop.
__device__ __forceinline__ ZXA_CONST TTab *tables(const TTab *tabp)
{
    asm volatile("" : "+s"(tabp));
    return (ZXA_CONST TTab *)tabp;
}

// tabp: register offsets and gate matrices; ttab: per-thread ST and G
// (kDiag), or the window table (kTab); tphase: per-tile E and D (kDiag only).
//
// Software pipeline over the workgroup's tiles t0, t1 = t0 + grid, ... with
// two register buffers: while tile k is worked on, tile k + 1 is in flight,
// and the store of tile k - 1 is issued after the loads of tile k + 1, so
// that waiting for a tile's loads never waits for the previous stores.  The
// prefetch is unconditional (past the last tile it re-reads the current one):
// with no branch around it the compiler's wait counts stay exact.
template <bool kC64, bool kL1, bool kL2, bool kDiag, bool kFlip, bool kFan = false, bool kTab = false>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(2))) void tile_kernel(
    void *__restrict__ state, const TPass P, const TTab *__restrict__ tabp, const double *__restrict__ ttab,
    const double *__restrict__ tphase)
{
    static_assert(!(kTab && (kDiag || kFan)), "one epilogue kind");
    using T = typename UnitT<kC64>::T;
    __shared__ T lds[(kL1 || kL2) ? kUnits : 1];
    // fan passes: every slot's thread-nibble factors (TFan::nib)
    __shared__ dvec2 fl[kFan ? (kUnitBits + 1) * 32 : 1];
    __shared__ WinT<kC64> wt[kTab ? (1 << kWinBits) : 1];
    const int t = threadIdx.x;
    T *p = static_cast<T *>(state);
    const uint64_t toff = thr_off<0>(t, P);
    Epi<kC64> epi;
    if constexpr (kDiag)
        epi.init(ttab, t);
    const uint64_t grid = gridDim.x;
    if constexpr (kFan)
    {
        for (int e = t; e < (kUnitBits + 1) * 32; e += kThreads)
            fl[e] = dvec2{tabp->fan[e >> 5].nib[e & 31][0], tabp->fan[e >> 5].nib[e & 31][1]};
        __syncthreads();
    }
    uint32_t yt = 0;
    if constexpr (kTab)
    {
        ZXA_CONST TTab *G0 = tables(tabp);
        win_load<kC64>(wt, ttab, G0->win, t);
        yt = win_y(G0->win, toff);
        __syncthreads();
    }

    T a[kRegs], b[kRegs];
    uint64_t t0 = blockIdx.x;
    uint64_t b0 = deposit(t0, P.sorted);
    ZXA_CONST TTab *G = tables(tabp);
    load_tile<kC64>(a, p, b0 ^ toff, G);
    uint64_t t1 = t0 + grid;
    bool m1 = t1 < P.ntiles;
    uint64_t b1 = next_base(b0, P);
    load_tile<kC64>(b, p, (m1 ? b1 : b0) ^ toff, G);
    tile_work<kC64, kL1, kL2, kDiag, kFlip, kFan, kTab>(a, lds, P, G, tphase, epi, t, toff, t0, b0, fl, wt, yt);
    for (;;)
    {
        G = tables(tabp);
        store_tile<kC64>(a, p, b0 ^ toff, G);
        if (!m1)
            break;
        const uint64_t t2 = t1 + grid;
        const bool m2 = t2 < P.ntiles;
        const uint64_t b2 = next_base(b1, P);
        load_tile<kC64>(a, p, (m2 ? b2 : b1) ^ toff, G);
        tile_work<kC64, kL1, kL2, kDiag, kFlip, kFan, kTab>(b, lds, P, G, tphase, epi, t, toff, t1, b1, fl, wt, yt);

        G = tables(tabp);
        store_tile<kC64>(b, p, b1 ^ toff, G);
        if (!m2)
            break;
        const uint64_t t3 = t2 + grid;
        const bool m3 = t3 < P.ntiles;
        const uint64_t b3 = next_base(b2, P);
        load_tile<kC64>(b, p, (m3 ? b3 : b2) ^ toff, G);
        tile_work<kC64, kL1, kL2, kDiag, kFlip, kFan, kTab>(a, lds, P, G, tphase, epi, t, toff, t2, b2, fl, wt, yt);

        t0 = t2;
        b0 = b2;
        t1 = t3;
        b1 = b3;
        m1 = m3;
    }
}

// The same pipeline with three register buffers: while tile k is worked on,
// tiles k + 1 and k + 2 are in flight (twice the loads in flight per
// workgroup; the workgroup count per CU is bound by LDS).  Past the last tile
// the prefetch re-reads the last valid one.
template <bool kC64, bool kL1, bool kL2, bool kDiag, bool kFlip>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(2))) void tile_kernel3(
    void *__restrict__ state, const TPass P, const TTab *__restrict__ tabp, const double *__restrict__ ttab,
    const double *__restrict__ tphase)
{
    using T = typename UnitT<kC64>::T;
    __shared__ T lds[(kL1 || kL2) ? kUnits : 1];
    const int t = threadIdx.x;
    T *p = static_cast<T *>(state);
    const uint64_t toff = thr_off<0>(t, P);
    Epi<kC64> epi;
    if constexpr (kDiag)
        epi.init(ttab, t);
    const uint64_t grid = gridDim.x;

    T a[kRegs], b[kRegs], c[kRegs];
    uint64_t t0 = blockIdx.x;
    uint64_t b0 = deposit(t0, P.sorted);
    ZXA_CONST TTab *G = tables(tabp);
    load_tile<kC64>(a, p, b0 ^ toff, G);
    uint64_t t1 = t0 + grid;
    bool m1 = t1 < P.ntiles;
    uint64_t b1 = m1 ? next_base(b0, P) : b0;
    load_tile<kC64>(b, p, b1 ^ toff, G);
    uint64_t t2 = t1 + grid;
    bool m2 = m1 && t2 < P.ntiles;
    uint64_t b2 = m2 ? next_base(b1, P) : b1;
    load_tile<kC64>(c, p, b2 ^ toff, G);
    tile_work<kC64, kL1, kL2, kDiag, kFlip>(a, lds, P, G, tphase, epi, t, toff, t0, b0);
    for (;;)
    {
        G = tables(tabp);
        store_tile<kC64>(a, p, b0 ^ toff, G);
        if (!m1)
            break;
        const uint64_t t3 = t2 + grid;
        const bool m3 = m2 && t3 < P.ntiles;
        const uint64_t b3 = m3 ? next_base(b2, P) : b2;
        load_tile<kC64>(a, p, b3 ^ toff, G);
        tile_work<kC64, kL1, kL2, kDiag, kFlip>(b, lds, P, G, tphase, epi, t, toff, t1, b1);

        G = tables(tabp);
        store_tile<kC64>(b, p, b1 ^ toff, G);
        if (!m2)
            break;
        const uint64_t t4 = t3 + grid;
        const bool m4 = m3 && t4 < P.ntiles;
        const uint64_t b4 = m4 ? next_base(b3, P) : b3;
        load_tile<kC64>(b, p, b4 ^ toff, G);
        tile_work<kC64, kL1, kL2, kDiag, kFlip>(c, lds, P, G, tphase, epi, t, toff, t2, b2);

        G = tables(tabp);
        store_tile<kC64>(c, p, b2 ^ toff, G);
        if (!m3)
            break;
        const uint64_t t5 = t4 + grid;
        const bool m5 = m4 && t5 < P.ntiles;
        const uint64_t b5 = m5 ? next_base(b4, P) : b4;
        load_tile<kC64>(c, p, b5 ^ toff, G);
        tile_work<kC64, kL1, kL2, kDiag, kFlip>(a, lds, P, G, tphase, epi, t, toff, t3, b3);

        t0 = t3;
        b0 = b3;
        t1 = t4;
        b1 = b4;
        m1 = m4;
        t2 = t5;
        b2 = b5;
        m2 = m5;
    }
}

// Exchanges lane bit 4 (k32 false: v_permlane16_swap, the odd 16-lane rows of
// a with the even rows of b) or lane bit 5 (k32 true: v_permlane32_swap, the
// upper 32 lanes of a with the lower 32 of b) with the register bit that
// tells a from b.  Each exchange is its own inverse.
template <bool k32, typename T>
__device__ __forceinline__ void lane_swap(T &a, T &b)
{
    static_assert(sizeof(T) == 16, "16-byte units");
    struct W
    {
        uint32_t w[4];
    };
    W x = __builtin_bit_cast(W, a), y = __builtin_bit_cast(W, b);
#pragma unroll
    for (int i = 0; i < 4; i++)
    {
        if constexpr (k32)
        {
            const auto r = __builtin_amdgcn_permlane32_swap(x.w[i], y.w[i], false, false);
            x.w[i] = r[0];
            y.w[i] = r[1];
        }
        else
        {
            const auto r = __builtin_amdgcn_permlane16_swap(x.w[i], y.w[i], false, false);
            x.w[i] = r[0];
            y.w[i] = r[1];
        }
    }
    a = __builtin_bit_cast(T, x);
    b = __builtin_bit_cast(T, y);
}

// The layout-0 thread whose units lane j of wave w holds in upper_kernel:
// unit bits 0..3 in lane bits 0..3, unit bits 4, 5 in the wave index, unit
// bits 6, 7 in lane bits 4, 5.
__device__ __forceinline__ int upper_thread(int t)
{
    const int j = t & 63;
    return (j & 15) | ((t >> 6) << 4) | ((j >> 4) << 6);
}

// Passes whose gates all sit on unit bits 6..11 (complex64: and the in-unit
// bit), without fans: one tile per workgroup and no LDS, so the workgroups per
// CU are bound by registers instead of by a 64 KiB tile buffer (more tiles in
// flight per CU).  Each thread holds the units of layout-0 thread
// upper_thread(t), registers = unit bits 8..11 as in layout 0.  After the
// layout-0 gates the lane swaps make unit bits 6, 7 register bits 0, 1 (and
// unit bits 8, 9 lane bits 4, 5) for their gates, and are undone before the
// diagonal epilogue and the store.  The gates run in the tile kernel's order
// (layout 0, then unit bits 6, 7) with the same arithmetic in complex128;
// complex64 gates run on the VALU here and on the matrix cores in the tile
// kernel, which agree to float rounding.  kFlip: per-pair
// selectors; the rpar masks of slots 6 and 7 are then those of the swapped
// registers (unit bits 6, 7, 10, 11), set by the host.
template <bool kC64, bool kDiag, bool kFlip, bool kTab = false>
__global__ __launch_bounds__(kThreads) void upper_kernel(void *__restrict__ state, const TPass P,
                                                         const TTab *__restrict__ tabp,
                                                         const double *__restrict__ ttab,
                                                         const double *__restrict__ tphase)
{
    static_assert(!(kDiag && kFlip), "no phase-polynomial epilogue with per-pair selectors");
    static_assert(!(kDiag && kTab), "one epilogue kind");
    using T = typename UnitT<kC64>::T;
    __shared__ WinT<kC64> wt[kTab ? (1 << kWinBits) : 1];
    const int tl = upper_thread(threadIdx.x);
    T *p = static_cast<T *>(state);
    const uint64_t tile = blockIdx.x;
    const uint64_t base = deposit(tile, P.sorted);
    const uint64_t toff = thr_off<0>(tl, P);
    const uint64_t at = base ^ toff;
    ZXA_CONST TTab *G = (ZXA_CONST TTab *)tabp;
    T v[kRegs];
    load_tile<kC64>(v, p, at, G);
    if constexpr (kTab)
        win_load<kC64>(wt, ttab, G->win, threadIdx.x); // synchronized before its use below
    // VALU arithmetic here: with the matrix cores the compiler moves the
    // results through AGPRs and the occupancy this kernel relies on halves
    apply_layout<kC64, kFlip, 0, false, false>(v, P, G, base, toff);
#pragma unroll
    for (int r = 0; r < kRegs; r += 2)
        lane_swap<false>(v[r], v[r + 1]);
#pragma unroll
    for (int r = 0; r < kRegs; r++)
        if (!(r & 2))
            lane_swap<true>(v[r], v[r + 2]);
    // unit address of register 0 in the swapped layout: lane bits 4, 5 now
    // stand for unit bits 8, 9 instead of 6, 7
    uint64_t toff_s = 0;
    if constexpr (kFlip)
        toff_s = toff ^ (((tl >> 6) & 1) ? P.vec[6] ^ P.vec[8] : 0) ^ (((tl >> 7) & 1) ? P.vec[7] ^ P.vec[9] : 0);
#pragma unroll
    for (int i = 0; i < 2; i++)
        gate_on_reg_bit<kC64, kFlip, false>(v, i, G->g[6 + i], base, toff_s);
#pragma unroll
    for (int r = 0; r < kRegs; r++)
        if (!(r & 2))
            lane_swap<true>(v[r], v[r + 2]);
#pragma unroll
    for (int r = 0; r < kRegs; r += 2)
        lane_swap<false>(v[r], v[r + 1]);
    if constexpr (kDiag)
    {
        Epi<kC64> epi;
        epi.init(ttab, tl);
        epi.apply(v, (ZXA_CONST double *)(tphase + tile * Epi<kC64>::kStride), tl);
    }
    if constexpr (kTab)
    {
        __syncthreads();
        win_apply<kC64>(v, wt, G->win, win_y(G->win, base) ^ win_y(G->win, toff));
    }
    store_tile<kC64>(v, p, at, G);
}

// ---------------------------------------------------------------------------
// Wide tile pass: 8192-unit tiles (13 unit bits) for passes of up to nine
// targets above the contiguous run.  Unit bits 0..3 are the run (256-byte
// runs of 16-byte units, where the 4096-unit tile with nine such targets
// has 128-byte runs); unit bits 4..12 are the nine slots.  512 threads hold
// 16 units each:
//   layout 0 (load, store): thread bits = unit bits 0..8, register bits =
//     unit bits 9..12; unit bit 4 is lane bit 4, reached by exchanging it
//     with register bit 0 (v_permlane16_swap);
//   layout 1: thread bits = unit bits 0..4 and 9..12, register bits = unit
//     bits 5..8, through LDS (the whole tile, 128 KiB: one workgroup per CU).
// Each thread writes back exactly the LDS units it read, so a transpose
// needs one barrier, after its writes.  Pipelined as tile_kernel.
constexpr int kW13Threads = 512;
constexpr int kW13Bits = 13;
constexpr int kW13Units = 1 << kW13Bits;
constexpr int kW13Slots = 9; // unit bits 4..12

struct P13
{
    uint32_t sorted[kW13Bits]; // the tile's unit-address bits, ascending
    uint64_t ntiles;
    uint64_t tmask;
    uint64_t gstep;
};

struct Tab13
{
    uint64_t roff[kRegs];      // layout-0 register offsets (unit bits 9..12)
    uint64_t toffv[9];         // unit-address offset of each thread bit (unit bits 0..8)
    TGate g[kW13Slots];        // slot s: unit bit 4 + s (the identity where no gate)
};

__host__ __device__ __forceinline__ uint64_t deposit13(uint64_t tile, const uint32_t *ub)
{
    for (int k = 0; k < kW13Bits; k++)
    {
        const unsigned p = ub[k];
        tile = ((tile >> p) << (p + 1)) | (tile & ((uint64_t(1) << p) - 1));
    }
    return tile;
}

__device__ __forceinline__ int w13_l1(int t, int r)
{
    return (t & 31) | (r << 5) | ((t >> 5) << 9);
}

template <bool kC64, bool kDiag>
__device__ __forceinline__ void w13_work(typename UnitT<kC64>::T (&v)[kRegs], typename UnitT<kC64>::T *lds,
                                         ZXA_CONST Tab13 *G, const double *tphase, const Epi<kC64, 9> &epi, int t,
                                         uint64_t tile)
{
#pragma unroll
    for (int i = 0; i < 4; i++)
        gate_on_reg_bit<kC64, false>(v, i, G->g[5 + i], 0, 0);
#pragma unroll
    for (int r = 0; r < kRegs; r += 2)
        lane_swap<false>(v[r], v[r + 1]);
    gate_on_reg_bit<kC64, false>(v, 0, G->g[0], 0, 0);
#pragma unroll
    for (int r = 0; r < kRegs; r += 2)
        lane_swap<false>(v[r], v[r + 1]);
#pragma unroll
    for (int r = 0; r < kRegs; r++)
        lds[t | (r << 9)] = v[r];
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kRegs; r++)
        v[r] = lds[w13_l1(t, r)];
#pragma unroll
    for (int i = 0; i < 4; i++)
        gate_on_reg_bit<kC64, false>(v, i, G->g[1 + i], 0, 0);
#pragma unroll
    for (int r = 0; r < kRegs; r++)
        lds[w13_l1(t, r)] = v[r];
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kRegs; r++)
        v[r] = lds[t | (r << 9)];
    if constexpr (kDiag)
        epi.apply(v, (ZXA_CONST double *)(tphase + tile * Epi<kC64, 9>::kStride), t);
}

template <bool kC64>
__device__ __forceinline__ void w13_load(typename UnitT<kC64>::T (&v)[kRegs], const typename UnitT<kC64>::T *p,
                                         uint64_t at, ZXA_CONST Tab13 *G)
{
#pragma unroll
    for (int r = 0; r < kRegs; r++)
        v[r] = __builtin_nontemporal_load(p + (at ^ G->roff[r]));
}

template <bool kC64>
__device__ __forceinline__ void w13_store(const typename UnitT<kC64>::T (&v)[kRegs], typename UnitT<kC64>::T *p,
                                          uint64_t at, ZXA_CONST Tab13 *G)
{
    asm volatile("" : "+v"(at));
#pragma unroll
    for (int r = 0; r < kRegs; r++)
        __builtin_nontemporal_store(v[r], p + (at ^ G->roff[r]));
}

// The wide pass's table travels as the kernel's first argument, so it sits at
// offset 0 of the kernel-argument segment (explicit arguments are laid out
// from offset 0 in declaration order) and reaches the device with the launch
// itself: no staging copy.  It is read in place through the segment pointer
// (scalar loads); the empty asm keeps each re-read from being hoisted.
__device__ __forceinline__ ZXA_CONST Tab13 *tables13()
{
    uint64_t a = reinterpret_cast<uint64_t>(__builtin_amdgcn_kernarg_segment_ptr());
    asm volatile("" : "+s"(a));
    return reinterpret_cast<ZXA_CONST Tab13 *>(a);
}

__device__ __forceinline__ uint64_t next_base13(uint64_t base, const P13 &P)
{
    return ((base | P.tmask) + P.gstep) & ~P.tmask;
}

template <bool kC64, bool kDiag>
__global__ __launch_bounds__(kW13Threads) __attribute__((amdgpu_waves_per_eu(2))) void wide13_kernel(
    const Tab13 tab, void *__restrict__ state, const P13 P, const double *__restrict__ ttab,
    const double *__restrict__ tphase)
{
    using T = typename UnitT<kC64>::T;
    __shared__ T lds[kW13Units];
    const int t = threadIdx.x;
    T *p = static_cast<T *>(state);
    ZXA_CONST Tab13 *G = tables13();
    uint64_t toff = 0;
#pragma unroll
    for (int j = 0; j < 9; j++)
        toff ^= ((t >> j) & 1) ? G->toffv[j] : 0;
    Epi<kC64, 9> epi;
    if constexpr (kDiag)
        epi.init(ttab, t);
    const uint64_t grid = gridDim.x;
    T a[kRegs], b[kRegs];
    uint64_t t0 = blockIdx.x;
    uint64_t b0 = deposit13(t0, P.sorted);
    w13_load<kC64>(a, p, b0 ^ toff, G);
    uint64_t t1 = t0 + grid;
    bool m1 = t1 < P.ntiles;
    uint64_t b1 = next_base13(b0, P);
    w13_load<kC64>(b, p, (m1 ? b1 : b0) ^ toff, G);
    w13_work<kC64, kDiag>(a, lds, G, tphase, epi, t, t0);
    for (;;)
    {
        G = tables13();
        w13_store<kC64>(a, p, b0 ^ toff, G);
        if (!m1)
            break;
        const uint64_t t2 = t1 + grid;
        const bool m2 = t2 < P.ntiles;
        const uint64_t b2 = next_base13(b1, P);
        w13_load<kC64>(a, p, (m2 ? b2 : b1) ^ toff, G);
        w13_work<kC64, kDiag>(b, lds, G, tphase, epi, t, t1);

        G = tables13();
        w13_store<kC64>(b, p, b1 ^ toff, G);
        if (!m2)
            break;
        const uint64_t t3 = t2 + grid;
        const bool m3 = t3 < P.ntiles;
        const uint64_t b3 = next_base13(b2, P);
        w13_load<kC64>(b, p, (m3 ? b3 : b2) ^ toff, G);
        w13_work<kC64, kDiag>(a, lds, G, tphase, epi, t, t2);

        t0 = t2;
        b0 = b2;
        t1 = t3;
        b1 = b3;
        m1 = m3;
    }
}

// ---------------------------------------------------------------------------
// Adjoint pass: the reverse of one forward tile pass, on two states at once.
// psi is the ansatz state and lam = (later gates)^dagger H psi, both taken at
// the point just after the forward pass's single-qubit gates.  A rotation's
// gradient is a transition value Im <lam| A |psi>, with A the rotation's
// generator conjugated by the gates composed after it on the same qubit: a
// traceless Hermitian 2x2 operator [[a, b], [conj(b), -a]] on the slot's
// qubit.  For such A
//     Im <lam|A|psi> = a S_d + Re(b) S_e + Im(b) S_f
// with three sums over the slot's amplitude pairs (p0, p1), (l0, l1):
//     S_d = Im(conj(l0) p0) - Im(conj(l1) p1)
//     S_e = Im(conj(l0) p1) + Im(conj(l1) p0)
//     S_f = Re(conj(l0) p1) - Re(conj(l1) p0)
// so the kernel accumulates the three sums of every slot (whatever the
// operators), and the collection combines them with each operator's (a, b).
// After a layout's sums the pass applies its slots' uncompute matrices
// M = U^dagger to both states.  The diagonal epilogue (if any) is the
// previous forward pass's window with negated angles, on both states.
// Per tile the sums are reduced over the wave by recursive halving: a step
// on lane bit m splits a lane's list of length len at h = ceil(len / 2);
// lanes with bit m keep (and add the partner's copy of) the upper part, the
// others the lower part.  Each layout's list (its 4 slots x 3 sums, plus the
// complex64 in-unit slot in layout 0) takes the steps of bits 32 and 16 right
// after it is formed; the three remainders, concatenated, take the steps of
// bits 8, 4, 2, 1.  Then every lane holds the wave total of one sum (or 0),
// which it adds into its own double accumulator.
// ---------------------------------------------------------------------------

constexpr int kObs = 2;  // gradient operators per slot
constexpr int kSums = 3; // bilinear sums per slot
constexpr int kWavesPerTile = kThreads / 64;
constexpr int kAccs = (kUnitBits + 1) * kObs; // gradient entries per pass

__host__ __device__ constexpr int half_up(int n)
{
    return (n + 1) / 2;
}
// list length of layout L's sums, and of its remainder after two steps
__host__ __device__ constexpr int group_len(bool c64, int L)
{
    return 4 * kSums + (c64 && L == 0 ? kSums : 0);
}
__host__ __device__ constexpr int group_rest(bool c64, int L)
{
    return half_up(half_up(group_len(c64, L)));
}
template <bool kC64>
constexpr int kRest = group_rest(kC64, 0) + group_rest(kC64, 1) + group_rest(kC64, 2);

// Lane holding the wave total of sum j of layout L's list.
__host__ __device__ constexpr int sum_lane(bool c64, int L, int j)
{
    int lane = 0;
    int len = group_len(c64, L);
    for (int m = 32; m >= 16; m >>= 1)
    {
        const int h = half_up(len);
        if (j >= h)
        {
            j -= h;
            lane |= m;
        }
        len = h;
    }
    for (int g = 0; g < L; g++)
        j += group_rest(c64, g);
    len = group_rest(c64, 0) + group_rest(c64, 1) + group_rest(c64, 2);
    for (int m = 8; m >= 1; m >>= 1)
    {
        const int h = half_up(len);
        if (j >= h)
        {
            j -= h;
            lane |= m;
        }
        len = h;
    }
    return lane;
}
static_assert(kRest<true> <= 16 && kRest<false> <= 16, "four final steps reduce the remainders to one");

template <int kLen, int kMask, typename R>
__device__ __forceinline__ void halve(const R (&x)[kLen], R (&y)[half_up(kLen)], int lane)
{
    constexpr int kH = half_up(kLen);
    const bool up = (lane & kMask) != 0;
#pragma unroll
    for (int i = 0; i < kH; i++)
    {
        const R lo = x[i];
        const R hi = i + kH < kLen ? x[i + kH] : R(0);
        y[i] = (up ? hi : lo) + R(__shfl_xor(up ? lo : hi, kMask));
    }
}

template <int kLen, int kMask>
__device__ __forceinline__ double halve_all(const double (&x)[kLen], int lane)
{
    if constexpr (kLen == 1)
    {
        return x[0];
    }
    else
    {
        double y[half_up(kLen)];
        halve<kLen, kMask, double>(x, y, lane);
        return halve_all<half_up(kLen), kMask / 2>(y, lane);
    }
}

// One gradient entry of a pass: param (< 0: none) += c[0] S_d + c[1] S_e +
// c[2] S_f of the slot whose sums sit in lanes lane[0..2].
struct TGrad
{
    int32_t param;
    int32_t lane[kSums];
    double c[kSums];
};

// The two off-diagonal sums of one amplitude pair (p0, p1), (l0, l1),
// accumulated: S_e and S_f.
template <typename R>
__device__ __forceinline__ void pair_cross(R p0x, R p0y, R p1x, R p1y, R l0x, R l0y, R l1x, R l1y, R &se, R &sf)
{
    se += l0x * p1y - l0y * p1x + l1x * p0y - l1y * p0x;
    sf += l0x * p1x + l0y * p1y - l1x * p0x - l1y * p0y;
}

// Im(conj(l) p) of one amplitude.
template <typename R>
__device__ __forceinline__ R im_lp(R px, R py, R lx, R ly)
{
    return lx * py - ly * px;
}

// One halving step on lane bit 32 or 16 (halve's result) through the lane
// exchange of two lists' words: lanes with the bit keep hi plus the partner's
// hi, the others lo plus the partner's lo (the partner's copy added first
// there; the sums are the same).
template <int kLen, int kMask, typename R>
__device__ __forceinline__ void halve_swap(const R (&x)[kLen], R (&y)[half_up(kLen)])
{
    static_assert(kMask == 32 || kMask == 16, "lane exchange steps");
    constexpr int kW = sizeof(R) / 4;
    struct Words
    {
        uint32_t w[kW];
    };
    constexpr int kH = half_up(kLen);
#pragma unroll
    for (int i = 0; i < kH; i++)
    {
        Words lo = __builtin_bit_cast(Words, x[i]);
        Words hi = __builtin_bit_cast(Words, i + kH < kLen ? x[i + kH] : R(0));
#pragma unroll
        for (int k = 0; k < kW; k++)
        {
            const auto r = kMask == 32 ? __builtin_amdgcn_permlane32_swap(lo.w[k], hi.w[k], false, false)
                                       : __builtin_amdgcn_permlane16_swap(lo.w[k], hi.w[k], false, false);
            lo.w[k] = r[0];
            hi.w[k] = r[1];
        }
        y[i] = __builtin_bit_cast(R, lo) + __builtin_bit_cast(R, hi);
    }
}

// Slots of layout L (register bit i at list position 3 i; the complex64
// in-unit slot of layout 0 at 12): every slot's sums on the state as it
// enters the layout, then the slots' matrices M on both states; the list's
// first two halving steps into rest.  A slot's sums do not depend on which
// of the other slots' M were applied before them: each M acts on another
// qubit, commutes with the slot's operator and is the same unitary on both
// states.  So S_d, a signed sum of Im(conj(l) p) over the registers by the
// slot's bit, shares one product per register among the slots.
template <bool kC64, int L>
__device__ __forceinline__ void adj_layout(typename UnitT<kC64>::T (&v)[kRegs], typename UnitT<kC64>::T (&w)[kRegs],
                                           ZXA_CONST TTab *G, Real<kC64> (&rest)[group_rest(kC64, L)])
{
    using R = Real<kC64>;
    constexpr int kN = group_len(kC64, L);
    R s[kN];
    // S_e, S_f per slot
#pragma unroll
    for (int i = 0; i < 4; i++)
    {
        R se = 0, sf = 0;
#pragma unroll
        for (int r = 0; r < kRegs; r++)
        {
            if ((r >> i) & 1)
                continue;
            const int r1 = r | (1 << i);
            pair_cross<R>(v[r].x, v[r].y, v[r1].x, v[r1].y, w[r].x, w[r].y, w[r1].x, w[r1].y, se, sf);
            if constexpr (kC64)
                pair_cross<float>(v[r].z, v[r].w, v[r1].z, v[r1].w, w[r].z, w[r].w, w[r1].z, w[r1].w, se, sf);
        }
        s[i * kSums + 1] = se;
        s[i * kSums + 2] = sf;
    }
    // S_d per slot: a butterfly over the registers' Im(conj(l) p), bit 0
    // first (d: pair sums, e: pair differences)
    R d[kRegs];
    R dlo = 0, dhi = 0; // complex64: each amplitude's part, for the in-unit slot
#pragma unroll
    for (int r = 0; r < kRegs; r++)
    {
        d[r] = im_lp<R>(v[r].x, v[r].y, w[r].x, w[r].y);
        if constexpr (kC64)
        {
            const float dh = im_lp<float>(v[r].z, v[r].w, w[r].z, w[r].w);
            if constexpr (L == 0)
            {
                dlo += d[r];
                dhi += dh;
            }
            d[r] += dh;
        }
    }
#pragma unroll
    for (int i = 0; i < 4; i++)
    {
        const int len = kRegs >> i;
        R e = 0;
#pragma unroll
        for (int r = 0; r < len; r += 2)
        {
            e += d[r] - d[r + 1];
            d[r >> 1] = d[r] + d[r + 1];
        }
        s[i * kSums + 0] = e;
    }
    if constexpr (kC64 && L == 0)
    {
        // the in-unit slot: pairs (first, second amplitude) of every unit
        R se = 0, sf = 0;
#pragma unroll
        for (int r = 0; r < kRegs; r++)
            pair_cross<float>(v[r].x, v[r].y, v[r].z, v[r].w, w[r].x, w[r].y, w[r].z, w[r].w, se, sf);
        s[4 * kSums + 0] = dlo - dhi;
        s[4 * kSums + 1] = se;
        s[4 * kSums + 2] = sf;
    }
    // the sums read the layout's inputs: formed here (not sunk towards their
    // use after the matrices, which would keep old copies of the states)
#pragma unroll
    for (int j = 0; j < kN; j++)
        asm volatile("" : "+v"(s[j]));
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int i = 0; i < 4; i++)
    {
        const int k = reg_bit(L, i);
        R m[8];
        load_matrix<kC64>(m, G->g[k]);
        if constexpr (kC64)
        {
            const MRow mr = mfma_row(m);
#pragma unroll
            for (int r = 0; r < kRegs; r++)
            {
                if ((r >> i) & 1)
                    continue;
                mfma_units(mr, v[r], v[r | (1 << i)]);
                mfma_units(mr, w[r], w[r | (1 << i)]);
            }
        }
        else
        {
#pragma unroll
            for (int r = 0; r < kRegs; r++)
            {
                if ((r >> i) & 1)
                    continue;
                unit_pair<kC64, false>(v[r], v[r | (1 << i)], m, false, false);
                unit_pair<kC64, false>(w[r], w[r | (1 << i)], m, false, false);
            }
        }
        // one slot at a time: interleaving slots only lengthens live ranges
        __builtin_amdgcn_sched_barrier(0);
    }
    if constexpr (kC64 && L == 0)
    {
        float m[8];
        load_matrix<true>(m, G->g[kIntra]);
        const MRow mr = mfma_row(m);
#pragma unroll
        for (int r = 0; r < kRegs; r++)
        {
            mfma_intra(mr, v[r]);
            mfma_intra(mr, w[r]);
        }
    }
    R h[half_up(kN)];
    halve_swap<kN, 32, R>(s, h);
    halve_swap<half_up(kN), 16, R>(h, rest);
#pragma unroll
    for (int i = 0; i < group_rest(kC64, L); i++)
        asm volatile("" : "+v"(rest[i]));
}

// Remainders of the three layouts (zeros for an unused layout).
template <bool kC64>
struct AdjRest
{
    Real<kC64> r0[group_rest(kC64, 0)], r1[group_rest(kC64, 1)], r2[group_rest(kC64, 2)];
    __device__ __forceinline__ double total(int lane) const
    {
        double c[kRest<kC64>];
        int n = 0;
#pragma unroll
        for (int i = 0; i < group_rest(kC64, 0); i++)
            c[n++] = double(r0[i]);
#pragma unroll
        for (int i = 0; i < group_rest(kC64, 1); i++)
            c[n++] = double(r1[i]);
#pragma unroll
        for (int i = 0; i < group_rest(kC64, 2); i++)
            c[n++] = double(r2[i]);
        return halve_all<kRest<kC64>, 8>(c, lane);
    }
};

// All layouts of one tile on both states.
template <bool kC64, bool kL1, bool kL2>
__device__ __forceinline__ void adj_tile(typename UnitT<kC64>::T (&v)[kRegs], typename UnitT<kC64>::T (&w)[kRegs],
                                         typename UnitT<kC64>::T *lds, ZXA_CONST TTab *G, int t, AdjRest<kC64> &R)
{
#pragma unroll
    for (int i = 0; i < group_rest(kC64, 1); i++)
        R.r1[i] = 0;
#pragma unroll
    for (int i = 0; i < group_rest(kC64, 2); i++)
        R.r2[i] = 0;
    adj_layout<kC64, 0>(v, w, G, R.r0);
    if constexpr (kL1)
    {
        transpose<kC64, 0, 1>(v, lds, t);
        transpose<kC64, 0, 1>(w, lds, t);
        adj_layout<kC64, 1>(v, w, G, R.r1);
        if constexpr (kL2)
        {
            transpose<kC64, 1, 2>(v, lds, t);
            transpose<kC64, 1, 2>(w, lds, t);
            adj_layout<kC64, 2>(v, w, G, R.r2);
            transpose<kC64, 2, 0>(v, lds, t);
            transpose<kC64, 2, 0>(w, lds, t);
        }
        else
        {
            transpose<kC64, 1, 0>(v, lds, t);
            transpose<kC64, 1, 0>(w, lds, t);
        }
    }
    else if constexpr (kL2)
    {
        transpose<kC64, 0, 2>(v, lds, t);
        transpose<kC64, 0, 2>(w, lds, t);
        adj_layout<kC64, 2>(v, w, G, R.r2);
        transpose<kC64, 2, 0>(v, lds, t);
        transpose<kC64, 2, 0>(w, lds, t);
    }
}

// partial: one row of 64 lane accumulators per wave.
template <bool kC64, bool kL1, bool kL2, bool kDiag>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(2))) void adj_kernel(
    void *__restrict__ psi, void *__restrict__ lam, const TPass P, const TTab *__restrict__ tabp,
    const double *__restrict__ ttab, const double *__restrict__ tphase, double *__restrict__ partial)
{
    using T = typename UnitT<kC64>::T;
    __shared__ T lds[(kL1 || kL2) ? kUnits : 1];
    const int t = threadIdx.x;
    T *pp = static_cast<T *>(psi);
    T *pl = static_cast<T *>(lam);
    const uint64_t toff = thr_off<0>(t, P);
    double acc = 0.0;
    uint64_t base = deposit(uint64_t(blockIdx.x), P.sorted);
    for (uint64_t tile = blockIdx.x; tile < P.ntiles; tile += gridDim.x)
    {
        ZXA_CONST TTab *G = tables(tabp);
        T v[kRegs], w[kRegs];
        load_tile<kC64>(v, pp, base ^ toff, G);
        load_tile<kC64>(w, pl, base ^ toff, G);
        AdjRest<kC64> R;
        adj_tile<kC64, kL1, kL2>(v, w, lds, G, t, R);
        if constexpr (kDiag)
        {
            // the per-thread factors (a few KiB, cached) read again per tile
            // rather than held across it: the sums and both states fill the
            // registers
            const double *et = ttab;
            asm volatile("" : "+s"(et));
            Epi<kC64> epi;
            epi.init(et, t);
            epi.template apply_n<true>(v, w, (ZXA_CONST double *)(tphase + tile * Epi<kC64>::kStride), t);
        }
        store_tile<kC64>(v, pp, base ^ toff, G);
        store_tile<kC64>(w, pl, base ^ toff, G);
        acc += R.total(t & 63);
        base = next_base(base, P);
    }
    partial[(size_t(blockIdx.x) * kWavesPerTile + (t >> 6)) * 64 + (t & 63)] = acc;
}

// Sums rows [0, nrows) of 64 lane accumulators in a fixed order, then adds
// every gradient entry into grads[param] (param < nparams) in entry order, so
// that entries of one pass sharing a parameter never race.
__device__ __forceinline__ void collect_rows(const double *__restrict__ partial, uint64_t nrows,
                                             const TGrad *__restrict__ tg, int64_t nparams, double *__restrict__ grads)
{
    static_assert(kWavesPerTile == 4, "four row groups");
    __shared__ double part[kWavesPerTile][64];
    __shared__ double tot[64];
    const int lane = threadIdx.x & 63;
    const int q = threadIdx.x >> 6;
    // four independent partial sums per thread (loads in flight), in a fixed
    // order: deterministic
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    uint64_t r = q;
    for (; r + 3 * kWavesPerTile < nrows; r += 4 * kWavesPerTile)
    {
        s0 += partial[r * 64 + lane];
        s1 += partial[(r + kWavesPerTile) * 64 + lane];
        s2 += partial[(r + 2 * kWavesPerTile) * 64 + lane];
        s3 += partial[(r + 3 * kWavesPerTile) * 64 + lane];
    }
    for (; r < nrows; r += kWavesPerTile)
        s0 += partial[r * 64 + lane];
    part[q][lane] = (s0 + s1) + (s2 + s3);
    __syncthreads();
    if (threadIdx.x < 64)
        tot[lane] = (part[0][lane] + part[1][lane]) + (part[2][lane] + part[3][lane]);
    __syncthreads();
    if (threadIdx.x == 0)
    {
        for (int a = 0; a < kAccs; a++)
        {
            const TGrad &g = tg[a];
            if (g.param < 0 || g.param >= nparams)
                continue;
            grads[g.param] += g.c[0] * tot[g.lane[0] & 63] + g.c[1] * tot[g.lane[1] & 63] +
                              g.c[2] * tot[g.lane[2] & 63];
        }
    }
}

__global__ __launch_bounds__(kThreads) void adj_collect_kernel(const double *__restrict__ partial, uint64_t nrows,
                                                               const TGrad *__restrict__ tg, int64_t nparams,
                                                               double *__restrict__ grads)
{
    collect_rows(partial, nrows, tg, nparams, grads);
}

__device__ __forceinline__ cd cis(double a)
{
    double s, c;
    sincos(a, &s, &c);
    return {c, s};
}

// Per-tile phase factors of the diagonal epilogue.
// Tile index bit i is address bit outer[i].  tab (nE tile element bits, nh
// outer bits, nR register elements):
//     cq[k * nh + i] = Q(tile bit k, outer i)   k < nE
//     hL[i] = L(outer i), hQ[i * nh + l] = Q(outer i, outer l)  (i < l)
//     rs[rho]        = static phase of register element rho
//     map[0..8)      = tile element index of thread bit j
//     map[8..8+nb)   = tile element index of register element bit b
// Output per tile: E_j = exp(i c(thr j)) (8), D[rho] = exp(i (C + rs[rho] +
// sum_{b in rho} c(reg b))) (nR), as complex pairs.

// c_k of a tile: the k-th element coefficient summed over the tile's bits.
__device__ __forceinline__ double tp_coef(const double *cq, int nh, uint64_t tile, int k)
{
    double s = 0.0;
    for (int i = 0; i < nh; i++)
        if ((tile >> i) & 1u)
            s += cq[k * nh + i];
    return s;
}

// The tile's constant: K plus its linear and quadratic terms.
__device__ __forceinline__ double tp_const(const double *hL, const double *hQ, int nh, uint64_t tile, double K)
{
    double C = K;
    for (int i = 0; i < nh; i++)
    {
        if (!((tile >> i) & 1u))
            continue;
        double s = hL[i];
        for (int l = i + 1; l < nh; l++)
            if ((tile >> l) & 1u)
                s += hQ[i * nh + l];
        C += s;
    }
    return C;
}

// Per-tile phase factors of the diagonal epilogue: 8 E factors and 2^nb D
// factors per tile.  spread = 0: one thread per tile; spread = 1: one thread
// per (tile, factor), for states of few tiles (the same sums in the same
// order, so the same factors).
__device__ __forceinline__ void tile_phase_work(const double *__restrict__ tab, int nE, int nh, int nb, double K,
                                                uint64_t ntiles, double *__restrict__ out, int spread, int nthr,
                                                uint64_t id, double *sh)
{
    const int nR = 1 << nb;
    const int ntab = nE * nh + nh + nh * nh + nR + nthr + nb;
    for (int i = threadIdx.x; i < ntab; i += blockDim.x)
        sh[i] = tab[i];
    __syncthreads();
    const double *cq = sh;
    const double *hL = cq + nE * nh;
    const double *hQ = hL + nh;
    const double *rs = hQ + nh * nh;
    const double *mp = rs + nR;
    const int per = nthr + nR;
    const int stride = 2 * per;
    if (spread)
    {
        if (id >= ntiles * uint64_t(per))
            return;
        const uint64_t tile = id / uint64_t(per);
        const int j = int(id % uint64_t(per));
        double a;
        if (j < nthr)
            a = tp_coef(cq, nh, tile, static_cast<int>(mp[j]));
        else
        {
            const int rho = j - nthr;
            a = tp_const(hL, hQ, nh, tile, K) + rs[rho];
            for (int b = 0; b < nb; b++)
                if ((rho >> b) & 1)
                    a += tp_coef(cq, nh, tile, static_cast<int>(mp[nthr + b]));
        }
        const cd e = cis(a);
        out[tile * stride + 2 * j] = e.x;
        out[tile * stride + 2 * j + 1] = e.y;
        return;
    }
    const uint64_t tile = id;
    if (tile >= ntiles)
        return;
    double c[14];
    for (int k = 0; k < nE; k++)
        c[k] = tp_coef(cq, nh, tile, k);
    const double C = tp_const(hL, hQ, nh, tile, K);
    double *o = out + tile * stride;
    for (int j = 0; j < nthr; j++)
    {
        const cd e = cis(c[static_cast<int>(mp[j])]);
        o[2 * j] = e.x;
        o[2 * j + 1] = e.y;
    }
    for (int rho = 0; rho < nR; rho++)
    {
        double a = C + rs[rho];
        for (int b = 0; b < nb; b++)
            if ((rho >> b) & 1)
                a += c[static_cast<int>(mp[nthr + b])];
        const cd d = cis(a);
        o[2 * nthr + 2 * rho] = d.x;
        o[2 * nthr + 2 * rho + 1] = d.y;
    }
}

__global__ __launch_bounds__(kThreads) void tile_phase_kernel(const double *__restrict__ tab, int nE, int nh, int nb,
                                                              double K, uint64_t ntiles, double *__restrict__ out,
                                                              int spread, int nthr)
{
    extern __shared__ double sh[];
    tile_phase_work(tab, nE, nh, nb, K, ntiles, out, spread, nthr,
                    blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x, sh);
}

unsigned tile_phase_blocks(uint64_t ntiles, int nb, int nthr, int *spread)
{
    *spread = ntiles < 4096 ? 1 : 0;
    const uint64_t threads = *spread ? ntiles * uint64_t(nthr + (1 << nb)) : ntiles;
    return unsigned((threads + kThreads - 1) / kThreads);
}

// Launches tile_phase_kernel, spread over (tile, factor) pairs when the state
// has few tiles (below 4096 one thread per tile leaves most CUs idle and each
// thread runs 8 + 2^nb sincos in a row).
void launch_tile_phase(hipStream_t stream, const double *tab, size_t tab_doubles, int nE, int nh, int nb, double K,
                       uint64_t ntiles, double *out, int nthr = 8)
{
    int spread = 0;
    const unsigned grid = tile_phase_blocks(ntiles, nb, nthr, &spread);
    hipLaunchKernelGGL(tile_phase_kernel, dim3(grid), dim3(kThreads), tab_doubles * sizeof(double), stream, tab, nE,
                       nh, nb, K, ntiles, out, spread, nthr);
}

template <typename K>
int persistent_grid(Ctx &ctx, K kernel, uint64_t tiles)
{
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kThreads, 0) != hipSuccess || per_cu <= 0)
        per_cu = 1;
    const uint64_t cap = uint64_t(per_cu) * uint64_t(ctx.num_cus > 0 ? ctx.num_cus : 256);
    return static_cast<int>(std::max<uint64_t>(1, std::min<uint64_t>(tiles, cap)));
}

// A fresh tile's work: the input |0...0> has its one nonzero element at
// tile-local unit 0, held by thread 0 in register 0 in every layout (the
// layouts' thread and register maps are linear), so it is formed directly in
// the first layout the pass uses beyond layout 0, and the gates run from
// there, ending in layout 0: one transpose fewer than tile_work.  The gates
// of a pass without fans commute, so their order across layouts is free; the
// window comes last, in layout 0, as in tile_work.
template <bool kC64, bool kL1, bool kL2, bool kDiag, bool kTab>
__device__ __forceinline__ void fresh_work(typename UnitT<kC64>::T (&v)[kRegs], typename UnitT<kC64>::T *lds,
                                           const TPass &P, ZXA_CONST TTab *G, const double *tphase,
                                           const Epi<kC64> &epi, int t, uint64_t toff, uint64_t tile, uint64_t base,
                                           const WinT<kC64> *wt, uint32_t yt)
{
    if constexpr (kL1)
    {
        apply_layout<kC64, false, 1>(v, P, G, base, thr_off<1>(t, P), t, nullptr, nullptr);
        if constexpr (kL2)
        {
            transpose<kC64, 1, 2>(v, lds, t);
            apply_layout<kC64, false, 2>(v, P, G, base, thr_off<2>(t, P), t, nullptr, nullptr);
            transpose<kC64, 2, 0>(v, lds, t);
        }
        else
        {
            transpose<kC64, 1, 0>(v, lds, t);
        }
    }
    else if constexpr (kL2)
    {
        apply_layout<kC64, false, 2>(v, P, G, base, thr_off<2>(t, P), t, nullptr, nullptr);
        transpose<kC64, 2, 0>(v, lds, t);
    }
    apply_layout<kC64, false, 0>(v, P, G, base, toff, t, nullptr, nullptr);
    if constexpr (kDiag)
        epi.apply(v, (ZXA_CONST double *)(tphase + tile * Epi<kC64>::kStride), t);
    if constexpr (kTab)
        win_apply<kC64>(v, wt, G->win, win_y(G->win, base) ^ yt);
}

// Fresh tile pass: the first pass over a state just reset to |0...0>
// (amplitude 0 is 1, every other 0).  Each tile's input is formed in
// registers instead of read: the reset's sweep and the pass's read merge into
// the pass's store, and every amplitude still goes through the pass's gates,
// transposes, fans and window before it is stored.  Stores only, so no
// prefetch.
template <bool kC64, bool kL1, bool kL2, bool kDiag, bool kTab, bool kFan = false>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(2))) void tile_fresh_kernel(
    void *__restrict__ state, const TPass P, const TTab *__restrict__ tabp, const double *__restrict__ ttab,
    const double *__restrict__ tphase)
{
    static_assert(!(kTab && (kDiag || kFan)), "one epilogue kind");
    static_assert(!(kFan && kC64), "fans need complex128");
    using T = typename UnitT<kC64>::T;
    __shared__ T lds[(kL1 || kL2) ? kUnits : 1];
    __shared__ dvec2 fl[kFan ? (kUnitBits + 1) * 32 : 1];
    __shared__ WinT<kC64> wt[kTab ? (1 << kWinBits) : 1];
    const int t = threadIdx.x;
    T *p = static_cast<T *>(state);
    const uint64_t toff = thr_off<0>(t, P);
    Epi<kC64> epi;
    if constexpr (kDiag)
        epi.init(ttab, t);
    if constexpr (kFan)
    {
        for (int e = t; e < (kUnitBits + 1) * 32; e += kThreads)
            fl[e] = dvec2{tabp->fan[e >> 5].nib[e & 31][0], tabp->fan[e >> 5].nib[e & 31][1]};
        __syncthreads();
    }
    uint32_t yt = 0;
    if constexpr (kTab)
    {
        ZXA_CONST TTab *G0 = tables(tabp);
        win_load<kC64>(wt, ttab, G0->win, t);
        yt = win_y(G0->win, toff);
        __syncthreads();
    }
    T one{}, zero{};
    one.x = 1;
    uint64_t base = deposit(uint64_t(blockIdx.x), P.sorted);
    for (uint64_t tile = blockIdx.x; tile < P.ntiles; tile += gridDim.x)
    {
        ZXA_CONST TTab *G = tables(tabp);
        const uint64_t at = base ^ toff;
        T v[kRegs];
        if constexpr (kFan)
        {
            // fans keep the circuit's order of the layouts: formed in layout 0
#pragma unroll
            for (int r = 0; r < kRegs; r++)
                v[r] = (at ^ G->roff[r]) == 0 ? one : zero;
            tile_work<kC64, kL1, kL2, kDiag, false, kFan, kTab>(v, lds, P, G, tphase, epi, t, toff, tile, base, fl,
                                                               wt, yt);
        }
        else
        {
#pragma unroll
            for (int r = 0; r < kRegs; r++)
                v[r] = zero;
            if (base == 0 && t == 0)
                v[0] = one;
            fresh_work<kC64, kL1, kL2, kDiag, kTab>(v, lds, P, G, tphase, epi, t, toff, tile, base, wt, yt);
        }
        store_tile<kC64>(v, p, at, G);
        base = next_base(base, P);
    }
}

template <bool kC64, bool kL1, bool kL2, bool kDiag, bool kFlip, bool kFan = false, bool kTab = false>
int launch_t(Ctx &ctx, void *state, const TPass &P0, const TTab *tab, const double *ttab, const double *tphase)
{
    // register buffers per workgroup (knob tile.bufs.<prec>); the flip, fan
    // and window-table variants keep two (their tables take the registers of
    // a third)
    auto k = tile_kernel<kC64, kL1, kL2, kDiag, kFlip, kFan, kTab>;
    if constexpr (!kFlip && !kFan && !kTab)
        if (tune_get(kC64 ? "tile.bufs.c64" : "tile.bufs.c128") == 3)
            k = tile_kernel3<kC64, kL1, kL2, kDiag, false>;
    const int grid = persistent_grid(ctx, k, P0.ntiles);
    TPass P = P0;
    P.gstep = deposit(uint64_t(grid), P.sorted);
    hipLaunchKernelGGL(k, dim3(grid), dim3(kThreads), 0, ctx.stream, state, P, tab, ttab, tphase);
    return hip_check(hipGetLastError(), "tile_kernel launch");
}

template <bool kC64, bool kDiag, bool kFlip, bool kFan = false, bool kTab = false>
int launch_l(Ctx &ctx, void *state, const TPass &P, bool l1, bool l2, const TTab *tab, const double *ttab,
             const double *tphase)
{
    if (l1 && l2)
        return launch_t<kC64, true, true, kDiag, kFlip, kFan, kTab>(ctx, state, P, tab, ttab, tphase);
    if (l1)
        return launch_t<kC64, true, false, kDiag, kFlip, kFan, kTab>(ctx, state, P, tab, ttab, tphase);
    if (l2)
        return launch_t<kC64, false, true, kDiag, kFlip, kFan, kTab>(ctx, state, P, tab, ttab, tphase);
    return launch_t<kC64, false, false, kDiag, kFlip, kFan, kTab>(ctx, state, P, tab, ttab, tphase);
}

// fan: complex128 passes without flips only (P.fphase set); win: ttab is a
// window table (TTab::win), with or without flips
template <bool kC64>
int launch_p(Ctx &ctx, void *state, const TPass &P, bool l1, bool l2, bool flip, const TTab *tab,
             const double *ttab, const double *tphase, bool fan = false, bool win = false)
{
    if (win)
    {
        if (fan)
        {
            set_error("tile pass: fans with a window table");
            return EINVAL_;
        }
        if (flip)
            return launch_l<kC64, false, true, false, true>(ctx, state, P, l1, l2, tab, ttab, nullptr);
        return launch_l<kC64, false, false, false, true>(ctx, state, P, l1, l2, tab, ttab, nullptr);
    }
    if constexpr (!kC64)
        if (fan)
        {
            if (flip)
            {
                set_error("tile pass: fans with per-pair selectors");
                return EINVAL_;
            }
            if (ttab)
                return launch_l<false, true, false, true>(ctx, state, P, l1, l2, tab, ttab, tphase);
            return launch_l<false, false, false, true>(ctx, state, P, l1, l2, tab, nullptr, nullptr);
        }
    if (fan)
    {
        set_error("tile pass: fans need a complex128 state");
        return EINVAL_;
    }
    if (ttab)
        return launch_l<kC64, true, false>(ctx, state, P, l1, l2, tab, ttab, tphase);
    if (flip)
        return launch_l<kC64, false, true>(ctx, state, P, l1, l2, tab, nullptr, nullptr);
    return launch_l<kC64, false, false>(ctx, state, P, l1, l2, tab, nullptr, nullptr);
}

template <bool kC64, bool kL1, bool kL2, bool kDiag, bool kTab, bool kFan>
int launch_f(Ctx &ctx, void *state, const TPass &P0, const TTab *tab, const double *ttab, const double *tphase)
{
    auto k = tile_fresh_kernel<kC64, kL1, kL2, kDiag, kTab, kFan>;
    const int grid = persistent_grid(ctx, k, P0.ntiles);
    TPass P = P0;
    P.gstep = deposit(uint64_t(grid), P.sorted);
    hipLaunchKernelGGL(k, dim3(grid), dim3(kThreads), 0, ctx.stream, state, P, tab, ttab, tphase);
    return hip_check(hipGetLastError(), "tile_fresh_kernel launch");
}

template <bool kC64, bool kDiag, bool kTab, bool kFan = false>
int launch_fl(Ctx &ctx, void *state, const TPass &P, bool l1, bool l2, const TTab *tab, const double *ttab,
              const double *tphase)
{
    if (l1 && l2)
        return launch_f<kC64, true, true, kDiag, kTab, kFan>(ctx, state, P, tab, ttab, tphase);
    if (l1)
        return launch_f<kC64, true, false, kDiag, kTab, kFan>(ctx, state, P, tab, ttab, tphase);
    if (l2)
        return launch_f<kC64, false, true, kDiag, kTab, kFan>(ctx, state, P, tab, ttab, tphase);
    return launch_f<kC64, false, false, kDiag, kTab, kFan>(ctx, state, P, tab, ttab, tphase);
}

// A fresh pass (no per-pair selectors): win and fan as in launch_p.
template <bool kC64>
int launch_fresh(Ctx &ctx, void *state, const TPass &P, bool l1, bool l2, const TTab *tab, const double *ttab,
                 const double *tphase, bool win, bool fan)
{
    if (fan)
    {
        if constexpr (!kC64)
        {
            if (!win)
                return ttab ? launch_fl<false, true, false, true>(ctx, state, P, l1, l2, tab, ttab, tphase)
                            : launch_fl<false, false, false, true>(ctx, state, P, l1, l2, tab, nullptr, nullptr);
        }
        set_error("fresh tile pass: fans need a complex128 state and a phase-polynomial window");
        return EINVAL_;
    }
    if (win)
        return launch_fl<kC64, false, true>(ctx, state, P, l1, l2, tab, ttab, nullptr);
    if (ttab)
        return launch_fl<kC64, true, false>(ctx, state, P, l1, l2, tab, ttab, tphase);
    return launch_fl<kC64, false, false>(ctx, state, P, l1, l2, tab, nullptr, nullptr);
}

// One upper_kernel workgroup per tile (ttab and tphase set: the diagonal
// epilogue; flip: per-pair selectors).
template <bool kC64>
void launch_u_prec(Ctx &ctx, void *state, const TPass &P, const TTab *tab, const double *ttab,
                   const double *tphase, bool flip, bool win)
{
    const dim3 grid(unsigned(P.ntiles));
    if (win && flip)
        hipLaunchKernelGGL((upper_kernel<kC64, false, true, true>), grid, dim3(kThreads), 0, ctx.stream, state, P,
                           tab, ttab, nullptr);
    else if (win)
        hipLaunchKernelGGL((upper_kernel<kC64, false, false, true>), grid, dim3(kThreads), 0, ctx.stream, state, P,
                           tab, ttab, nullptr);
    else if (ttab)
        hipLaunchKernelGGL((upper_kernel<kC64, true, false>), grid, dim3(kThreads), 0, ctx.stream, state, P, tab,
                           ttab, tphase);
    else if (flip)
        hipLaunchKernelGGL((upper_kernel<kC64, false, true>), grid, dim3(kThreads), 0, ctx.stream, state, P, tab,
                           nullptr, nullptr);
    else
        hipLaunchKernelGGL((upper_kernel<kC64, false, false>), grid, dim3(kThreads), 0, ctx.stream, state, P, tab,
                           nullptr, nullptr);
}

int launch_u(Ctx &ctx, bool c64, void *state, const TPass &P, const TTab *tab, const double *ttab,
             const double *tphase, bool flip, bool win)
{
    if (flip && ttab && !win)
    {
        set_error("tile pass: phase-polynomial epilogue with per-pair selectors");
        return EINVAL_;
    }
    if (c64)
        launch_u_prec<true>(ctx, state, P, tab, ttab, tphase, flip, win);
    else
        launch_u_prec<false>(ctx, state, P, tab, ttab, tphase, flip, win);
    return hip_check(hipGetLastError(), "upper_kernel launch");
}

// A gate's target is the address bit its tile slot stands for: the bit of
// its pair offset for a single-bit offset, else the pivot chosen for it
// (pivot_passes), and the slot's address offset is then the whole offset.
struct SingleGate
{
    unsigned target;
    const U3Gate *g;
};

// Tile plan: the address bit of every tile element bit e (complex128: e is
// the unit bit; complex64: e = 0 is address bit 0 inside a unit and e = k + 1
// is unit bit k), and the unit bit of every target.  Unit bits 0..seg-1 are
// the lowest unit-address bits (seg = 3, 4, 5 or 6: 128 B to 1 KiB
// contiguous per wave access); targets above them go to the register bits of
// layout 0 (unit bits 8..11), then of layout 1 from unit bit 7 down, then
// unit bit 3 (layout 2); the remaining slots take the lowest other bits in
// ascending order (so that in the load layout the contiguous run continues
// across lanes where it can).
struct Plan
{
    std::vector<unsigned> ebit; // tile element bit -> address bit (pivot)
    std::vector<uint64_t> evec; // tile element bit -> address offset
    std::vector<int> slot;      // per gate: unit bit, or kIntra
};

// Staged unit bits (complex128, single-bit targets in circuit order): the
// gates take unit bits in the kernel's processing order, layout 0's
// register bits 8..11, then layout 1's 4..7, then layout 2's 0..3.  A target
// at or above the low bits needs a unit bit at or above seg, a low target one
// below: the contiguous run may hold its address bits in any order, since a
// wave still reads the same 64 contiguous units.  Earliest fit is optimal;
// false if the order cannot be kept.  The unused run bits take the unused
// low address bits; other unit bits stay -1.
bool staged_units(int prec, int geometry, const std::vector<unsigned> &targets, std::vector<int> &unit_addr)
{
    if (prec == C64)
        return false;
    const int seg = int(tile_seg(prec, geometry));
    static const int order[kUnitBits] = {8, 9, 10, 11, 4, 5, 6, 7, 0, 1, 2, 3};
    unit_addr.assign(kUnitBits, -1);
    std::vector<bool> low_used(size_t(seg), false);
    int pos = 0;
    for (unsigned t : targets)
    {
        const bool up = t >= unsigned(seg);
        while (pos < kUnitBits && (order[pos] >= seg) != up)
            pos++;
        if (pos == kUnitBits)
            return false;
        unit_addr[order[pos]] = int(t);
        if (!up)
            low_used[t] = true;
        pos++;
    }
    int next = 0;
    for (int k = 0; k < seg; k++)
    {
        if (unit_addr[k] >= 0)
            continue;
        while (low_used[size_t(next)])
            next++;
        unit_addr[k] = next;
        low_used[size_t(next)] = true;
    }
    return true;
}

// staged: the gates' order is the circuit's (fan passes), unit bits from
// staged_units.
bool plan_tile(int prec, unsigned n, const std::vector<SingleGate> &gs, Plan &pl, int geometry = -1,
               bool staged = false)
{
    const bool c64 = prec == C64;
    const int seg = int(tile_seg(prec, geometry));
    const unsigned low = tile_low_bits(prec, geometry);
    const unsigned shift = c64 ? 1u : 0u;
    std::vector<int> unit_addr(kUnitBits, -1); // unit bit -> address bit
    for (int k = 0; k < seg; k++)
        unit_addr[k] = int(k + shift);
    std::vector<bool> used(64, false);
    for (unsigned b = 0; b < low; b++)
        used[b] = true;
    std::vector<unsigned> upper;
    for (const SingleGate &s : gs)
        if (s.target >= low)
            upper.push_back(s.target);
    if (n < tile_min_bits(prec))
        return false;
    if (staged)
    {
        std::vector<unsigned> targets;
        for (const SingleGate &s : gs)
            targets.push_back(s.target);
        if (!staged_units(prec, geometry, targets, unit_addr))
            return false;
        for (unsigned t : upper)
            used[t] = true;
    }
    else
    {
        std::sort(upper.begin(), upper.end());
        if (upper.size() > tile_upper_slots(prec, geometry))
            return false;
        // targets: layout-0 register bits, then layout-1 register bits from
        // the top, then (geometry 3) layout 2's register bit 3
        std::vector<int> order;
        for (int k : {8, 9, 10, 11, 7, 6, 5, 4, 3})
            if (k >= seg)
                order.push_back(k);
        for (size_t u = 0; u < upper.size(); u++)
        {
            unit_addr[order[u]] = int(upper[u]);
            used[upper[u]] = true;
        }
    }
    unsigned next = low;
    for (int f = seg; f < kUnitBits; f++)
    {
        if (unit_addr[f] >= 0)
            continue;
        while (used[next])
            next++;
        unit_addr[f] = int(next);
        used[next] = true;
    }
    pl.ebit.clear();
    if (c64)
        pl.ebit.push_back(0);
    for (int k = 0; k < kUnitBits; k++)
        pl.ebit.push_back(unsigned(unit_addr[k]));
    for (unsigned e : pl.ebit)
        if (e >= n)
            return false;
    pl.evec.clear();
    for (unsigned e : pl.ebit)
        pl.evec.push_back(uint64_t(1) << e);
    pl.slot.clear();
    for (const SingleGate &s : gs)
    {
        const int e = int(std::find(pl.ebit.begin(), pl.ebit.end(), s.target) - pl.ebit.begin());
        pl.slot.push_back(c64 ? (e == 0 ? kIntra : e - 1) : e);
        if (__builtin_popcountll(s.g->delta) > 1)
            pl.evec[size_t(e)] = s.g->delta;
    }
    return true;
}

// Host tables of the diagonal epilogue: the per-thread table (ST, G) and the
// packed input of tile_phase_kernel.
void epilogue_tables(int prec, unsigned n, const std::vector<unsigned> &pos, const Poly &poly,
                     std::vector<double> &ttab, std::vector<double> &ptab, int *nh_out, int *nb_out, int nthr = 8)
// pos: address bit of each tile element bit (Plan::ebit); nthr thread bits
// (element bits 0.. of the thread index, above the complex64 in-unit bit),
// the register bits above them
{
    const bool c64 = prec == C64;
    const int nE = int(pos.size());
    const int nb = c64 ? 5 : 4;
    const int nR = 1 << nb;
    const int nthreads = 1 << nthr;
    int thr_k[9], reg_k[5];
    for (int j = 0; j < nthr; j++)
        thr_k[j] = c64 ? j + 1 : j;
    for (int b = 0; b < nb; b++)
        reg_k[b] = c64 ? (b == 0 ? 0 : nthr + b) : nthr + b;
    auto Lt = [&](int k) { return poly.L[pos[k]]; };
    auto Qt = [&](int k, int l) { return poly.q(int(pos[k]), int(pos[l])); };

    // Thread t's factors exp(i s(t)), s(t) = sum_{j in t} L_j + sum_{j<l in
    // t} Q_jl, and exp(i d_b(t)), d_b(t) = sum_{j in t} Q(j, reg b), as
    // products over t's bits: t = t' + (highest bit h) multiplies t''s
    // entries by exp(i L_h) and exp(i Q_jh) for j in t' (resp. by
    // exp(i Q(h, reg b))).  A few complex products per entry instead of a
    // cosine and a sine; the rounding error grows by about 1e-16 per factor.
    using cd = std::complex<double>;
    auto cis = [](double a) { return cd(std::cos(a), std::sin(a)); };
    cd eL[9], eQ[9][9], eD[9][5];
    for (int j = 0; j < nthr; j++)
    {
        eL[j] = cis(Lt(thr_k[j]));
        for (int l = j + 1; l < nthr; l++)
            eQ[j][l] = cis(Qt(thr_k[j], thr_k[l]));
        for (int b = 0; b < nb; b++)
            eD[j][b] = cis(Qt(thr_k[j], reg_k[b]));
    }
    const int stride = 2 * (1 + nb);
    ttab.assign(size_t(nthreads) * stride, 0.0);
    for (int b = 0; b <= nb; b++)
        ttab[size_t(2 * b)] = 1.0;
    for (int t = 1; t < nthreads; t++)
    {
        const int h = 31 - __builtin_clz(unsigned(t));
        const int tp = t ^ (1 << h);
        const double *ap = ttab.data() + size_t(tp) * stride;
        double *a = ttab.data() + size_t(t) * stride;
        cd v = cd(ap[0], ap[1]) * eL[h];
        for (int j = 0; j < h; j++)
            if ((tp >> j) & 1)
                v *= eQ[j][h];
        a[0] = v.real();
        a[1] = v.imag();
        for (int b = 0; b < nb; b++)
        {
            const cd d = cd(ap[2 + 2 * b], ap[3 + 2 * b]) * eD[h][b];
            a[2 + 2 * b] = d.real();
            a[3 + 2 * b] = d.imag();
        }
    }

    std::vector<int> outer;
    for (unsigned b = 0; b < n; b++)
        if (std::find(pos.begin(), pos.end(), b) == pos.end())
            outer.push_back(int(b));
    const int nh = int(outer.size());
    ptab.assign(size_t(nE) * nh + nh + size_t(nh) * nh + nR + nthr + nb, 0.0);
    double *cq = ptab.data();
    double *hL = cq + nE * nh;
    double *hQ = hL + nh;
    double *rs = hQ + nh * nh;
    double *mp = rs + nR;
    for (int k = 0; k < nE; k++)
        for (int i = 0; i < nh; i++)
            cq[k * nh + i] = poly.q(int(pos[k]), outer[i]);
    for (int i = 0; i < nh; i++)
    {
        hL[i] = poly.L[outer[i]];
        for (int l = i + 1; l < nh; l++)
            hQ[i * nh + l] = poly.q(outer[i], outer[l]);
    }
    for (int rho = 0; rho < nR; rho++)
    {
        double s = 0.0;
        for (int b = 0; b < nb; b++)
        {
            if (!((rho >> b) & 1))
                continue;
            s += Lt(reg_k[b]);
            for (int c = b + 1; c < nb; c++)
                if ((rho >> c) & 1)
                    s += Qt(reg_k[b], reg_k[c]);
        }
        rs[rho] = s;
    }
    for (int j = 0; j < nthr; j++)
        mp[j] = thr_k[j];
    for (int b = 0; b < nb; b++)
        mp[nthr + b] = reg_k[b];
    *nh_out = nh;
    *nb_out = nb;
}

// Per-tile factors of the fans: out[tile * kFanStride + 2 s] = exp(i (c_s +
// sum_i theta_s[outer i] * tile bit i)), one thread per tile; tab holds, per
// slot s < kUnitBits + 1, c_s then the nh outer coefficients.
// spread = 1: one thread per (tile, slot) instead of one per tile (few tiles).
__device__ __forceinline__ void fan_phase_work(const double *__restrict__ tab, int nh, uint64_t ntiles,
                                               double *__restrict__ out, int spread, uint64_t id)
{
    const uint64_t tile = spread ? id / uint64_t(kUnitBits + 1) : id;
    if (tile >= ntiles)
        return;
    const int first = spread ? int(id % uint64_t(kUnitBits + 1)) : 0;
    const int last = spread ? first + 1 : kUnitBits + 1;
    for (int sl = first; sl < last; sl++)
    {
        const double *row = tab + size_t(sl) * (1 + nh);
        double ph = row[0];
        for (int i = 0; i < nh; i++)
            if ((tile >> i) & 1u)
                ph += row[1 + i];
        double sn, cs;
        sincos(ph, &sn, &cs);
        out[tile * kFanStride + 2 * sl] = cs;
        out[tile * kFanStride + 2 * sl + 1] = sn;
    }
}

__global__ __launch_bounds__(kThreads) void fan_phase_kernel(const double *__restrict__ tab, int nh, uint64_t ntiles,
                                                             double *__restrict__ out, int spread)
{
    fan_phase_work(tab, nh, ntiles, out, spread, blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x);
}

unsigned fan_phase_blocks(uint64_t ntiles, int *spread)
{
    *spread = ntiles < 4096 ? 1 : 0;
    const uint64_t threads = *spread ? ntiles * uint64_t(kUnitBits + 1) : ntiles;
    return unsigned((threads + kThreads - 1) / kThreads);
}

// Both per-tile factor kernels of a pass with an epilogue and fans in one
// launch: blocks below tblocks run tile_phase_kernel's work, the others
// fan_phase_kernel's (a block-uniform branch, so the epilogue's LDS staging
// and barrier see whole blocks).
__global__ __launch_bounds__(kThreads) void pass_phase_kernel(const double *__restrict__ ttab, int nE, int nh, int nb,
                                                              double K, uint64_t ntiles, double *__restrict__ tout,
                                                              int tspread, int nthr, unsigned tblocks,
                                                              const double *__restrict__ ftab, int fnh,
                                                              double *__restrict__ fout, int fspread)
{
    extern __shared__ double sh[];
    if (blockIdx.x < tblocks)
        tile_phase_work(ttab, nE, nh, nb, K, ntiles, tout, tspread, nthr,
                        blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x, sh);
    else
        fan_phase_work(ftab, fnh, ntiles, fout, fspread,
                       (blockIdx.x - tblocks) * static_cast<uint64_t>(blockDim.x) + threadIdx.x);
}

// The per-tile factor launches of a pass, recorded while its tables are
// packed (setup_epilogue, setup_fans) and made after the batch's flush: one
// pass_phase_kernel when the pass has both.
struct PhaseLaunches
{
    bool tile = false;
    const double *ttab = nullptr;
    size_t tdoubles = 0;
    int nE = 0, nh = 0, nb = 0, nthr = 8;
    double K = 0.0;
    uint64_t tntiles = 0;
    double *tout = nullptr;
    bool fan = false;
    const double *ftab = nullptr;
    int fnh = 0;
    uint64_t fntiles = 0;
    double *fout = nullptr;
};

int launch_phases(hipStream_t stream, const PhaseLaunches &ph)
{
    if (ph.tile && ph.fan && ph.tntiles == ph.fntiles)
    {
        int ts = 0, fs = 0;
        const unsigned tb = tile_phase_blocks(ph.tntiles, ph.nb, ph.nthr, &ts);
        const unsigned fb = fan_phase_blocks(ph.fntiles, &fs);
        hipLaunchKernelGGL(pass_phase_kernel, dim3(tb + fb), dim3(kThreads), ph.tdoubles * sizeof(double), stream,
                           ph.ttab, ph.nE, ph.nh, ph.nb, ph.K, ph.tntiles, ph.tout, ts, ph.nthr, tb, ph.ftab, ph.fnh,
                           ph.fout, fs);
        return hip_check(hipGetLastError(), "pass_phase_kernel launch");
    }
    if (ph.tile)
    {
        launch_tile_phase(stream, ph.ttab, ph.tdoubles, ph.nE, ph.nh, ph.nb, ph.K, ph.tntiles, ph.tout, ph.nthr);
        ZXA_TRY(hip_check(hipGetLastError(), "tile_phase_kernel launch"));
    }
    if (ph.fan)
    {
        int fs = 0;
        const unsigned fb = fan_phase_blocks(ph.fntiles, &fs);
        hipLaunchKernelGGL(fan_phase_kernel, dim3(fb), dim3(kThreads), 0, stream, ph.ftab, ph.fnh, ph.fntiles, ph.fout,
                           fs);
        ZXA_TRY(hip_check(hipGetLastError(), "fan_phase_kernel launch"));
    }
    return OK;
}

// A pass's tables in one copy: begin() opens the ring's batch, the per-tile
// factor launches that read its blocks wait in ph until run() flushes it; the
// destructor drops a batch that an error return left open.
struct PassBatch
{
    Ctx &ctx;
    PhaseLaunches ph;
    explicit PassBatch(Ctx &c) : ctx(c)
    {
    }
    ~PassBatch()
    {
        ctx.ring.end_batch();
    }
    int begin()
    {
        return ctx.ring.begin_batch();
    }
    int run()
    {
        ZXA_TRY(ctx.ring.flush(ctx.stream));
        const PhaseLaunches ph_now = ph;
        ph = PhaseLaunches();
        return launch_phases(ctx.stream, ph_now);
    }
};

// Stages the diagonal epilogue of a pass: the per-thread table and the
// per-tile phase factors (tile_phase_kernel into scratch_a, recorded in ph
// and launched after the batch's flush).
int setup_epilogue(Ctx &ctx, int prec, unsigned n, const std::vector<unsigned> &pos, const DiagGate *diag,
                   size_t ndiag, uint64_t ntiles, const double **ttab, const double **tphase, int nthr,
                   PhaseLaunches &ph)
{
    Poly poly;
    window_polynomial(n, diag, ndiag, poly);
    std::vector<double> tt, pt;
    int nh = 0, nb = 0;
    epilogue_tables(prec, n, pos, poly, tt, pt, &nh, &nb, nthr);
    void *dtt = nullptr, *dpt = nullptr;
    ZXA_TRY(ctx.ring.stage(tt.data(), tt.size() * sizeof(double), ctx.stream, &dtt));
    ZXA_TRY(ctx.ring.stage(pt.data(), pt.size() * sizeof(double), ctx.stream, &dpt));
    const size_t stride = 2 * (size_t(nthr) + (size_t(1) << nb));
    ZXA_TRY(ctx.scratch_a.ensure(ntiles * stride * sizeof(double)));
    ph.tile = true;
    ph.ttab = static_cast<const double *>(dpt);
    ph.tdoubles = pt.size();
    ph.nE = int(pos.size());
    ph.nh = nh;
    ph.nb = nb;
    ph.nthr = nthr;
    ph.K = poly.K;
    ph.tntiles = ntiles;
    ph.tout = static_cast<double *>(ctx.scratch_a.ptr);
    *ttab = static_cast<const double *>(dtt);
    *tphase = static_cast<const double *>(ctx.scratch_a.ptr);
    return OK;
}

// Kernel arguments and tables of a pass from its plan (targets' matrices in
// gs order; every other slot the identity).  use[L]: layout L has a target.
void fill_pass(int prec, unsigned n, const Plan &pl, TPass &P, TTab &G, bool use[3])
{
    const bool c64 = prec == C64;
    const std::vector<unsigned> &pos = pl.ebit;
    std::memset(&P, 0, sizeof(P));
    std::memset(&G, 0, sizeof(G));
    const unsigned shift = c64 ? 1u : 0u;
    for (int k = 0; k < kUnitBits; k++)
    {
        P.ub[k] = pos[k + shift] - shift;
        P.sorted[k] = P.ub[k];
        P.vec[k] = pl.evec[k + shift] >> shift;
        P.tmask |= uint64_t(1) << P.ub[k];
    }
    std::sort(P.sorted, P.sorted + kUnitBits);
    for (int r = 0; r < kRegs; r++)
        for (int i = 0; i < 4; i++)
            if ((r >> i) & 1)
                G.roff[r] ^= P.vec[reg_bit(0, i)];
    P.ntiles = uint64_t(1) << (n - shift - kUnitBits);
    for (TGate &tg : G.g)
    {
        tg.u[0] = tg.u[6] = 1.0;
        tg.uf[0] = tg.uf[6] = 1.0f;
    }
    use[0] = use[1] = use[2] = false;
    for (int slot : pl.slot)
        use[slot == kIntra ? 0 : layout_of_unit_bit(slot)] = true;
}

// Fan tables of a complex128 pass (single-bit slots, planned staged): each
// fan gate's thread and register factors in its layout, and the per-tile
// factors over the outer bits (fan_phase_kernel into scratch_g, recorded in
// ph and launched after the batch's flush).
int setup_fans(Ctx &ctx, unsigned n, const std::vector<SingleGate> &gs, const GateFan *fans, const Plan &pl,
               TPass &P, TTab &G, PhaseLaunches &ph)
{
    std::vector<int> outer;
    for (unsigned b = 0; b < n; b++)
        if (std::find(pl.ebit.begin(), pl.ebit.end(), b) == pl.ebit.end())
            outer.push_back(int(b));
    const int nh = int(outer.size());
    std::vector<double> tab(size_t(kUnitBits + 1) * (1 + nh), 0.0);
    auto cis = [](double a, double *o) {
        o[0] = std::cos(a);
        o[1] = std::sin(a);
    };
    for (size_t gi = 0; gi < gs.size(); gi++)
    {
        if (!fans[gi].on)
            continue;
        const int k = pl.slot[gi];
        const int lay = layout_of_unit_bit(k);
        int own = -1;
        for (int i = 0; i < 4; i++)
            if (reg_bit(lay, i) == k)
                own = i;
        const std::vector<double> &th = fans[gi].theta;
        TFan &f = G.fan[k];
        for (int h = 0; h < 2; h++)
            for (int e = 0; e < 16; e++)
            {
                double a = 0.0;
                for (int j = 0; j < 4; j++)
                    if ((e >> j) & 1)
                        a += th[P.ub[thr_bit(lay, 4 * h + j)]];
                cis(a, f.nib[16 * h + e]);
            }
        for (int r = 0; r < kRegs; r++)
        {
            double a = 0.0;
            for (int i = 0; i < 4; i++)
                if (i != own && ((r >> i) & 1))
                    a += th[P.ub[reg_bit(lay, i)]];
            cis(a, f.reg[r]);
        }
        G.g[k].fan = 1u;
        double *row = tab.data() + size_t(k) * (1 + nh);
        row[0] = fans[gi].c;
        for (int i = 0; i < nh; i++)
            row[1 + i] = th[size_t(outer[i])];
    }
    void *dt = nullptr;
    ZXA_TRY(ctx.ring.stage(tab.data(), tab.size() * sizeof(double), ctx.stream, &dt));
    ZXA_TRY(ctx.scratch_g.ensure(P.ntiles * kFanStride * sizeof(double)));
    ph.fan = true;
    ph.ftab = static_cast<const double *>(dt);
    ph.fnh = nh;
    ph.fntiles = P.ntiles;
    ph.fout = static_cast<double *>(ctx.scratch_g.ptr);
    P.fphase = static_cast<const double *>(ctx.scratch_g.ptr);
    return OK;
}

// A window's selector forms reduced to a basis (see TWin): fully reduced
// rows, row k the only one with bit piv[k] set, so a form's coordinate k is
// its bit piv[k].  False when the forms span more than kWinBits dimensions
// or reach above the state.
struct WinBasis
{
    int rank = 0;
    uint64_t a[kWinBits] = {};
    int piv[kWinBits] = {};
};

bool win_basis(unsigned n, const DiagGate *gates, size_t ngates, WinBasis &B)
{
    const uint64_t amask = (uint64_t(1) << 63) - 1;
    B = WinBasis();
    for (size_t i = 0; i < ngates; i++)
        for (int k = 0; k < 2; k++)
        {
            const uint64_t s = k == 0 ? gates[i].s0 : gates[i].s1;
            if (k == 1 && s == kNoSel)
                continue;
            uint64_t v = s & amask;
            if (n < 64 && (v >> n) != 0)
                return false;
            for (int r = 0; r < B.rank; r++)
                if ((v >> B.piv[r]) & 1)
                    v ^= B.a[r];
            if (v == 0)
                continue;
            if (B.rank == kWinBits)
                return false;
            const int p = 63 - __builtin_clzll(v);
            for (int r = 0; r < B.rank; r++)
                if ((B.a[r] >> p) & 1)
                    B.a[r] ^= v;
            B.a[B.rank] = v;
            B.piv[B.rank] = p;
            B.rank++;
        }
    return true;
}

// The window's phase table over y (2^rank entries of (cos, sin)): theta(y)
// sums each gate's angle where its selectors hold, in gate order.
void win_table(const WinBasis &B, const DiagGate *gates, size_t ngates, std::vector<double> &tab)
{
    const uint64_t amask = (uint64_t(1) << 63) - 1;
    auto coord = [&](uint64_t s) {
        uint32_t c = 0;
        for (int r = 0; r < B.rank; r++)
            c |= uint32_t(((s & amask) >> B.piv[r]) & 1) << r;
        return c;
    };
    std::vector<uint32_t> c0(ngates), c1(ngates);
    for (size_t i = 0; i < ngates; i++)
    {
        c0[i] = coord(gates[i].s0);
        c1[i] = gates[i].s1 == kNoSel ? 0 : coord(gates[i].s1);
    }
    const uint32_t nent = uint32_t(1) << B.rank;
    tab.assign(2 * size_t(nent), 0.0);
    for (uint32_t y = 0; y < nent; y++)
    {
        double th = 0.0;
        for (size_t i = 0; i < ngates; i++)
        {
            const bool on0 = ((__builtin_popcount(c0[i] & y) & 1) ^ int(gates[i].s0 >> 63)) != 0;
            const bool on1 = gates[i].s1 == kNoSel ||
                             ((__builtin_popcount(c1[i] & y) & 1) ^ int(gates[i].s1 >> 63)) != 0;
            if (on0 && on1)
                th += gates[i].theta;
        }
        tab[2 * size_t(y)] = std::cos(th);
        tab[2 * size_t(y) + 1] = std::sin(th);
    }
}

// The wide (13-unit-bit) tile: single-bit gates (alpha == delta) on distinct
// targets, all above the 16-unit contiguous run, at most nine; a state of
// at least 13 unit bits.  Slots take the targets' unit bits and, where fewer
// than nine, the lowest free unit bits above the run (which lengthen the
// runs), ascending.  ebit: the address bit of each tile element bit (as
// Plan::ebit) for the epilogue tables.
bool wide13_plan(int prec, unsigned n, const std::vector<SingleGate> &gs, P13 &P, Tab13 &G,
                 std::vector<unsigned> &ebit)
{
    const bool c64 = prec == C64;
    const unsigned shift = c64 ? 1u : 0u;
    if (n < unsigned(kW13Bits) + shift || n > 40 || gs.size() > size_t(kW13Slots))
        return false;
    const unsigned nu = n - shift;
    std::vector<unsigned> S;
    for (const SingleGate &g : gs)
    {
        const uint64_t alpha = g.g->packed_sel & ((uint64_t(1) << 63) - 1);
        if (g.g->delta != (uint64_t(1) << g.target) || alpha != g.g->delta || g.target < 4 + shift)
            return false;
        const unsigned u = g.target - shift;
        if (std::find(S.begin(), S.end(), u) != S.end())
            return false;
        S.push_back(u);
    }
    for (unsigned u = 4; u < nu && S.size() < size_t(kW13Slots); u++)
        if (std::find(S.begin(), S.end(), u) == S.end())
            S.push_back(u);
    if (S.size() != size_t(kW13Slots))
        return false;
    std::sort(S.begin(), S.end());
    uint32_t ub[kW13Bits];
    for (int u = 0; u < 4; u++)
        ub[u] = uint32_t(u);
    for (int k = 0; k < kW13Slots; k++)
        ub[4 + k] = S[size_t(k)];
    std::memset(&P, 0, sizeof(P));
    std::memset(&G, 0, sizeof(G));
    for (int k = 0; k < kW13Bits; k++)
    {
        P.sorted[k] = ub[k];
        P.tmask |= uint64_t(1) << ub[k];
    }
    std::sort(P.sorted, P.sorted + kW13Bits);
    P.ntiles = uint64_t(1) << (nu - unsigned(kW13Bits));
    for (int j = 0; j < 9; j++)
        G.toffv[j] = uint64_t(1) << ub[j];
    for (int r = 0; r < kRegs; r++)
        for (int i = 0; i < 4; i++)
            if ((r >> i) & 1)
                G.roff[r] ^= uint64_t(1) << ub[9 + i];
    for (TGate &tg : G.g)
    {
        tg.u[0] = tg.u[6] = 1.0;
        tg.uf[0] = tg.uf[6] = 1.0f;
    }
    for (const SingleGate &g : gs)
    {
        const unsigned u = g.target - shift;
        const int slot = int(std::find(S.begin(), S.end(), u) - S.begin());
        const double *m = g.g->u;
        // beta = 1: the pair's first element is the target-bit-one amplitude
        const bool swap = (g.g->packed_sel >> 63) != 0;
        TGate &tg = G.g[slot];
        for (int e = 0; e < 8; e++)
        {
            tg.u[e] = swap ? m[e ^ 6] : m[e];
            tg.uf[e] = float(tg.u[e]);
        }
    }
    ebit.clear();
    if (c64)
        ebit.push_back(0);
    for (int u = 0; u < kW13Bits; u++)
        ebit.push_back(ub[u] + shift);
    return true;
}

template <bool kC64, bool kDiag>
int launch_wide13(Ctx &ctx, void *state, const P13 &P0, const Tab13 &tab, const double *ttab, const double *tphase)
{
    auto k = wide13_kernel<kC64, kDiag>;
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, kW13Threads, 0) != hipSuccess || per_cu <= 0)
        per_cu = 1;
    const uint64_t cap = uint64_t(per_cu) * uint64_t(ctx.num_cus > 0 ? ctx.num_cus : 256);
    const uint64_t grid = std::max<uint64_t>(1, std::min<uint64_t>(P0.ntiles, cap));
    P13 P = P0;
    P.gstep = deposit13(grid, P.sorted);
    hipLaunchKernelGGL(k, dim3(unsigned(grid)), dim3(kW13Threads), 0, ctx.stream, tab, state, P, ttab, tphase);
    return hip_check(hipGetLastError(), "wide13_kernel launch");
}

// One wide tile pass, then the window (a phase polynomial) as its epilogue.
int run_wide13_pass(Ctx &ctx, void *state, int prec, unsigned n, const std::vector<SingleGate> &gs,
                    const DiagGate *diag, size_t ndiag)
{
    const bool c64 = prec == C64;
    P13 P;
    Tab13 G;
    std::vector<unsigned> ebit;
    if (!wide13_plan(prec, n, gs, P, G, ebit))
    {
        set_error("wide tile pass: the gates do not fit the wide tile");
        return EINVAL_;
    }
    const double *ttab = nullptr;
    const double *tphase = nullptr;
    // the gate table goes with the launch; only an epilogue stages tables
    PassBatch batch(ctx);
    if (ndiag > 0)
    {
        if (!window_is_polynomial(n, diag, ndiag))
        {
            set_error("wide tile pass: the window is not a phase polynomial");
            return EINVAL_;
        }
        ZXA_TRY(batch.begin());
        ZXA_TRY(setup_epilogue(ctx, prec, n, ebit, diag, ndiag, P.ntiles, &ttab, &tphase, 9, batch.ph));
        ZXA_TRY(batch.run());
    }
    const Tab13 &tab = G;
    const uint64_t bytes = 2ull * (uint64_t(1) << n) * elem_bytes(prec);
    ZXA_TRY(ctx.tic());
    int rc;
    if (c64)
        rc = ndiag ? launch_wide13<true, true>(ctx, state, P, tab, ttab, tphase)
                   : launch_wide13<true, false>(ctx, state, P, tab, ttab, tphase);
    else
        rc = ndiag ? launch_wide13<false, true>(ctx, state, P, tab, ttab, tphase)
                   : launch_wide13<false, false>(ctx, state, P, tab, ttab, tphase);
    ZXA_TRY(rc);
    ZXA_TRY(ctx.toc(ndiag > 0 ? F_FUSED : F_U3TILE, bytes));
    return ctx.ring.done(ctx.stream);
}

// One tile pass over the given targets, then the diagonal window (if any).
// fresh: the state is the reset |0...0> of 2^n amplitudes, not yet written
// (see tile_fresh_kernel; passes with per-pair selectors reset it first).
int run_tile_pass(Ctx &ctx, void *state, int prec, unsigned n, const std::vector<SingleGate> &gs,
                  const DiagGate *diag = nullptr, size_t ndiag = 0, int geometry = -1,
                  const GateFan *fans = nullptr, bool fresh = false)
{
    const bool c64 = prec == C64;
    bool any_fan = false;
    if (fans)
        for (size_t gi = 0; gi < gs.size(); gi++)
            any_fan = any_fan || fans[gi].on;
    if (any_fan)
    {
        if (c64)
        {
            set_error("tile pass: fans need a complex128 state");
            return EINVAL_;
        }
        for (size_t gi = 0; gi < gs.size(); gi++)
        {
            const uint64_t alpha = gs[gi].g->packed_sel & ((uint64_t(1) << 63) - 1);
            if (alpha != gs[gi].g->delta || (gs[gi].g->delta & (gs[gi].g->delta - 1)) != 0)
            {
                set_error("tile pass: fans need single-bit selectors equal to the pair offsets");
                return EINVAL_;
            }
            if (fans[gi].on && fans[gi].theta.size() != n)
            {
                set_error("tile pass: a fan needs one coefficient per address bit");
                return EINVAL_;
            }
        }
    }
    Plan pl;
    if (!plan_tile(prec, n, gs, pl, geometry, any_fan))
    {
        set_error(any_fan ? "tile pass: the gates of a pass with fans do not fit the kernel's order"
                          : "tile pass: state narrower than a tile or too many targets");
        return EINVAL_;
    }
    const std::vector<unsigned> &pos = pl.ebit;
    TPass P;
    TTab G;
    bool use[3];
    fill_pass(prec, n, pl, P, G, use);
    const unsigned shift = c64 ? 1u : 0u; // element bit -> unit-address bit
    bool flip = false;
    // the side of a pair's first element is the selector's parity over that
    // element's address.  With single-bit slot offsets the target bit is
    // clear there and drops out of the selector; a wider offset may cover
    // other gates' target bits, so then every gate keeps its whole selector
    bool wide = false;
    for (const SingleGate &s : gs)
        wide = wide || __builtin_popcountll(s.g->delta) > 1;
    bool any_rest = false;
    for (const SingleGate &s : gs)
    {
        const uint64_t alpha = s.g->packed_sel & ((uint64_t(1) << 63) - 1);
        any_rest = any_rest || (wide ? alpha : alpha & ~s.g->delta) != 0;
    }
    if (any_rest)
    {
        // Filler unit bits (neither contiguous nor a slot) take the offsets
        // of the gates whose selector sees them: v ^= sum_j [alpha_j . v]
        // delta_j leaves the tile's span (so the tiles) unchanged and makes
        // alpha_k . v = 0 for every gate k (alpha_k . delta_j = [k == j]),
        // so a gate's side no longer changes across those bits; where it
        // changes across none of a layout's register bits the kernel picks
        // the matrix once per thread instead of once per pair.
        const int seg = int(tile_seg(prec, geometry));
        std::vector<bool> is_slot(kUnitBits, false);
        for (int sl : pl.slot)
            if (sl != kIntra)
                is_slot[size_t(sl)] = true;
        for (int k = seg; k < kUnitBits; k++)
        {
            if (is_slot[size_t(k)])
                continue;
            const uint64_t v0 = P.vec[k];
            uint64_t v = v0;
            for (size_t gi = 0; gi < gs.size(); gi++)
            {
                if (pl.slot[gi] == kIntra)
                    continue;
                const uint64_t au = (gs[gi].g->packed_sel & ((uint64_t(1) << 63) - 1)) >> shift;
                if (__builtin_popcountll(au & v0) & 1)
                    v ^= gs[gi].g->delta >> shift;
            }
            P.vec[k] = v;
        }
        for (int r = 0; r < kRegs; r++)
        {
            G.roff[r] = 0;
            for (int i = 0; i < 4; i++)
                if ((r >> i) & 1)
                    G.roff[r] ^= P.vec[reg_bit(0, i)];
        }
    }
    for (size_t gi = 0; gi < gs.size(); gi++)
    {
        const SingleGate &s = gs[gi];
        const int slot = pl.slot[gi];
        TGate &tg = G.g[slot];
        const uint64_t alpha = s.g->packed_sel & ((uint64_t(1) << 63) - 1);
        // filler offsets may now cover target bits too: whole selectors
        const uint64_t rest = any_rest ? alpha : alpha & ~s.g->delta;
        uint32_t beta = uint32_t(s.g->packed_sel >> 63);
        const double *u = s.g->u;
        double m[8];
        if (rest == 0 && beta)
        {
            // the even side is always the target-bit-one amplitude: swap roles
            const double sw[8] = {u[6], u[7], u[4], u[5], u[2], u[3], u[0], u[1]};
            std::memcpy(m, sw, sizeof(m));
            beta = 0;
        }
        else
        {
            std::memcpy(m, u, sizeof(m));
        }
        for (int i = 0; i < 8; i++)
        {
            tg.u[i] = m[i];
            tg.uf[i] = float(m[i]);
        }
        tg.amask = rest >> shift;
        tg.a0 = c64 ? uint32_t(rest & 1u) : 0u;
        tg.beta = beta;
        const int lay = slot == kIntra ? 0 : layout_of_unit_bit(slot);
        // only the registers holding a pair's first element (the gate's own
        // register bit clear) are read, so rpar == 0 when the side is the
        // same across all of them
        int own = -1;
        for (int i = 0; i < 4; i++)
            if (slot != kIntra && reg_bit(lay, i) == slot)
                own = i;
        for (int r = 0; r < kRegs; r++)
        {
            if (own >= 0 && ((r >> own) & 1))
                continue;
            uint64_t off = 0;
            for (int i = 0; i < 4; i++)
                if ((r >> i) & 1)
                    off ^= P.vec[reg_bit(lay, i)];
            tg.rpar |= uint32_t(__builtin_popcountll(off & tg.amask) & 1) << r;
        }
        if (rest != 0)
            flip = true;
    }
    const double *ttab = nullptr;
    const double *tphase = nullptr;
    // the window: a phase polynomial on a pass without per-pair selectors
    // (and with every fan), else a table over its selector forms
    bool win = false;
    std::vector<double> wtab;
    std::memset(&G.win, 0, sizeof(G.win));
    PassBatch batch(ctx);
    ZXA_TRY(batch.begin());
    if (ndiag > 0)
    {
        if (!flip && !wide && (any_fan || window_is_polynomial(n, diag, ndiag)))
        {
            ZXA_TRY(setup_epilogue(ctx, prec, n, pos, diag, ndiag, P.ntiles, &ttab, &tphase, 8, batch.ph));
        }
        else
        {
            WinBasis B;
            if (any_fan || !win_basis(n, diag, ndiag, B))
            {
                set_error("tile pass: the diagonal window is neither a phase polynomial nor tabulable");
                return EINVAL_;
            }
            win_table(B, diag, ndiag, wtab);
            win = true;
            G.win.rank = uint32_t(B.rank);
            for (int k = 0; k < B.rank; k++)
            {
                G.win.a[k] = B.a[k] >> shift;
                if (c64)
                    G.win.yin |= uint32_t(B.a[k] & 1u) << k;
            }
        }
    }
    if (any_fan)
        ZXA_TRY(setup_fans(ctx, n, gs, fans, pl, P, G, batch.ph));
    // passes whose targets all sit on unit bits 6..11 run on upper_kernel (one
    // tile per workgroup, no LDS) when their highest target unit-address bit
    // is at least tile.upper.<prec>: spread over more than 2^k units, a tile's
    // runs are fetched faster with more tiles in flight per CU than the LDS
    // allows the tile kernel
    bool upper = !fresh && !any_fan && P.ntiles <= uint64_t(INT_MAX);
    unsigned high = 0;
    for (size_t gi = 0; gi < gs.size() && upper; gi++)
    {
        const int slot = pl.slot[gi];
        if (slot == kIntra)
            continue;
        upper = slot >= 6;
        high = std::max(high, P.ub[slot]);
    }
    upper = upper && high >= unsigned(tune_get(c64 ? "tile.upper.c64" : "tile.upper.c128"));
    if (upper && flip)
    {
        // the upper kernel applies slots 6 and 7 with registers = unit bits
        // 6, 7, 10, 11: their parity masks over those register offsets
        const uint64_t sv[4] = {P.vec[6], P.vec[7], P.vec[10], P.vec[11]};
        for (int own = 0; own < 2; own++)
        {
            TGate &tg = G.g[6 + own];
            tg.rpar = 0;
            for (int r = 0; r < kRegs; r++)
            {
                if ((r >> own) & 1)
                    continue;
                uint64_t off = 0;
                for (int i = 0; i < 4; i++)
                    if ((r >> i) & 1)
                        off ^= sv[i];
                tg.rpar |= uint32_t(__builtin_popcountll(off & tg.amask) & 1) << r;
            }
        }
    }
    if (win)
    {
        // y of every register offset, over the final (filler-cleaned) offsets
        for (int r = 0; r < kRegs; r++)
        {
            uint32_t y = 0;
            for (int k = 0; k < kWinBits; k++)
                y |= uint32_t(__builtin_popcountll(G.roff[r] & G.win.a[k]) & 1) << k;
            G.win.yr[r] = y;
        }
        void *dw = nullptr;
        ZXA_TRY(ctx.ring.stage(wtab.data(), wtab.size() * sizeof(double), ctx.stream, &dw));
        ttab = static_cast<const double *>(dw);
    }
    void *dtab = nullptr;
    ZXA_TRY(ctx.ring.stage(&G, sizeof(G), ctx.stream, &dtab));
    const TTab *tab = static_cast<const TTab *>(dtab);
    ZXA_TRY(batch.run());
    const bool fresh_k = fresh && !flip;
    if (fresh && !fresh_k)
        ZXA_TRY(state_reset(ctx, state, prec, size_t(1) << n));
    // algorithmic bytes: the state read and written, or written only
    const uint64_t bytes = (fresh_k ? 1ull : 2ull) * (uint64_t(1) << n) * elem_bytes(prec);
    ZXA_TRY(ctx.tic());
    if (fresh_k)
    {
        if (c64)
            ZXA_TRY(launch_fresh<true>(ctx, state, P, use[1], use[2], tab, ttab, tphase, win, any_fan));
        else
            ZXA_TRY(launch_fresh<false>(ctx, state, P, use[1], use[2], tab, ttab, tphase, win, any_fan));
    }
    else if (upper)
        ZXA_TRY(launch_u(ctx, c64, state, P, tab, ttab, tphase, flip, win));
    else if (c64)
        ZXA_TRY(launch_p<true>(ctx, state, P, use[1], use[2], flip, tab, ttab, tphase, false, win));
    else
        ZXA_TRY(launch_p<false>(ctx, state, P, use[1], use[2], flip, tab, ttab, tphase, any_fan, win));
    ZXA_TRY(ctx.toc(ndiag > 0 || any_fan ? F_FUSED : F_U3TILE, bytes));
    if (upper && ctx.timing)
        ctx.fam[F_UPPER].launches++;
    return ctx.ring.done(ctx.stream);
}

template <bool kC64, bool kL1, bool kL2, bool kDiag>
int launch_adj_t(Ctx &ctx, void *psi, void *lam, const TPass &P0, const TTab *tab, const double *ttab,
                 const double *tphase, double *partial, int *nwaves)
{
    auto k = adj_kernel<kC64, kL1, kL2, kDiag>;
    const int grid = persistent_grid(ctx, k, P0.ntiles);
    // the partial buffer holds *nwaves (on entry: its capacity in waves)
    if (grid * (kThreads / 64) > *nwaves)
    {
        set_error("adjoint pass: partial buffer too small for the grid");
        return EINVAL_;
    }
    TPass P = P0;
    P.gstep = deposit(uint64_t(grid), P.sorted);
    *nwaves = grid * (kThreads / 64);
    hipLaunchKernelGGL(k, dim3(grid), dim3(kThreads), 0, ctx.stream, psi, lam, P, tab, ttab, tphase, partial);
    return hip_check(hipGetLastError(), "adj_kernel launch");
}

template <bool kC64, bool kDiag>
int launch_adj_l(Ctx &ctx, void *psi, void *lam, const TPass &P, bool l1, bool l2, const TTab *tab,
                 const double *ttab, const double *tphase, double *partial, int *nwaves)
{
    if (l1 && l2)
        return launch_adj_t<kC64, true, true, kDiag>(ctx, psi, lam, P, tab, ttab, tphase, partial, nwaves);
    if (l1)
        return launch_adj_t<kC64, true, false, kDiag>(ctx, psi, lam, P, tab, ttab, tphase, partial, nwaves);
    if (l2)
        return launch_adj_t<kC64, false, true, kDiag>(ctx, psi, lam, P, tab, ttab, tphase, partial, nwaves);
    return launch_adj_t<kC64, false, false, kDiag>(ctx, psi, lam, P, tab, ttab, tphase, partial, nwaves);
}

// Gradient entries of one target in slot `slot` (out: kAccs entries).  The
// operators must be traceless and Hermitian (rotation generators conjugated
// by unitaries); a = (A00 - A11) / 2, b = (A01 + conj(A10)) / 2.
int grad_entries(bool c64, int slot, const AdjTarget &tg, size_t nparams, TGrad *out)
{
    // the slot's layout and list position
    int L = 0, pos = 4 * kSums;
    if (slot != kIntra)
        for (int l = 0; l < 3; l++)
            for (int i = 0; i < 4; i++)
                if (reg_bit(l, i) == slot)
                {
                    L = l;
                    pos = i * kSums;
                }
    for (int o = 0; o < kObs; o++)
    {
        TGrad &g = out[slot * kObs + o];
        const double *A = tg.a[o];
        g.param = tg.param[o];
        for (int j = 0; j < kSums; j++)
            g.lane[j] = sum_lane(c64, L, pos + j);
        g.c[0] = 0.5 * (A[0] - A[6]);
        g.c[1] = 0.5 * (A[2] + A[4]);
        g.c[2] = 0.5 * (A[3] - A[5]);
        if (g.param < 0)
            continue;
        if (size_t(g.param) >= nparams)
        {
            set_error("adjoint pass: parameter index outside the gradient array");
            return EINVAL_;
        }
        double scale = 1.0;
        for (int k = 0; k < 8; k++)
            scale = std::max(scale, std::fabs(A[k]));
        const double off = std::max({std::fabs(A[0] + A[6]), std::fabs(A[1]), std::fabs(A[7]), std::fabs(A[2] - A[4]),
                                     std::fabs(A[3] + A[5])});
        if (off > 1e-9 * scale)
        {
            set_error("adjoint pass: transition operator is not traceless Hermitian");
            return EINVAL_;
        }
    }
    return OK;
}

void clear_grads(TGrad *g)
{
    for (int a = 0; a < kAccs; a++)
    {
        std::memset(&g[a], 0, sizeof(TGrad));
        g[a].param = -1;
    }
}

// ---------------------------------------------------------------------------
// Batched tile passes: one pass of many independent states (different widths
// and tables) in one persistent launch.  Work item = one tile of one state;
// the launch's items are the states' tile ranges, concatenated.  Every tile
// goes through all three layouts (identity matrices where a state's pass has
// no target there) and the epilogue when the state's pass has a window.
// ---------------------------------------------------------------------------
struct BItem
{
    uint64_t first;   // first global tile index of this state's pass
    uint64_t ntiles;
    uint64_t psi;     // state (16-byte units)
    uint64_t lam;     // adjoint: second state
    uint64_t tab;     // TTab
    uint64_t ttab;    // epilogue per-thread table (the identity table: no window)
    uint64_t tphase;  // epilogue per-tile factors (one identity row: no window)
    uint64_t tstride; // doubles per tile in tphase; 0: no window
    uint64_t partial; // adjoint: per tile and wave, a row of 64 lane sums
    uint32_t ub[kUnitBits];
    uint32_t sorted[kUnitBits];
};

// The item holding global tile g: binary search over the items' first-tile
// indices (scalar loads).
__device__ __forceinline__ int batch_find(ZXA_CONST BItem *IT, int nitems, uint64_t g)
{
    int lo = 0, hi = nitems - 1;
    while (lo < hi)
    {
        const int mid = (lo + hi + 1) >> 1;
        if (IT[mid].first <= g)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ void batch_pass(ZXA_CONST BItem &I, TPass &P)
{
#pragma unroll
    for (int k = 0; k < kUnitBits; k++)
    {
        P.ub[k] = I.ub[k];
        P.sorted[k] = I.sorted[k];
        P.vec[k] = uint64_t(1) << I.ub[k];
    }
}

template <bool kC64>
__device__ __forceinline__ void batch_load(typename UnitT<kC64>::T (&v)[kRegs], ZXA_CONST BItem &I, uint64_t g, int t)
{
    using T = typename UnitT<kC64>::T;
    TPass P;
    batch_pass(I, P);
    const uint64_t base = deposit(g - I.first, P.sorted);
    load_tile<kC64>(v, reinterpret_cast<const T *>(I.psi), base ^ thr_off<0>(t, P), (ZXA_CONST TTab *)I.tab);
}

template <bool kC64>
__device__ __forceinline__ void batch_store(const typename UnitT<kC64>::T (&v)[kRegs], ZXA_CONST BItem &I, uint64_t g,
                                            int t)
{
    using T = typename UnitT<kC64>::T;
    TPass P;
    batch_pass(I, P);
    const uint64_t base = deposit(g - I.first, P.sorted);
    store_tile<kC64>(v, reinterpret_cast<T *>(I.psi), base ^ thr_off<0>(t, P), (ZXA_CONST TTab *)I.tab);
}

// Gates and transposes of one forward tile, then the item's window (tphase
// set) with the per-thread table e loaded before the tile's work.
template <bool kC64, bool kL1, bool kL2>
__device__ __forceinline__ void batch_work(typename UnitT<kC64>::T (&v)[kRegs], typename UnitT<kC64>::T *lds,
                                           ZXA_CONST BItem &I, uint64_t g, int t, const Epi<kC64> &e)
{
    TPass P;
    batch_pass(I, P);
    const uint64_t local = g - I.first;
    const uint64_t base = deposit(local, P.sorted);
    const uint64_t toff = thr_off<0>(t, P);
    ZXA_CONST TTab *G = (ZXA_CONST TTab *)I.tab;
    apply_layout<kC64, false, 0>(v, P, G, base, toff);
    if constexpr (kL1)
    {
        transpose<kC64, 0, 1>(v, lds, t);
        apply_layout<kC64, false, 1>(v, P, G, base, toff);
        if constexpr (kL2)
        {
            transpose<kC64, 1, 2>(v, lds, t);
            apply_layout<kC64, false, 2>(v, P, G, base, toff);
            transpose<kC64, 2, 0>(v, lds, t);
        }
        else
        {
            transpose<kC64, 1, 0>(v, lds, t);
        }
    }
    else if constexpr (kL2)
    {
        transpose<kC64, 0, 2>(v, lds, t);
        apply_layout<kC64, false, 2>(v, P, G, base, toff);
        transpose<kC64, 2, 0>(v, lds, t);
    }
    if (I.tstride != 0)
        e.apply(v, (ZXA_CONST double *)(reinterpret_cast<const double *>(I.tphase) + local * Epi<kC64>::kStride), t);
}

// Forward passes of many states, pipelined as tile_kernel: two register
// buffers, the loads of the workgroup's next tile in flight while the
// current one is worked on, its store issued after those loads.  The next
// tile may belong to another item.  Every item has a valid per-thread
// epilogue table (the identity where it has no window), so that table's
// loads are unconditional and issued before the prefetch: waiting for them
// never waits for the prefetched tile.  Past the workgroup's last tile the
// prefetch re-reads the current one.
template <bool kC64, bool kL1, bool kL2>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(2))) void batch_fwd_kernel(
    const BItem *__restrict__ items, int nitems, uint64_t total)
{
    using T = typename UnitT<kC64>::T;
    __shared__ T lds[(kL1 || kL2) ? kUnits : 1];
    const int t = threadIdx.x;
    const uint64_t grid = gridDim.x;
    const BItem *I0 = items;
    asm volatile("" : "+s"(I0));
    ZXA_CONST BItem *IT = (ZXA_CONST BItem *)I0;
    T a[kRegs], b[kRegs];
    Epi<kC64> e;
    uint64_t g0 = blockIdx.x;
    int i0 = batch_find(IT, nitems, g0);
    batch_load<kC64>(a, IT[i0], g0, t);
    e.init(reinterpret_cast<const double *>(IT[i0].ttab), t);
    uint64_t g1 = g0 + grid;
    bool m1 = g1 < total;
    int i1 = m1 ? batch_find(IT, nitems, g1) : i0;
    uint64_t h1 = m1 ? g1 : g0;
    batch_load<kC64>(b, IT[i1], h1, t);
    batch_work<kC64, kL1, kL2>(a, lds, IT[i0], g0, t, e);
    for (;;)
    {
        batch_store<kC64>(a, IT[i0], g0, t);
        if (!m1)
            break;
        const uint64_t g2 = g1 + grid;
        const bool m2 = g2 < total;
        const int i2 = m2 ? batch_find(IT, nitems, g2) : i1;
        const uint64_t h2 = m2 ? g2 : g1;
        e.init(reinterpret_cast<const double *>(IT[i1].ttab), t);
        batch_load<kC64>(a, IT[i2], h2, t);
        batch_work<kC64, kL1, kL2>(b, lds, IT[i1], g1, t, e);

        batch_store<kC64>(b, IT[i1], g1, t);
        if (!m2)
            break;
        const uint64_t g3 = g2 + grid;
        const bool m3 = g3 < total;
        const int i3 = m3 ? batch_find(IT, nitems, g3) : i2;
        const uint64_t h3 = m3 ? g3 : g2;
        e.init(reinterpret_cast<const double *>(IT[i2].ttab), t);
        batch_load<kC64>(b, IT[i3], h3, t);
        batch_work<kC64, kL1, kL2>(a, lds, IT[i2], g2, t, e);

        g0 = g2;
        i0 = i2;
        g1 = g3;
        i1 = i3;
        m1 = m3;
    }
}

// Adjoint passes of many states, one tile per step (two states per tile
// leave no registers for a second buffer).  kL1 / kL2: the launch's items use
// layout 1 / 2 (any of them); items that do not get identity matrices there.
template <bool kC64, bool kL1, bool kL2>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(2))) void batch_adj_kernel(
    const BItem *__restrict__ items, int nitems, uint64_t total)
{
    using T = typename UnitT<kC64>::T;
    __shared__ T lds[(kL1 || kL2) ? kUnits : 1];
    const int t = threadIdx.x;
    const int lane = t & 63;
    const BItem *I0 = items;
    asm volatile("" : "+s"(I0));
    ZXA_CONST BItem *IT = (ZXA_CONST BItem *)I0;
    for (uint64_t g = blockIdx.x; g < total; g += gridDim.x)
    {
        ZXA_CONST BItem &I = IT[batch_find(IT, nitems, g)];
        TPass P;
        batch_pass(I, P);
        const uint64_t local = g - I.first;
        const uint64_t base = deposit(local, P.sorted);
        const uint64_t toff = thr_off<0>(t, P);
        ZXA_CONST TTab *G = (ZXA_CONST TTab *)I.tab;
        T *pp = reinterpret_cast<T *>(I.psi);
        T *pl = reinterpret_cast<T *>(I.lam);
        T v[kRegs], w[kRegs];
        load_tile<kC64>(v, pp, base ^ toff, G);
        load_tile<kC64>(w, pl, base ^ toff, G);
        AdjRest<kC64> R;
        adj_tile<kC64, kL1, kL2>(v, w, lds, G, t, R);
        // unconditional (identity tables where the item has no window): a
        // branch around the epilogue made the compiler spill about 90 VGPRs.
        // The per-thread factors are read here, not held across the tile.
        const double *et = reinterpret_cast<const double *>(I.ttab);
        asm volatile("" : "+s"(et));
        Epi<kC64> epi;
        epi.init(et, t);
        epi.template apply_n<true>(
            v, w, (ZXA_CONST double *)(reinterpret_cast<const double *>(I.tphase) + local * I.tstride), t);
        store_tile<kC64>(v, pp, base ^ toff, G);
        store_tile<kC64>(w, pl, base ^ toff, G);
        reinterpret_cast<double *>(I.partial)[(local * kWavesPerTile + (t >> 6)) * 64 + lane] = R.total(lane);
    }
}

template <bool kC64, bool kAdj>
auto batch_kernel_for(bool l1, bool l2)
{
    if constexpr (kAdj)
    {
        if (l1 && l2)
            return batch_adj_kernel<kC64, true, true>;
        if (l1)
            return batch_adj_kernel<kC64, true, false>;
        if (l2)
            return batch_adj_kernel<kC64, false, true>;
        return batch_adj_kernel<kC64, false, false>;
    }
    else
    {
        if (l1 && l2)
            return batch_fwd_kernel<kC64, true, true>;
        if (l1)
            return batch_fwd_kernel<kC64, true, false>;
        if (l2)
            return batch_fwd_kernel<kC64, false, true>;
        return batch_fwd_kernel<kC64, false, false>;
    }
}

using BatchKernel = void (*)(const BItem *, int, uint64_t);

BatchKernel batch_kernel(bool c64, bool adj, bool l1, bool l2)
{
    if (c64)
        return adj ? batch_kernel_for<true, true>(l1, l2) : batch_kernel_for<true, false>(l1, l2);
    return adj ? batch_kernel_for<false, true>(l1, l2) : batch_kernel_for<false, false>(l1, l2);
}

// Per item (one workgroup each): the rows of its tiles and waves summed in a
// fixed order, then its gradient entries added in order.
struct BCollect
{
    uint64_t partial;
    uint64_t ntiles;
    uint64_t grad;    // TGrad[kAccs]
    uint64_t grads;   // double*
    int64_t nparams;  // length of grads
};

__global__ __launch_bounds__(kThreads) void batch_collect_kernel(const BCollect *__restrict__ items)
{
    const BCollect &I = items[blockIdx.x];
    collect_rows(reinterpret_cast<const double *>(I.partial), I.ntiles * kWavesPerTile,
                 reinterpret_cast<const TGrad *>(I.grad), I.nparams, reinterpret_cast<double *>(I.grads));
}

} // namespace

unsigned tile_seg(int prec, int geometry)
{
    if (geometry >= 3 && geometry <= 6)
        return unsigned(geometry);
    const int s = tune_get(prec == C64 ? "tile.seg.c64" : "tile.seg.c128");
    return s >= 3 && s <= 6 ? unsigned(s) : 4u;
}

unsigned tile_low_bits(int prec, int geometry)
{
    return tile_seg(prec, geometry) + (prec == C64 ? 1u : 0u);
}

unsigned tile_upper_slots(int prec, int geometry)
{
    return unsigned(kUnitBits) - tile_seg(prec, geometry);
}

bool tile_wide_fits(int prec, unsigned n, const std::vector<unsigned> &targets)
{
    const unsigned shift = prec == C64 ? 1u : 0u;
    if (tune_get(prec == C64 ? "tile.wide.c64" : "tile.wide.c128") != 1)
        return false;
    if (n < unsigned(kW13Bits) + shift || n > 40 || targets.size() > size_t(kW13Slots))
        return false;
    for (size_t i = 0; i < targets.size(); i++)
    {
        if (targets[i] < 4 + shift || targets[i] >= n)
            return false;
        for (size_t j = 0; j < i; j++)
            if (targets[j] == targets[i])
                return false;
    }
    return true;
}

// Which address bits a pass's slots sit on matters more than its run length.
// Nine slots on consecutive address bits that begin 8-12 unit bits above the
// run (strides 64 KiB-1 MiB) move the state at about 51% of HBM; nine slots
// two bits apart, anywhere, at 61-64%; nine consecutive slots right above the
// run (one contiguous tile) at 70%.  So a batch that needs two or more wide
// passes gives its nine lowest upper targets, when consecutive from the
// bottom of the wide tile's slots, a wide pass of their own, and deals the
// others out in turn, ascending, among the remaining wide passes and the
// first pass.
bool tile_layer_plan(int prec, unsigned n, int geometry, const std::vector<unsigned> &targets,
                     std::vector<std::vector<unsigned>> &passes)
{
    passes.clear();
    std::vector<unsigned> ts(targets);
    std::sort(ts.begin(), ts.end());
    if (std::adjacent_find(ts.begin(), ts.end()) != ts.end() || (!ts.empty() && ts.back() >= n))
        return false;
    const unsigned lowb = tile_low_bits(prec, geometry);
    const size_t slots = tile_upper_slots(prec, geometry);
    const unsigned wlow = 4 + (prec == C64 ? 1u : 0u);
    std::vector<unsigned> first, pool;
    size_t first_high = 0;
    for (unsigned t : ts)
    {
        if (t < lowb)
            first.push_back(t);
        else if (t < wlow)
        {
            first.push_back(t);
            first_high++;
        }
        else
            pool.push_back(t);
    }
    const size_t nhigh = first_high + pool.size();
    if (nhigh <= slots || first_high > slots)
        return false;
    const size_t W = size_t(kW13Slots);
    size_t nw = (nhigh - slots + W - 1) / W;
    // no target below the wide tile's run: wide passes alone when they take
    // fewer passes (nine targets above the run are one wide pass)
    const bool wide_only = !ts.empty() && ts.front() >= wlow && (ts.size() + W - 1) / W < 1 + nw;
    if (wide_only)
    {
        nw = (ts.size() + W - 1) / W;
        pool = ts;
        first.clear();
        first_high = 0;
    }
    std::vector<std::vector<unsigned>> wide(nw);
    size_t w0 = 0;
    if (nw >= 2 && pool.size() >= W && pool[W - 1] == wlow + unsigned(W) - 1)
    {
        wide[0].assign(pool.begin(), pool.begin() + long(W));
        pool.erase(pool.begin(), pool.begin() + long(W));
        w0 = 1;
    }
    // in turn, in proportion to the room each pass has (the wide passes w0..,
    // then the first pass): target j goes to the pass furthest behind its
    // share (j + 1) cap / total, the earlier pass among equals
    const size_t np = nw - w0 + (wide_only ? 0 : 1);
    std::vector<int64_t> cap(np, int64_t(W)), got(np, 0);
    if (!wide_only)
        cap[np - 1] = int64_t(slots - first_high);
    int64_t total = 0;
    for (int64_t c : cap)
        total += c;
    if (total < int64_t(pool.size()))
        return false;
    for (size_t j = 0; j < pool.size(); j++)
    {
        size_t best = np;
        int64_t lead = 0;
        for (size_t p = 0; p < np; p++)
        {
            const int64_t d = cap[p] * int64_t(j + 1) - got[p] * total;
            if (got[p] < cap[p] && (best == np || d > lead))
            {
                best = p;
                lead = d;
            }
        }
        got[best]++;
        if (wide_only || best + 1 < np)
            wide[w0 + best].push_back(pool[j]);
        else
            first.push_back(pool[j]);
    }
    for (const std::vector<unsigned> &w : wide)
        if (w.empty() || !tile_wide_fits(prec, n, w))
            return false;
    if (!wide_only)
        passes.push_back(first);
    passes.insert(passes.end(), wide.begin(), wide.end());
    return true;
}

unsigned tile_min_bits(int prec)
{
    return unsigned(kUnitBits) + (prec == C64 ? 1u : 0u);
}

bool tile_staged_fits(int prec, unsigned n, const std::vector<unsigned> &targets, int geometry)
{
    std::vector<int> ua;
    if (n < tile_min_bits(prec) || n > 40)
        return false;
    for (unsigned t : targets)
        if (t >= n)
            return false;
    return staged_units(prec, geometry, targets, ua);
}

// Pivots of a set of pair offsets: each offset reduced by the earlier ones
// (elimination over GF(2)) and its top bit taken, so that the offsets
// restricted to the pivots form an invertible matrix.  A tile whose slots
// carry the offsets and whose outer bits skip the pivots is then a coset of
// the offsets' span, and the tiles cover the state once.  Taking the top bit
// keeps the contiguous low bits free wherever an offset reaches above them.
// False if the offsets are linearly dependent.
bool pivot_bits(const std::vector<const U3Gate *> &gs, unsigned low, std::vector<unsigned> &piv, unsigned *nhigh)
{
    std::vector<uint64_t> w;
    piv.clear();
    *nhigh = 0;
    for (const U3Gate *g : gs)
    {
        uint64_t x = g->delta;
        for (size_t i = 0; i < w.size(); i++)
            if ((x >> piv[i]) & 1)
                x ^= w[i];
        if (x == 0)
            return false;
        const unsigned p = 63u - unsigned(__builtin_clzll(x));
        w.push_back(x);
        piv.push_back(p);
        if (p >= low)
            (*nhigh)++;
    }
    return true;
}

// A batch acts on distinct virtual qubits when every gate's selector sees
// its own pair offset and no other (alpha_i . delta_j = [i == j]): then the
// gates commute and each is a single-qubit gate in the coordinates the
// offsets span, so any grouping into passes gives the batch's product.
bool independent_batch(const U3Gate *gates, size_t ngates)
{
    for (size_t i = 0; i < ngates; i++)
    {
        const uint64_t alpha = gates[i].packed_sel & ((uint64_t(1) << 63) - 1);
        for (size_t j = 0; j < ngates; j++)
            if (unsigned(__builtin_popcountll(alpha & gates[j].delta) & 1) != unsigned(i == j))
                return false;
    }
    return true;
}

// Passes of a batch for one tile geometry.  Single-bit offsets in ascending
// order (the low ones ride with the first pass), then the wider ones.  A pass
// takes every remaining gate whose pivot fits: the pivots above the tile's
// low bits within its upper slots, and no offset reaching above the low bits
// on a pivot below the first whole 128-byte line of the contiguous run (its
// lanes would read partial lines of two places).  A pass's first gate
// always fits, so every pass takes at least one.
bool plan_passes(int prec, int geometry, std::vector<const U3Gate *> rest,
                 std::vector<std::vector<SingleGate>> &passes)
{
    const bool c64 = prec == C64;
    const unsigned low = tile_low_bits(prec, geometry);
    const unsigned slots = tile_upper_slots(prec, geometry);
    const unsigned whole = 3u + (c64 ? 1u : 0u);
    std::vector<unsigned> piv;
    unsigned nhigh = 0;
    passes.clear();
    while (!rest.empty())
    {
        std::vector<const U3Gate *> pass, later;
        for (const U3Gate *g : rest)
        {
            pass.push_back(g);
            const bool fits = pivot_bits(pass, low, piv, &nhigh) && nhigh <= slots &&
                              !(piv.back() < std::min(low, whole) && (g->delta >> low) != 0);
            if (!fits)
            {
                pass.pop_back();
                later.push_back(g);
            }
        }
        if (pass.empty())
            return false;
        rest.swap(later);
        pivot_bits(pass, low, piv, &nhigh);
        std::vector<SingleGate> gs;
        for (size_t k = 0; k < pass.size(); k++)
            gs.push_back({piv[k], pass[k]});
        passes.push_back(gs);
    }
    return true;
}

bool window_tabulable(unsigned n, const DiagGate *gates, size_t ngates)
{
    WinBasis B;
    return n <= 40 && win_basis(n, gates, ngates, B);
}

bool window_table_host(unsigned n, const DiagGate *gates, size_t ngates, uint64_t *basis, int *rank,
                       std::vector<double> &table)
{
    WinBasis B;
    if (n > 40 || !win_basis(n, gates, ngates, B))
        return false;
    win_table(B, gates, ngates, table);
    for (int k = 0; k < kWinBits; k++)
        basis[k] = k < B.rank ? B.a[k] : 0;
    *rank = B.rank;
    return true;
}

int u3_apply(Ctx &ctx, void *state, int prec, unsigned n, const U3Gate *gates, size_t ngates, const DiagGate *win,
             size_t nwin)
{
    const bool c64 = prec == C64;
    // the window rides the last tile pass when it is tabulable; else (or
    // when no tile pass runs) it gets a sweep of its own
    const bool ride = nwin > 0 && window_tabulable(n, win, nwin);
    auto rest = [&]() { return nwin > 0 ? diag_apply(ctx, state, prec, n, win, nwin) : OK; };
    if (!independent_batch(gates, ngates))
    {
        // gates that do not commute: one sweep each, in order
        for (size_t i = 0; i < ngates; i++)
            ZXA_TRY(u3_pair_apply(ctx, state, prec, n, gates[i]));
        return rest();
    }
    std::vector<const U3Gate *> single, wide;
    for (size_t i = 0; i < ngates; i++)
    {
        const U3Gate &g = gates[i];
        const uint64_t alpha = g.packed_sel & ((uint64_t(1) << 63) - 1);
        const bool in_state = g.delta != 0 && (g.delta >> n) == 0 && (alpha >> n) == 0;
        // complex64 units hold address bit 0: an offset may use it only alone
        const bool in_units = !c64 || (g.delta & 1u) == 0 || g.delta == 1;
        if (n >= tile_min_bits(prec) && n <= 40 && in_state && in_units)
            ((g.delta & (g.delta - 1)) == 0 ? single : wide).push_back(&g);
        else
            ZXA_TRY(u3_pair_apply(ctx, state, prec, n, g));
    }
    if (single.empty() && wide.empty())
        return rest();
    std::sort(single.begin(), single.end(), [](const U3Gate *a, const U3Gate *b) { return a->delta < b->delta; });
    std::vector<const U3Gate *> order = single;
    order.insert(order.end(), wide.begin(), wide.end());
    // single-bit batches: the configured geometry.  Wide batches: the
    // geometry with the fewest passes (wide offsets take pivots above the
    // low bits, which shorter contiguous runs give more slots for), among
    // equal counts the longest runs
    const int own = int(tile_seg(prec));
    std::vector<std::vector<SingleGate>> best, cand;
    int geo = -1;
    for (int g = wide.empty() ? own : 6; g >= (wide.empty() ? own : 3); g--)
        if (plan_passes(prec, g, order, cand) && (geo < 0 || cand.size() < best.size()))
        {
            best.swap(cand);
            geo = g;
        }
    if (geo < 0)
    {
        set_error("u3_apply: dependent pair offsets in one batch");
        return EINVAL_;
    }
    // single-bit batches: the first pass as planned, the remaining targets in
    // wide passes of nine when that takes no more passes (their runs are
    // longer)
    std::vector<bool> wide13(best.size(), false);
    if (wide.empty() && best.size() >= 2)
    {
        // the first pass as a 4096-unit tile pass, the others wide
        // (tile_layer_plan)
        std::vector<SingleGate> all;
        for (const std::vector<SingleGate> &ps : best)
            all.insert(all.end(), ps.begin(), ps.end());
        std::vector<unsigned> ts;
        for (const SingleGate &g : all)
            ts.push_back(g.target);
        std::vector<std::vector<unsigned>> plan;
        std::vector<std::vector<SingleGate>> alt;
        const bool ok = tile_layer_plan(prec, n, geo, ts, plan);
        for (size_t k = 0; ok && k < plan.size(); k++)
        {
            alt.emplace_back();
            for (unsigned t : plan[k])
                alt.back().push_back(*std::find_if(all.begin(), all.end(),
                                                   [&](const SingleGate &g) { return g.target == t; }));
        }
        if (ok && alt.size() <= best.size())
        {
            best.swap(alt);
            // the wide passes (every pass but a 4096-unit first pass)
            wide13.assign(best.size(), false);
            for (size_t k = 0; k < plan.size(); k++)
                wide13[k] = tile_wide_fits(prec, n, plan[k]);
        }
    }
    bool rode = false;
    for (size_t k = 0; k < best.size(); k++)
    {
        const bool last = k + 1 == best.size();
        if (wide13[k])
        {
            const bool take = last && ride && window_is_polynomial(n, win, nwin);
            ZXA_TRY(run_wide13_pass(ctx, state, prec, n, best[k], take ? win : nullptr, take ? nwin : 0));
            rode = rode || take;
        }
        else
        {
            ZXA_TRY(run_tile_pass(ctx, state, prec, n, best[k], last && ride ? win : nullptr,
                                  last && ride ? nwin : 0, geo));
            rode = rode || (last && ride);
        }
    }
    return rode ? OK : rest();
}

int fused_pass_apply(Ctx &ctx, void *state, int prec, unsigned n, const U3Gate *gates, size_t ngates,
                     const DiagGate *diag, size_t ndiag, int geometry, const GateFan *fans, bool fresh)
{
    if (n < tile_min_bits(prec) || n > 40)
    {
        set_error("fused pass: state narrower than a tile");
        return EINVAL_;
    }
    bool any_fan = false;
    for (size_t i = 0; fans && i < ngates; i++)
        any_fan = any_fan || fans[i].on;
    if (ndiag > 0 && !window_is_polynomial(n, diag, ndiag) && (any_fan || !window_tabulable(n, diag, ndiag)))
    {
        set_error(any_fan ? "fused pass: a pass with fans needs a phase-polynomial window"
                          : "fused pass: diagonal window is neither a phase polynomial nor tabulable");
        return EINVAL_;
    }
    const unsigned low = tile_low_bits(prec, geometry);
    std::vector<SingleGate> gs;
    int high = 0;
    for (size_t i = 0; i < ngates; i++)
    {
        const U3Gate &g = gates[i];
        const uint64_t alpha = g.packed_sel & ((uint64_t(1) << 63) - 1);
        if (g.delta == 0 || (g.delta & (g.delta - 1)) != 0 || alpha != g.delta || (g.delta >> n) != 0)
        {
            set_error("fused pass: gate is not a single-bit U3");
            return EINVAL_;
        }
        const unsigned t = unsigned(__builtin_ctzll(g.delta));
        for (const SingleGate &o : gs)
            if (o.target == t)
            {
                set_error("fused pass: two gates on one target");
                return EINVAL_;
            }
        if (t >= low)
            high++;
        gs.push_back({t, &g});
    }
    if (unsigned(high) > tile_upper_slots(prec, geometry))
    {
        // the wide tile: nine targets above a 16-unit run
        std::vector<unsigned> ts;
        for (const SingleGate &g : gs)
            ts.push_back(g.target);
        if (!any_fan && tile_wide_fits(prec, n, ts) && (ndiag == 0 || window_is_polynomial(n, diag, ndiag)))
        {
            if (fresh)
                ZXA_TRY(state_reset(ctx, state, prec, size_t(1) << n));
            return run_wide13_pass(ctx, state, prec, n, gs, diag, ndiag);
        }
        set_error("fused pass: more targets above the tile's low bits than the tile has slots");
        return EINVAL_;
    }
    if (gs.empty())
    {
        if (fresh)
            ZXA_TRY(state_reset(ctx, state, prec, size_t(1) << n));
        return diag_apply(ctx, state, prec, n, diag, ndiag);
    }
    return run_tile_pass(ctx, state, prec, n, gs, diag, ndiag, geometry, fans, fresh);
}

} // namespace zxa

namespace zxa
{

int adjoint_pass_apply(Ctx &ctx, void *psi, void *lam, int prec, unsigned n, const AdjTarget *targets,
                       size_t ntargets, const DiagGate *diag, size_t ndiag, double *dgrads, size_t nparams,
                       int geometry)
{
    const bool c64 = prec == C64;
    if (n < tile_min_bits(prec) || n > 40)
    {
        set_error("adjoint pass: state narrower than a tile");
        return EINVAL_;
    }
    if (ndiag > 0 && !window_is_polynomial(n, diag, ndiag))
    {
        set_error("adjoint pass: diagonal window is not a phase polynomial");
        return EINVAL_;
    }
    std::vector<U3Gate> ug(ntargets);
    std::vector<SingleGate> gs;
    for (size_t i = 0; i < ntargets; i++)
    {
        const unsigned q = targets[i].target;
        if (q >= n)
        {
            set_error("adjoint pass: target outside the state");
            return EINVAL_;
        }
        for (size_t j = 0; j < i; j++)
            if (targets[j].target == q)
            {
                set_error("adjoint pass: two entries on one target");
                return EINVAL_;
            }
        ug[i].packed_sel = uint64_t(1) << q;
        ug[i].delta = uint64_t(1) << q;
        std::memcpy(ug[i].u, targets[i].m, sizeof(ug[i].u));
        gs.push_back({q, &ug[i]});
    }
    Plan pl;
    if (!plan_tile(prec, n, gs, pl, geometry))
    {
        set_error("adjoint pass: more targets than the tile has slots");
        return EINVAL_;
    }
    TPass P;
    TTab G;
    bool use[3];
    fill_pass(prec, n, pl, P, G, use);
    TGrad tg[kAccs];
    clear_grads(tg);
    for (size_t i = 0; i < ntargets; i++)
    {
        const int slot = pl.slot[i];
        for (int k = 0; k < 8; k++)
        {
            G.g[slot].u[k] = targets[i].m[k];
            G.g[slot].uf[k] = float(targets[i].m[k]);
        }
        ZXA_TRY(grad_entries(c64, slot, targets[i], nparams, tg));
    }
    const double *ttab = nullptr;
    const double *tphase = nullptr;
    PassBatch batch(ctx);
    ZXA_TRY(batch.begin());
    if (ndiag > 0)
        ZXA_TRY(setup_epilogue(ctx, prec, n, pl.ebit, diag, ndiag, P.ntiles, &ttab, &tphase, 8, batch.ph));
    void *dtab = nullptr, *dgr = nullptr;
    ZXA_TRY(ctx.ring.stage(&G, sizeof(G), ctx.stream, &dtab));
    ZXA_TRY(ctx.ring.stage(tg, sizeof(tg), ctx.stream, &dgr));
    ZXA_TRY(batch.run());
    // per-wave rows of 64 lane accumulators: at most 8 workgroups per CU
    const size_t maxw = size_t(8) * size_t(ctx.num_cus > 0 ? ctx.num_cus : 256) * kWavesPerTile;
    ZXA_TRY(ctx.scratch_f.ensure(maxw * 64 * sizeof(double)));
    double *partial = static_cast<double *>(ctx.scratch_f.ptr);
    const TTab *tab = static_cast<const TTab *>(dtab);
    const uint64_t bytes = 4ull * (uint64_t(1) << n) * elem_bytes(prec);
    int nwaves = int(maxw);
    ZXA_TRY(ctx.tic());
    int rc;
    if (c64)
        rc = ndiag ? launch_adj_l<true, true>(ctx, psi, lam, P, use[1], use[2], tab, ttab, tphase, partial, &nwaves)
                   : launch_adj_l<true, false>(ctx, psi, lam, P, use[1], use[2], tab, ttab, tphase, partial, &nwaves);
    else
        rc = ndiag ? launch_adj_l<false, true>(ctx, psi, lam, P, use[1], use[2], tab, ttab, tphase, partial, &nwaves)
                   : launch_adj_l<false, false>(ctx, psi, lam, P, use[1], use[2], tab, ttab, tphase, partial,
                                                &nwaves);
    ZXA_TRY(rc);
    ZXA_TRY(ctx.toc(F_ADJ, bytes));
    hipLaunchKernelGGL(adj_collect_kernel, dim3(1), dim3(kThreads), 0, ctx.stream, partial, uint64_t(nwaves),
                       static_cast<const TGrad *>(dgr), int64_t(nparams), dgrads);
    ZXA_HIP(hipGetLastError());
    return ctx.ring.done(ctx.stream);
}

} // namespace zxa

namespace zxa
{

BatchLaunch::~BatchLaunch()
{
    // d_adjs points into the d_tabs allocation
    for (void *p : {d_items, d_tabs, d_partial, d_collect, d_epi})
        if (p)
            (void)hipFree(p);
    if (h_tables)
        (void)hipHostFree(h_tables);
}

namespace
{
template <typename T>
int dmalloc(T **p, size_t bytes)
{
    void *v = nullptr;
    hipError_t e = hipMalloc(&v, std::max<size_t>(bytes, 16));
    if (e != hipSuccess)
    {
        set_error(std::string("batch: device allocation failed: ") + hipGetErrorString(e));
        return ENOMEM_;
    }
    *p = static_cast<T *>(v);
    return OK;
}
} // namespace

int batch_prepare(Ctx &ctx, int prec, bool adj, const std::vector<BatchPassSpec> &specs, BatchLaunch &b)
{
    const bool c64 = prec == C64;
    b.prec = prec;
    b.adj = adj;
    b.nitems = specs.size();
    b.slots.assign(specs.size(), {});
    if (specs.empty())
        return OK;
    std::vector<BItem> items(specs.size());
    std::vector<TTab> tabs(specs.size());
    std::vector<TGrad> grads(adj ? specs.size() * kAccs : 0);
    for (size_t i = 0; i < specs.size() && adj; i++)
        clear_grads(&grads[i * kAccs]);
    b.nparams.assign(specs.size(), 0);
    // epilogue tables: per item, per-thread table then per-tile factors.  The
    // items without a window share an identity per-thread table and one
    // identity row of per-tile factors at the start (tstride 0), so that
    // kernels can load and apply an epilogue without a branch.
    const int nb = c64 ? 5 : 4;
    const size_t ttab_doubles = size_t(kThreads) * 2 * (1 + nb);
    const size_t stride = 2 * (8 + (size_t(1) << nb));
    std::vector<size_t> epi_off(specs.size(), 0), epi_tt(specs.size(), 0);
    size_t epi_bytes = (ttab_doubles + stride) * sizeof(double);
    uint64_t first = 0, partial_doubles = 0;
    std::vector<Plan> plans(specs.size());
    for (size_t i = 0; i < specs.size(); i++)
    {
        const BatchPassSpec &sp = specs[i];
        if (sp.n < tile_min_bits(prec) || sp.n > 40)
        {
            set_error("batch: state narrower than a tile");
            return EINVAL_;
        }
        std::vector<U3Gate> ug(sp.targets.size());
        std::vector<SingleGate> gs;
        for (size_t k = 0; k < sp.targets.size(); k++)
        {
            ug[k].packed_sel = ug[k].delta = uint64_t(1) << sp.targets[k];
            gs.push_back({sp.targets[k], &ug[k]});
        }
        if (!plan_tile(prec, sp.n, gs, plans[i], sp.geometry))
        {
            set_error("batch: pass does not fit one tile");
            return EINVAL_;
        }
        TPass P;
        bool use[3];
        fill_pass(prec, sp.n, plans[i], P, tabs[i], use);
        b.l1 = b.l1 || use[1];
        b.l2 = b.l2 || use[2];
        b.slots[i] = plans[i].slot;
        BItem &it = items[i];
        std::memset(&it, 0, sizeof(it));
        it.first = first;
        it.ntiles = P.ntiles;
        it.psi = reinterpret_cast<uint64_t>(sp.psi);
        it.lam = reinterpret_cast<uint64_t>(sp.lam);
        for (int k = 0; k < kUnitBits; k++)
        {
            it.ub[k] = P.ub[k];
            it.sorted[k] = P.sorted[k];
        }
        first += P.ntiles;
        if (adj)
        {
            it.partial = partial_doubles; // offset for now
            partial_doubles += P.ntiles * kWavesPerTile * 64;
            b.nparams[i] = sp.nparams;
        }
        if (!sp.window.empty())
        {
            if (!window_is_polynomial(sp.n, sp.window.data(), sp.window.size()))
            {
                set_error("batch: window is not a phase polynomial");
                return EINVAL_;
            }
            epi_tt[i] = epi_bytes;
            epi_bytes += ttab_doubles * sizeof(double);
            epi_off[i] = epi_bytes;
            epi_bytes += P.ntiles * stride * sizeof(double);
        }
    }
    b.total_tiles = first;
    b.tab_bytes = specs.size() * sizeof(TTab);
    b.adj_bytes = grads.size() * sizeof(TGrad);
    char *dtabs = nullptr;
    ZXA_TRY(dmalloc(&dtabs, b.tab_bytes + b.adj_bytes));
    b.d_tabs = dtabs;
    b.d_adjs = adj ? dtabs + b.tab_bytes : nullptr;
    ZXA_HIP(hipHostMalloc(&b.h_tables, b.tab_bytes + b.adj_bytes, hipHostMallocDefault));
    std::memcpy(b.h_tables, tabs.data(), b.tab_bytes);
    if (adj)
        std::memcpy(static_cast<char *>(b.h_tables) + b.tab_bytes, grads.data(), b.adj_bytes);
    // the device tables start valid (no gradient entries), so a launch
    // before the first set/upload is harmless
    ZXA_HIP(hipMemcpy(dtabs, b.h_tables, b.tab_bytes + b.adj_bytes, hipMemcpyHostToDevice));
    char *depi = nullptr;
    ZXA_TRY(dmalloc(&depi, epi_bytes));
    b.d_epi = depi;
    {
        std::vector<double> ident(ttab_doubles + stride, 0.0);
        for (size_t k = 0; k < ident.size(); k += 2)
            ident[k] = 1.0;
        ZXA_HIP(hipMemcpy(depi, ident.data(), ident.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    double *dpart = nullptr;
    if (adj)
        ZXA_TRY(dmalloc(&dpart, partial_doubles * sizeof(double)));
    b.d_partial = dpart;
    // epilogue tables (fixed): per-thread table uploaded, per-tile factors by tile_phase_kernel
    for (size_t i = 0; i < specs.size(); i++)
    {
        BItem &it = items[i];
        it.tab = reinterpret_cast<uint64_t>(dtabs + i * sizeof(TTab));
        if (adj)
            it.partial = reinterpret_cast<uint64_t>(dpart + it.partial);
        const BatchPassSpec &sp = specs[i];
        it.ttab = reinterpret_cast<uint64_t>(depi);
        it.tphase = reinterpret_cast<uint64_t>(depi + ttab_doubles * sizeof(double));
        it.tstride = 0;
        if (sp.window.empty())
            continue;
        Poly poly;
        window_polynomial(sp.n, sp.window.data(), sp.window.size(), poly);
        std::vector<double> tt, pt;
        int nh = 0, nb2 = 0;
        epilogue_tables(prec, sp.n, plans[i].ebit, poly, tt, pt, &nh, &nb2);
        ZXA_HIP(hipMemcpy(depi + epi_tt[i], tt.data(), tt.size() * sizeof(double), hipMemcpyHostToDevice));
        double *dpt = nullptr;
        ZXA_TRY(dmalloc(&dpt, pt.size() * sizeof(double)));
        ZXA_HIP(hipMemcpy(dpt, pt.data(), pt.size() * sizeof(double), hipMemcpyHostToDevice));
        launch_tile_phase(ctx.stream, dpt, pt.size(), int(plans[i].ebit.size()), nh, nb2, poly.K, it.ntiles,
                          reinterpret_cast<double *>(depi + epi_off[i]));
        hipError_t e = hipGetLastError();
        if (e == hipSuccess)
            e = hipStreamSynchronize(ctx.stream);
        (void)hipFree(dpt);
        ZXA_HIP(e);
        it.ttab = reinterpret_cast<uint64_t>(depi + epi_tt[i]);
        it.tphase = reinterpret_cast<uint64_t>(depi + epi_off[i]);
        it.tstride = stride;
    }
    ZXA_TRY(dmalloc(&b.d_items, items.size() * sizeof(BItem)));
    ZXA_HIP(hipMemcpy(b.d_items, items.data(), items.size() * sizeof(BItem), hipMemcpyHostToDevice));
    if (adj)
    {
        std::vector<BCollect> col(specs.size());
        for (size_t i = 0; i < specs.size(); i++)
        {
            col[i].partial = items[i].partial;
            col[i].ntiles = items[i].ntiles;
            col[i].grad = reinterpret_cast<uint64_t>(dtabs + b.tab_bytes + i * kAccs * sizeof(TGrad));
            col[i].grads = reinterpret_cast<uint64_t>(specs[i].grads);
            col[i].nparams = int64_t(specs[i].nparams);
        }
        ZXA_TRY(dmalloc(&b.d_collect, col.size() * sizeof(BCollect)));
        ZXA_HIP(hipMemcpy(b.d_collect, col.data(), col.size() * sizeof(BCollect), hipMemcpyHostToDevice));
    }
    b.grid = persistent_grid(ctx, batch_kernel(c64, adj, b.l1, b.l2), b.total_tiles);
    return OK;
}

void batch_set_gates(BatchLaunch &b, size_t item, const U3Gate *gates, size_t ngates)
{
    TTab *tab = reinterpret_cast<TTab *>(b.h_tables) + item;
    for (size_t k = 0; k < ngates && k < b.slots[item].size(); k++)
    {
        TGate &tg = tab->g[b.slots[item][k]];
        for (int i = 0; i < 8; i++)
        {
            tg.u[i] = gates[k].u[i];
            tg.uf[i] = float(gates[k].u[i]);
        }
    }
}

int batch_set_adjoint(BatchLaunch &b, size_t item, const AdjTarget *targets, size_t ntargets)
{
    TTab *tab = reinterpret_cast<TTab *>(b.h_tables) + item;
    TGrad *gr = reinterpret_cast<TGrad *>(static_cast<char *>(b.h_tables) + b.tab_bytes) + item * kAccs;
    clear_grads(gr);
    for (size_t k = 0; k < ntargets && k < b.slots[item].size(); k++)
    {
        const int slot = b.slots[item][k];
        TGate &tg = tab->g[slot];
        for (int i = 0; i < 8; i++)
        {
            tg.u[i] = targets[k].m[i];
            tg.uf[i] = float(targets[k].m[i]);
        }
        ZXA_TRY(grad_entries(b.prec == C64, slot, targets[k], b.nparams[item], gr));
    }
    return OK;
}

int batch_upload(Ctx &ctx, BatchLaunch &b)
{
    if (!b.nitems)
        return OK;
    return hip_check(hipMemcpyAsync(b.d_tabs, b.h_tables, b.tab_bytes + b.adj_bytes,
                                    hipMemcpyHostToDevice, ctx.stream),
                     "batch table upload");
}

int batch_launch(Ctx &ctx, BatchLaunch &b)
{
    if (!b.nitems)
        return OK;
    const bool c64 = b.prec == C64;
    hipLaunchKernelGGL(batch_kernel(c64, b.adj, b.l1, b.l2), dim3(b.grid), dim3(kThreads), 0, ctx.stream, static_cast<const BItem *>(b.d_items),
                       int(b.nitems), b.total_tiles);
    ZXA_HIP(hipGetLastError());
    if (b.adj)
    {
        hipLaunchKernelGGL(batch_collect_kernel, dim3(unsigned(b.nitems)), dim3(kThreads), 0, ctx.stream,
                           static_cast<const BCollect *>(b.d_collect));
        ZXA_HIP(hipGetLastError());
    }
    return OK;
}

} // namespace zxa
