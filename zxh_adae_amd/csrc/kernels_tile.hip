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
// chosen "unit bits": the four lowest unit-address bits, so that each 16-lane
// group of a wave moves 256 contiguous bytes, plus up to eight target bits.
// A complex64 tile therefore spans address bits 0..4 and eight more.  Each
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
#include <cmath>
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
    uint32_t pad;
};

// Kernel arguments of a pass.  Unit bits 0..3 are unit-address bits 0..3;
// the host assigns the other eight freely (targets first into the register
// bits of layout 0, then of layout 1), so ub is not sorted; sorted holds the
// same positions ascending.
struct TPass
{
    uint32_t ub[kUnitBits];     // unit-address bit of each unit bit
    uint32_t sorted[kUnitBits]; // the same, ascending
    uint64_t ntiles;
    uint64_t tmask; // unit-address mask of the 12 unit bits
    uint64_t gstep; // the grid size deposited into the non-tile bits
};

// Per-pass tables in device memory.  The kernel re-reads them through the
// scalar cache at every tile instead of holding them in SGPRs across the
// tile loop (12 matrices and 16 offsets would not fit).  Every slot of a
// layout the pass uses holds a matrix (the identity where there is no gate):
// the kernel applies all of them without branches.
struct TTab
{
    uint64_t roff[kRegs]; // layout-0 register offsets (unit addresses)
    TGate g[kUnitBits + 1];
};

// Register bits of the three layouts (unit bits); the thread bits are the
// remaining eight in ascending order.  Layout 0 is the load/store layout: its
// thread bits start with unit bits 0..3, the contiguous ones.
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
        o |= uint64_t((t >> j) & 1) << P.ub[thr_bit(L, j)];
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

// Gates on the register bits of layout L (and, in layout 0, the complex64
// in-unit gate).  base | toff is the unit address of the thread's register 0.
template <bool kC64, bool kFlip, int L>
__device__ __forceinline__ void apply_layout(typename UnitT<kC64>::T (&v)[kRegs], const TPass &P, ZXA_CONST TTab *G,
                                             uint64_t base, uint64_t toff)
{
#pragma unroll
    for (int i = 0; i < 4; i++)
    {
        const int k = reg_bit(L, i);
        ZXA_CONST TGate &g = G->g[k];
        Real<kC64> m[8];
        load_matrix<kC64>(m, g);
        const uint64_t amask = kFlip ? g.amask : 0;
        const uint32_t pt = kFlip ? (par((base | toff) & amask) ^ g.beta) : 0u;
        const bool a0 = kFlip && g.a0 != 0;
#pragma unroll
        for (int r = 0; r < kRegs; r++)
        {
            if ((r >> i) & 1)
                continue;
            const bool flip = kFlip && ((pt ^ (g.rpar >> r)) & 1u) != 0u;
            unit_pair<kC64, kFlip>(v[r], v[r | (1 << i)], m, flip, a0);
        }
    }
    if constexpr (kC64 && L == 0)
    {
        ZXA_CONST TGate &g = G->g[kIntra];
        float m[8];
        load_matrix<true>(m, g);
        const uint64_t amask = kFlip ? g.amask : 0;
        const uint32_t pt = kFlip ? (par((base | toff) & amask) ^ g.beta) : 0u;
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
template <bool kC64>
struct Epi
{
    static constexpr int kNB = kC64 ? 5 : 4;         // register element bits
    static constexpr int kNR = 1 << kNB;             // register elements
    static constexpr int kStride = 2 * (8 + kNR);    // doubles per tile in tphase
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
        for (int j = 0; j < 8; j++)
        {
            const cd e = {tp[2 * j], tp[2 * j + 1]};
            const cd we = cmul(w, e);
            const bool on = (t >> j) & 1;
            w.x = on ? we.x : w.x;
            w.y = on ? we.y : w.y;
        }
        ZXA_CONST double *D = tp + 16;
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
        v[r] = __builtin_nontemporal_load(p + (at | G->roff[r]));
}

template <bool kC64>
__device__ __forceinline__ void store_tile(const typename UnitT<kC64>::T (&v)[kRegs], typename UnitT<kC64>::T *p,
                                           uint64_t at, ZXA_CONST TTab *G)
{
#pragma unroll
    for (int r = 0; r < kRegs; r++)
        __builtin_nontemporal_store(v[r], p + (at | G->roff[r]));
}

// Gates, transposes and the diagonal epilogue of one tile held in v.
template <bool kC64, bool kL1, bool kL2, bool kDiag, bool kFlip>
__device__ __forceinline__ void tile_work(typename UnitT<kC64>::T (&v)[kRegs], typename UnitT<kC64>::T *lds,
                                          const TPass &P, ZXA_CONST TTab *G, const double *tphase,
                                          const Epi<kC64> &epi, int t, uint64_t toff, uint64_t tile, uint64_t base)
{
    apply_layout<kC64, kFlip, 0>(v, P, G, base, toff);
    if constexpr (kL1)
    {
        transpose<kC64, 0, 1>(v, lds, t);
        apply_layout<kC64, kFlip, 1>(v, P, G, base, thr_off<1>(t, P));
        if constexpr (kL2)
        {
            transpose<kC64, 1, 2>(v, lds, t);
            apply_layout<kC64, kFlip, 2>(v, P, G, base, thr_off<2>(t, P));
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
        apply_layout<kC64, kFlip, 2>(v, P, G, base, thr_off<2>(t, P));
        transpose<kC64, 2, 0>(v, lds, t);
    }
    if constexpr (kDiag)
        epi.apply(v, (ZXA_CONST double *)(tphase + tile * Epi<kC64>::kStride), t);
}

// The tables are re-read at every tile (see TTab): the pointer is laundered
// so that the loads are not hoisted out of the tile loop.
__device__ __forceinline__ ZXA_CONST TTab *tables(const TTab *tabp)
{
    asm volatile("" : "+s"(tabp));
    return (ZXA_CONST TTab *)tabp;
}

// tabp: register offsets and gate matrices; ttab: per-thread ST and G
// (kDiag only); tphase: per-tile E and D (kDiag only).
//
// Software pipeline over the workgroup's tiles t0, t1 = t0 + grid, ... with
// two register buffers: while tile k is worked on, tile k + 1 is in flight,
// and the store of tile k - 1 is issued after the loads of tile k + 1, so
// that waiting for a tile's loads never waits for the previous stores.  The
// prefetch is unconditional (past the last tile it re-reads the current one):
// with no branch around it the compiler's wait counts stay exact.
template <bool kC64, bool kL1, bool kL2, bool kDiag, bool kFlip>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(2))) void tile_kernel(
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

    T a[kRegs], b[kRegs];
    uint64_t t0 = blockIdx.x;
    uint64_t b0 = deposit(t0, P.sorted);
    ZXA_CONST TTab *G = tables(tabp);
    load_tile<kC64>(a, p, b0 | toff, G);
    uint64_t t1 = t0 + grid;
    bool m1 = t1 < P.ntiles;
    uint64_t b1 = next_base(b0, P);
    load_tile<kC64>(b, p, (m1 ? b1 : b0) | toff, G);
    tile_work<kC64, kL1, kL2, kDiag, kFlip>(a, lds, P, G, tphase, epi, t, toff, t0, b0);
    for (;;)
    {
        G = tables(tabp);
        store_tile<kC64>(a, p, b0 | toff, G);
        if (!m1)
            break;
        const uint64_t t2 = t1 + grid;
        const bool m2 = t2 < P.ntiles;
        const uint64_t b2 = next_base(b1, P);
        load_tile<kC64>(a, p, (m2 ? b2 : b1) | toff, G);
        tile_work<kC64, kL1, kL2, kDiag, kFlip>(b, lds, P, G, tphase, epi, t, toff, t1, b1);

        G = tables(tabp);
        store_tile<kC64>(b, p, b1 | toff, G);
        if (!m2)
            break;
        const uint64_t t3 = t2 + grid;
        const bool m3 = t3 < P.ntiles;
        const uint64_t b3 = next_base(b2, P);
        load_tile<kC64>(b, p, (m3 ? b3 : b2) | toff, G);
        tile_work<kC64, kL1, kL2, kDiag, kFlip>(a, lds, P, G, tphase, epi, t, toff, t2, b2);

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
// the point just after the forward pass's single-qubit gates.  For every
// target slot the pass accumulates up to two transition values
//     Im <lam| A |psi>      (A: a 2x2 Hermitian operator on the slot's qubit,
//                            the generator of a rotation conjugated by the
//                            gates composed after it; the rotation's gradient)
// and then applies the slot's uncompute matrix M = U^dagger to both states.
// The diagonal epilogue (if any) is the previous forward pass's window with
// negated angles, applied to both states.
// ---------------------------------------------------------------------------

constexpr int kObs = 2;                     // transition operators per slot
constexpr int kWavesPerTile = kThreads / 64;
constexpr int kAccs = (kUnitBits + 1) * kObs; // accumulator lanes (< 64)

struct TObs
{
    double a[kObs][8]; // A (re, im) per operator; zero when unused
    float af[kObs][8];
};
struct TAdj
{
    TObs o[kUnitBits + 1];
};

// Im(conj(l0) (A00 p0 + A01 p1) + conj(l1) (A10 p0 + A11 p1))
template <typename R>
__device__ __forceinline__ R pair_im(R p0x, R p0y, R p1x, R p1y, R l0x, R l0y, R l1x, R l1y, const R (&a)[8])
{
    const R a0x = a[0] * p0x - a[1] * p0y + a[2] * p1x - a[3] * p1y;
    const R a0y = a[0] * p0y + a[1] * p0x + a[2] * p1y + a[3] * p1x;
    const R a1x = a[4] * p0x - a[5] * p0y + a[6] * p1x - a[7] * p1y;
    const R a1y = a[4] * p0y + a[5] * p0x + a[6] * p1y + a[7] * p1x;
    return l0x * a0y - l0y * a0x + l1x * a1y - l1y * a1x;
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int m = 32; m > 0; m >>= 1)
        v += __shfl_xor(v, m);
    return v;
}

template <bool kC64>
__device__ __forceinline__ void load_obs(Real<kC64> (&a)[8], ZXA_CONST TObs &o, int k)
{
#pragma unroll
    for (int i = 0; i < 8; i++)
    {
        if constexpr (kC64)
            a[i] = o.af[k][i];
        else
            a[i] = o.a[k][i];
    }
}

// Slot s of layout L (register bit i; s = kIntra for the complex64 in-unit
// bit): transition values into the lane-indexed accumulator, then M on both.
// emit(a, tot): the wave-wide transition value tot of accumulator a (slot
// * kObs + operator), identical in every lane.
template <bool kC64, int L, typename Emit>
__device__ __forceinline__ void adj_layout(typename UnitT<kC64>::T (&v)[kRegs], typename UnitT<kC64>::T (&w)[kRegs],
                                           ZXA_CONST TTab *G, ZXA_CONST TAdj *A, Emit &&emit)
{
#pragma unroll
    for (int i = 0; i < 4; i++)
    {
        const int k = reg_bit(L, i);
#pragma unroll
        for (int o = 0; o < kObs; o++)
        {
            Real<kC64> a[8];
            load_obs<kC64>(a, A->o[k], o);
            double part = 0.0;
#pragma unroll
            for (int r = 0; r < kRegs; r++)
            {
                if ((r >> i) & 1)
                    continue;
                const int r1 = r | (1 << i);
                if constexpr (kC64)
                {
                    part += double(pair_im<float>(v[r].x, v[r].y, v[r1].x, v[r1].y, w[r].x, w[r].y, w[r1].x,
                                                  w[r1].y, a));
                    part += double(pair_im<float>(v[r].z, v[r].w, v[r1].z, v[r1].w, w[r].z, w[r].w, w[r1].z,
                                                  w[r1].w, a));
                }
                else
                {
                    part += pair_im<double>(v[r].x, v[r].y, v[r1].x, v[r1].y, w[r].x, w[r].y, w[r1].x, w[r1].y, a);
                }
            }
            emit(k * kObs + o, wave_sum(part));
        }
        Real<kC64> m[8];
        load_matrix<kC64>(m, G->g[k]);
#pragma unroll
        for (int r = 0; r < kRegs; r++)
        {
            if ((r >> i) & 1)
                continue;
            unit_pair<kC64, false>(v[r], v[r | (1 << i)], m, false, false);
            unit_pair<kC64, false>(w[r], w[r | (1 << i)], m, false, false);
        }
    }
    if constexpr (kC64 && L == 0)
    {
#pragma unroll
        for (int o = 0; o < kObs; o++)
        {
            float a[8];
            load_obs<true>(a, A->o[kIntra], o);
            double part = 0.0;
#pragma unroll
            for (int r = 0; r < kRegs; r++)
                part += double(pair_im<float>(v[r].x, v[r].y, v[r].z, v[r].w, w[r].x, w[r].y, w[r].z, w[r].w, a));
            emit(kIntra * kObs + o, wave_sum(part));
        }
        float m[8];
        load_matrix<true>(m, G->g[kIntra]);
#pragma unroll
        for (int r = 0; r < kRegs; r++)
        {
            const Pair<float> pv = pair2<float>(v[r].x, v[r].y, v[r].z, v[r].w, m, false);
            v[r] = fvec4{pv.ax, pv.ay, pv.bx, pv.by};
            const Pair<float> pw = pair2<float>(w[r].x, w[r].y, w[r].z, w[r].w, m, false);
            w[r] = fvec4{pw.ax, pw.ay, pw.bx, pw.by};
        }
    }
}

// partial: per wave, kAccs doubles (lane-indexed accumulators).
template <bool kC64, bool kL1, bool kL2, bool kDiag>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(2))) void adj_kernel(
    void *__restrict__ psi, void *__restrict__ lam, const TPass P, const TTab *__restrict__ tabp,
    const TAdj *__restrict__ adjp, const double *__restrict__ ttab, const double *__restrict__ tphase,
    double *__restrict__ partial)
{
    using T = typename UnitT<kC64>::T;
    __shared__ T lds[(kL1 || kL2) ? kUnits : 1];
    const int t = threadIdx.x;
    const int lane = t & 63;
    T *pp = static_cast<T *>(psi);
    T *pl = static_cast<T *>(lam);
    const uint64_t toff = thr_off<0>(t, P);
    Epi<kC64> epi;
    if constexpr (kDiag)
        epi.init(ttab, t);
    double acc = 0.0;
    auto emit = [&](int a, double tot) { acc += lane == a ? tot : 0.0; };
    uint64_t base = deposit(uint64_t(blockIdx.x), P.sorted);
    for (uint64_t tile = blockIdx.x; tile < P.ntiles; tile += gridDim.x)
    {
        ZXA_CONST TTab *G = tables(tabp);
        const TAdj *A0 = adjp;
        asm volatile("" : "+s"(A0));
        ZXA_CONST TAdj *A = (ZXA_CONST TAdj *)A0;
        T v[kRegs], w[kRegs];
        load_tile<kC64>(v, pp, base | toff, G);
        load_tile<kC64>(w, pl, base | toff, G);
        adj_layout<kC64, 0>(v, w, G, A, emit);
        if constexpr (kL1)
        {
            transpose<kC64, 0, 1>(v, lds, t);
            transpose<kC64, 0, 1>(w, lds, t);
            adj_layout<kC64, 1>(v, w, G, A, emit);
            if constexpr (kL2)
            {
                transpose<kC64, 1, 2>(v, lds, t);
                transpose<kC64, 1, 2>(w, lds, t);
                adj_layout<kC64, 2>(v, w, G, A, emit);
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
            adj_layout<kC64, 2>(v, w, G, A, emit);
            transpose<kC64, 2, 0>(v, lds, t);
            transpose<kC64, 2, 0>(w, lds, t);
        }
        if constexpr (kDiag)
            epi.template apply_n<true>(v, w, (ZXA_CONST double *)(tphase + tile * Epi<kC64>::kStride), t);
        store_tile<kC64>(v, pp, base | toff, G);
        store_tile<kC64>(w, pl, base | toff, G);
        base = next_base(base, P);
    }
    if (lane < kAccs)
        partial[(size_t(blockIdx.x) * (kThreads / 64) + (t >> 6)) * kAccs + lane] = acc;
}

// Sums the per-wave accumulators in a fixed order and adds slot/operator
// (s, o)'s total into grads[param[s * kObs + o]] (param < 0: not a
// gradient).  One thread per accumulator; distinct accumulators of one pass
// may share a parameter, so the adds into grads run in one thread, in order.
__global__ __launch_bounds__(64) void adj_collect_kernel(const double *__restrict__ partial, int nwaves,
                                                         const int *__restrict__ param, double *__restrict__ grads)
{
    __shared__ double tot[64];
    const int j = threadIdx.x;
    double s = 0.0;
    if (j < kAccs)
        for (int w = 0; w < nwaves; w++)
            s += partial[size_t(w) * kAccs + j];
    tot[j] = s;
    __syncthreads();
    if (j == 0)
        for (int a = 0; a < kAccs; a++)
            if (param[a] >= 0)
                grads[param[a]] += tot[a];
}

__device__ __forceinline__ cd cis(double a)
{
    double s, c;
    sincos(a, &s, &c);
    return {c, s};
}

// Per-tile phase factors of the diagonal epilogue, one thread per tile.
// Tile index bit i is address bit outer[i].  tab (nE tile element bits, nh
// outer bits, nR register elements):
//     cq[k * nh + i] = Q(tile bit k, outer i)   k < nE
//     hL[i] = L(outer i), hQ[i * nh + l] = Q(outer i, outer l)  (i < l)
//     rs[rho]        = static phase of register element rho
//     map[0..8)      = tile element index of thread bit j
//     map[8..8+nb)   = tile element index of register element bit b
// Output per tile: E_j = exp(i c(thr j)) (8), D[rho] = exp(i (C + rs[rho] +
// sum_{b in rho} c(reg b))) (nR), as complex pairs.
__global__ __launch_bounds__(kThreads) void tile_phase_kernel(const double *__restrict__ tab, int nE, int nh, int nb,
                                                              double K, uint64_t ntiles, double *__restrict__ out)
{
    extern __shared__ double sh[];
    const int nR = 1 << nb;
    const int ntab = nE * nh + nh + nh * nh + nR + 8 + nb;
    for (int i = threadIdx.x; i < ntab; i += blockDim.x)
        sh[i] = tab[i];
    __syncthreads();
    const uint64_t tile = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    if (tile >= ntiles)
        return;
    const double *cq = sh;
    const double *hL = cq + nE * nh;
    const double *hQ = hL + nh;
    const double *rs = hQ + nh * nh;
    const double *mp = rs + nR;
    double c[13];
    for (int k = 0; k < nE; k++)
    {
        double s = 0.0;
        for (int i = 0; i < nh; i++)
            if ((tile >> i) & 1u)
                s += cq[k * nh + i];
        c[k] = s;
    }
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
    const int stride = 2 * (8 + nR);
    double *o = out + tile * stride;
    for (int j = 0; j < 8; j++)
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
                a += c[static_cast<int>(mp[8 + b])];
        const cd d = cis(a);
        o[16 + 2 * rho] = d.x;
        o[16 + 2 * rho + 1] = d.y;
    }
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

template <bool kC64, bool kL1, bool kL2, bool kDiag, bool kFlip>
int launch_t(Ctx &ctx, void *state, const TPass &P0, const TTab *tab, const double *ttab, const double *tphase)
{
    auto k = tile_kernel<kC64, kL1, kL2, kDiag, kFlip>;
    const int grid = persistent_grid(ctx, k, P0.ntiles);
    TPass P = P0;
    P.gstep = deposit(uint64_t(grid), P.sorted);
    hipLaunchKernelGGL(k, dim3(grid), dim3(kThreads), 0, ctx.stream, state, P, tab, ttab, tphase);
    return hip_check(hipGetLastError(), "tile_kernel launch");
}

template <bool kC64, bool kDiag, bool kFlip>
int launch_l(Ctx &ctx, void *state, const TPass &P, bool l1, bool l2, const TTab *tab, const double *ttab,
             const double *tphase)
{
    if (l1 && l2)
        return launch_t<kC64, true, true, kDiag, kFlip>(ctx, state, P, tab, ttab, tphase);
    if (l1)
        return launch_t<kC64, true, false, kDiag, kFlip>(ctx, state, P, tab, ttab, tphase);
    if (l2)
        return launch_t<kC64, false, true, kDiag, kFlip>(ctx, state, P, tab, ttab, tphase);
    return launch_t<kC64, false, false, kDiag, kFlip>(ctx, state, P, tab, ttab, tphase);
}

template <bool kC64>
int launch_p(Ctx &ctx, void *state, const TPass &P, bool l1, bool l2, bool flip, const TTab *tab,
             const double *ttab, const double *tphase)
{
    if (ttab)
        return launch_l<kC64, true, false>(ctx, state, P, l1, l2, tab, ttab, tphase);
    if (flip)
        return launch_l<kC64, false, true>(ctx, state, P, l1, l2, tab, nullptr, nullptr);
    return launch_l<kC64, false, false>(ctx, state, P, l1, l2, tab, nullptr, nullptr);
}

struct SingleGate
{
    unsigned target;
    const U3Gate *g;
};

// Tile plan: the address bit of every tile element bit e (complex128: e is
// the unit bit; complex64: e = 0 is address bit 0 inside a unit and e = k + 1
// is unit bit k), and the unit bit of every target.  Unit bits 0..seg-1 are
// the lowest unit-address bits (seg = 4, 5 or 6: 256 B, 512 B or 1 KiB
// contiguous per wave access); targets above them go to the register bits of
// layout 0 (unit bits 8..11), then of layout 1 from unit bit 7 down; the
// remaining slots take the lowest other bits in ascending order (so that in
// the load layout the contiguous run continues across lanes where it can).
struct Plan
{
    std::vector<unsigned> ebit; // tile element bit -> address bit
    std::vector<int> slot;      // per gate: unit bit, or kIntra
};

bool plan_tile(int prec, unsigned n, const std::vector<SingleGate> &gs, Plan &pl)
{
    const bool c64 = prec == C64;
    const int seg = int(tile_seg(prec));
    const unsigned low = tile_low_bits(prec);
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
    std::sort(upper.begin(), upper.end());
    if (upper.size() > tile_upper_slots(prec) || n < tile_min_bits(prec))
        return false;
    // targets: layout-0 register bits, then layout-1 register bits from the top
    std::vector<int> order;
    for (int k : {8, 9, 10, 11, 7, 6, 5, 4})
        if (k >= seg)
            order.push_back(k);
    for (size_t u = 0; u < upper.size(); u++)
    {
        unit_addr[order[u]] = int(upper[u]);
        used[upper[u]] = true;
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
    pl.slot.clear();
    for (const SingleGate &s : gs)
    {
        const int e = int(std::find(pl.ebit.begin(), pl.ebit.end(), s.target) - pl.ebit.begin());
        pl.slot.push_back(c64 ? (e == 0 ? kIntra : e - 1) : e);
    }
    return true;
}

// Host tables of the diagonal epilogue: the per-thread table (ST, G) and the
// packed input of tile_phase_kernel.
void epilogue_tables(int prec, unsigned n, const std::vector<unsigned> &pos, const Poly &poly,
                     std::vector<double> &ttab, std::vector<double> &ptab, int *nh_out, int *nb_out)
// pos: address bit of each tile element bit (Plan::ebit)
{
    const bool c64 = prec == C64;
    const int nE = int(pos.size());
    const int nb = c64 ? 5 : 4;
    const int nR = 1 << nb;
    int thr_k[8], reg_k[5];
    for (int j = 0; j < 8; j++)
        thr_k[j] = c64 ? j + 1 : j;
    for (int b = 0; b < nb; b++)
        reg_k[b] = c64 ? (b == 0 ? 0 : 8 + b) : 8 + b;
    auto Lt = [&](int k) { return poly.L[pos[k]]; };
    auto Qt = [&](int k, int l) { return poly.q(int(pos[k]), int(pos[l])); };

    ttab.assign(size_t(kThreads) * 2 * (1 + nb), 0.0);
    for (int t = 0; t < kThreads; t++)
    {
        double s = 0.0;
        for (int j = 0; j < 8; j++)
        {
            if (!((t >> j) & 1))
                continue;
            s += Lt(thr_k[j]);
            for (int l = j + 1; l < 8; l++)
                if ((t >> l) & 1)
                    s += Qt(thr_k[j], thr_k[l]);
        }
        double *a = ttab.data() + size_t(t) * 2 * (1 + nb);
        a[0] = std::cos(s);
        a[1] = std::sin(s);
        for (int b = 0; b < nb; b++)
        {
            double d = 0.0;
            for (int j = 0; j < 8; j++)
                if ((t >> j) & 1)
                    d += Qt(thr_k[j], reg_k[b]);
            a[2 + 2 * b] = std::cos(d);
            a[3 + 2 * b] = std::sin(d);
        }
    }

    std::vector<int> outer;
    for (unsigned b = 0; b < n; b++)
        if (std::find(pos.begin(), pos.end(), b) == pos.end())
            outer.push_back(int(b));
    const int nh = int(outer.size());
    ptab.assign(size_t(nE) * nh + nh + size_t(nh) * nh + nR + 8 + nb, 0.0);
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
    for (int j = 0; j < 8; j++)
        mp[j] = thr_k[j];
    for (int b = 0; b < nb; b++)
        mp[8 + b] = reg_k[b];
    *nh_out = nh;
    *nb_out = nb;
}

// Stages the diagonal epilogue of a pass: the per-thread table and the
// per-tile phase factors (tile_phase_kernel into scratch_a).
int setup_epilogue(Ctx &ctx, int prec, unsigned n, const std::vector<unsigned> &pos, const DiagGate *diag,
                   size_t ndiag, uint64_t ntiles, const double **ttab, const double **tphase)
{
    Poly poly;
    window_polynomial(n, diag, ndiag, poly);
    std::vector<double> tt, pt;
    int nh = 0, nb = 0;
    epilogue_tables(prec, n, pos, poly, tt, pt, &nh, &nb);
    void *dtt = nullptr, *dpt = nullptr;
    ZXA_TRY(ctx.ring.stage(tt.data(), tt.size() * sizeof(double), ctx.stream, &dtt));
    ZXA_TRY(ctx.ring.stage(pt.data(), pt.size() * sizeof(double), ctx.stream, &dpt));
    const size_t stride = 2 * (8 + (size_t(1) << nb));
    ZXA_TRY(ctx.scratch_a.ensure(ntiles * stride * sizeof(double)));
    const unsigned grid = unsigned((ntiles + kThreads - 1) / kThreads);
    hipLaunchKernelGGL(tile_phase_kernel, dim3(grid), dim3(kThreads), pt.size() * sizeof(double), ctx.stream,
                       static_cast<const double *>(dpt), int(pos.size()), nh, nb, poly.K, ntiles,
                       static_cast<double *>(ctx.scratch_a.ptr));
    ZXA_HIP(hipGetLastError());
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
        P.tmask |= uint64_t(1) << P.ub[k];
    }
    std::sort(P.sorted, P.sorted + kUnitBits);
    for (int r = 0; r < kRegs; r++)
        for (int i = 0; i < 4; i++)
            if ((r >> i) & 1)
                G.roff[r] |= uint64_t(1) << P.ub[reg_bit(0, i)];
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

// One tile pass over the given targets, then the diagonal window (if any).
int run_tile_pass(Ctx &ctx, void *state, int prec, unsigned n, const std::vector<SingleGate> &gs,
                  const DiagGate *diag = nullptr, size_t ndiag = 0)
{
    const bool c64 = prec == C64;
    Plan pl;
    if (!plan_tile(prec, n, gs, pl))
    {
        set_error("tile pass: state narrower than a tile or too many targets");
        return EINVAL_;
    }
    const std::vector<unsigned> &pos = pl.ebit;
    TPass P;
    TTab G;
    bool use[3];
    fill_pass(prec, n, pl, P, G, use);
    const unsigned shift = c64 ? 1u : 0u; // element bit -> unit-address bit
    bool flip = false;
    for (size_t gi = 0; gi < gs.size(); gi++)
    {
        const SingleGate &s = gs[gi];
        const int slot = pl.slot[gi];
        TGate &tg = G.g[slot];
        const uint64_t alpha = s.g->packed_sel & ((uint64_t(1) << 63) - 1);
        const uint64_t rest = alpha & ~s.g->delta;
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
        for (int r = 0; r < kRegs; r++)
        {
            uint64_t off = 0;
            for (int i = 0; i < 4; i++)
                if ((r >> i) & 1)
                    off |= uint64_t(1) << P.ub[reg_bit(lay, i)];
            tg.rpar |= uint32_t(__builtin_popcountll(off & tg.amask) & 1) << r;
        }
        if (rest != 0)
            flip = true;
    }
    const double *ttab = nullptr;
    const double *tphase = nullptr;
    if (ndiag > 0)
    {
        if (flip)
        {
            set_error("tile pass: diagonal epilogue with per-pair selectors");
            return EINVAL_;
        }
        ZXA_TRY(setup_epilogue(ctx, prec, n, pos, diag, ndiag, P.ntiles, &ttab, &tphase));
    }
    void *dtab = nullptr;
    ZXA_TRY(ctx.ring.stage(&G, sizeof(G), ctx.stream, &dtab));
    const TTab *tab = static_cast<const TTab *>(dtab);
    const uint64_t bytes = 2ull * (uint64_t(1) << n) * elem_bytes(prec);
    ZXA_TRY(ctx.tic());
    if (c64)
        ZXA_TRY(launch_p<true>(ctx, state, P, use[1], use[2], flip, tab, ttab, tphase));
    else
        ZXA_TRY(launch_p<false>(ctx, state, P, use[1], use[2], flip, tab, ttab, tphase));
    ZXA_TRY(ctx.toc(ndiag > 0 ? F_FUSED : F_U3TILE, bytes));
    return ctx.ring.done(ctx.stream);
}

template <bool kC64, bool kL1, bool kL2, bool kDiag>
int launch_adj_t(Ctx &ctx, void *psi, void *lam, const TPass &P0, const TTab *tab, const TAdj *adj,
                 const double *ttab, const double *tphase, double *partial, int *nwaves)
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
    hipLaunchKernelGGL(k, dim3(grid), dim3(kThreads), 0, ctx.stream, psi, lam, P, tab, adj, ttab, tphase, partial);
    return hip_check(hipGetLastError(), "adj_kernel launch");
}

template <bool kC64, bool kDiag>
int launch_adj_l(Ctx &ctx, void *psi, void *lam, const TPass &P, bool l1, bool l2, const TTab *tab,
                 const TAdj *adj, const double *ttab, const double *tphase, double *partial, int *nwaves)
{
    if (l1 && l2)
        return launch_adj_t<kC64, true, true, kDiag>(ctx, psi, lam, P, tab, adj, ttab, tphase, partial, nwaves);
    if (l1)
        return launch_adj_t<kC64, true, false, kDiag>(ctx, psi, lam, P, tab, adj, ttab, tphase, partial, nwaves);
    if (l2)
        return launch_adj_t<kC64, false, true, kDiag>(ctx, psi, lam, P, tab, adj, ttab, tphase, partial, nwaves);
    return launch_adj_t<kC64, false, false, kDiag>(ctx, psi, lam, P, tab, adj, ttab, tphase, partial, nwaves);
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
    uint64_t adj;     // adjoint: TAdj
    uint64_t ttab;    // epilogue per-thread table, 0: no window
    uint64_t tphase;  // epilogue per-tile factors
    uint64_t partial; // adjoint: per tile and wave, kAccs accumulators
    uint32_t ub[kUnitBits];
    uint32_t sorted[kUnitBits];
};

template <bool kC64, bool kAdj>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(2))) void batch_tile_kernel(
    const BItem *__restrict__ items, int nitems, uint64_t total)
{
    using T = typename UnitT<kC64>::T;
    __shared__ T lds[kUnits];
    const int t = threadIdx.x;
    const int lane = t & 63;
    for (uint64_t g = blockIdx.x; g < total; g += gridDim.x)
    {
        const BItem *I0 = items;
        asm volatile("" : "+s"(I0));
        ZXA_CONST BItem *IT = (ZXA_CONST BItem *)I0;
        int lo = 0, hi = nitems - 1;
        while (lo < hi)
        {
            const int mid = (lo + hi + 1) >> 1;
            if (IT[mid].first <= g)
                lo = mid;
            else
                hi = mid - 1;
        }
        ZXA_CONST BItem &I = IT[lo];
        TPass P;
#pragma unroll
        for (int k = 0; k < kUnitBits; k++)
        {
            P.ub[k] = I.ub[k];
            P.sorted[k] = I.sorted[k];
        }
        const uint64_t local = g - I.first;
        const uint64_t base = deposit(local, P.sorted);
        const uint64_t toff = thr_off<0>(t, P);
        ZXA_CONST TTab *G = (ZXA_CONST TTab *)I.tab;
        T *pp = reinterpret_cast<T *>(I.psi);
        T v[kRegs];
        load_tile<kC64>(v, pp, base | toff, G);
        if constexpr (kAdj)
        {
            T *pl = reinterpret_cast<T *>(I.lam);
            ZXA_CONST TAdj *A = (ZXA_CONST TAdj *)I.adj;
            T w[kRegs];
            load_tile<kC64>(w, pl, base | toff, G);
            // one wave-wide value per accumulator, stored by lane 0 at once
            double *pw = reinterpret_cast<double *>(I.partial) + (local * kWavesPerTile + (t >> 6)) * kAccs;
            auto emit = [&](int a, double tot) {
                if (lane == 0)
                    pw[a] = tot;
            };
            adj_layout<kC64, 0>(v, w, G, A, emit);
            transpose<kC64, 0, 1>(v, lds, t);
            transpose<kC64, 0, 1>(w, lds, t);
            adj_layout<kC64, 1>(v, w, G, A, emit);
            transpose<kC64, 1, 2>(v, lds, t);
            transpose<kC64, 1, 2>(w, lds, t);
            adj_layout<kC64, 2>(v, w, G, A, emit);
            transpose<kC64, 2, 0>(v, lds, t);
            transpose<kC64, 2, 0>(w, lds, t);
            if (I.ttab != 0)
            {
                Epi<kC64> epi;
                epi.init(reinterpret_cast<const double *>(I.ttab), t);
                epi.template apply_n<true>(
                    v, w, (ZXA_CONST double *)(reinterpret_cast<const double *>(I.tphase) + local * Epi<kC64>::kStride),
                    t);
            }
            store_tile<kC64>(v, pp, base | toff, G);
            store_tile<kC64>(w, pl, base | toff, G);
        }
        else
        {
            apply_layout<kC64, false, 0>(v, P, G, base, toff);
            transpose<kC64, 0, 1>(v, lds, t);
            apply_layout<kC64, false, 1>(v, P, G, base, toff);
            transpose<kC64, 1, 2>(v, lds, t);
            apply_layout<kC64, false, 2>(v, P, G, base, toff);
            transpose<kC64, 2, 0>(v, lds, t);
            if (I.ttab != 0)
            {
                Epi<kC64> epi;
                epi.init(reinterpret_cast<const double *>(I.ttab), t);
                epi.apply(v, (ZXA_CONST double *)(reinterpret_cast<const double *>(I.tphase) + local * Epi<kC64>::kStride),
                          t);
            }
            store_tile<kC64>(v, pp, base | toff, G);
        }
    }
}

// Per item: the accumulators of its tiles and waves summed in a fixed order,
// then added into grads[param[a]] in order (one thread).
struct BCollect
{
    uint64_t partial;
    uint64_t ntiles;
    uint64_t params;  // int32[kAccs]
    uint64_t grads;   // double*
    int64_t nparams;  // length of grads
};

__global__ __launch_bounds__(64) void batch_collect_kernel(const BCollect *__restrict__ items)
{
    __shared__ double tot[64];
    const BCollect &I = items[blockIdx.x];
    const int j = threadIdx.x;
    double s = 0.0;
    if (j < kAccs)
    {
        const double *pp = reinterpret_cast<const double *>(I.partial);
        const uint64_t nw = I.ntiles * kWavesPerTile;
        for (uint64_t w = 0; w < nw; w++)
            s += pp[w * kAccs + j];
    }
    tot[j] = s;
    __syncthreads();
    if (j == 0)
    {
        const int32_t *pm = reinterpret_cast<const int32_t *>(I.params);
        double *gr = reinterpret_cast<double *>(I.grads);
        for (int a = 0; a < kAccs; a++)
            if (pm[a] >= 0 && pm[a] < I.nparams)
                gr[pm[a]] += tot[a];
    }
}

} // namespace

unsigned tile_seg(int prec)
{
    const int s = tune_get(prec == C64 ? "tile.seg.c64" : "tile.seg.c128");
    return s >= 4 && s <= 6 ? unsigned(s) : 4u;
}

unsigned tile_low_bits(int prec)
{
    return tile_seg(prec) + (prec == C64 ? 1u : 0u);
}

unsigned tile_upper_slots(int prec)
{
    return unsigned(kUnitBits) - tile_seg(prec);
}

unsigned tile_min_bits(int prec)
{
    return tile_low_bits(prec) + tile_upper_slots(prec);
}

int u3_apply(Ctx &ctx, void *state, int prec, unsigned n, const U3Gate *gates, size_t ngates)
{
    const unsigned low = tile_low_bits(prec);
    std::vector<SingleGate> singles;
    for (size_t i = 0; i < ngates; i++)
    {
        const U3Gate &g = gates[i];
        const uint64_t alpha = g.packed_sel & ((uint64_t(1) << 63) - 1);
        const bool one_bit = g.delta != 0 && (g.delta & (g.delta - 1)) == 0 && (g.delta >> n) == 0;
        const bool paired = (__builtin_popcountll(alpha & g.delta) & 1) == 1;
        if (n >= tile_min_bits(prec) && n <= 40 && one_bit && paired && (alpha >> n) == 0)
            singles.push_back({unsigned(__builtin_ctzll(g.delta)), &g});
        else
            ZXA_TRY(u3_pair_apply(ctx, state, prec, n, g));
    }
    if (singles.empty())
        return OK;
    std::sort(singles.begin(), singles.end(),
              [](const SingleGate &a, const SingleGate &b) { return a.target < b.target; });
    for (size_t i = 1; i < singles.size(); i++)
        if (singles[i].target == singles[i - 1].target)
        {
            set_error("u3_apply: two gates on one target in one batch");
            return EINVAL_;
        }
    // the low targets ride with the first pass; the others fill the slots
    std::vector<SingleGate> lows, highs;
    for (const SingleGate &s : singles)
        (s.target < low ? lows : highs).push_back(s);
    size_t hi = 0;
    bool first = true;
    while (first || hi < highs.size())
    {
        std::vector<SingleGate> pass;
        if (first)
            pass = lows;
        const size_t take = std::min<size_t>(tile_upper_slots(prec), highs.size() - hi);
        for (size_t k = 0; k < take; k++)
            pass.push_back(highs[hi + k]);
        hi += take;
        first = false;
        if (!pass.empty())
            ZXA_TRY(run_tile_pass(ctx, state, prec, n, pass));
    }
    return OK;
}

int fused_pass_apply(Ctx &ctx, void *state, int prec, unsigned n, const U3Gate *gates, size_t ngates,
                     const DiagGate *diag, size_t ndiag)
{
    if (n < tile_min_bits(prec) || n > 40)
    {
        set_error("fused pass: state narrower than a tile");
        return EINVAL_;
    }
    if (ndiag > 0 && !window_is_polynomial(n, diag, ndiag))
    {
        set_error("fused pass: diagonal window is not a phase polynomial");
        return EINVAL_;
    }
    const unsigned low = tile_low_bits(prec);
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
    if (unsigned(high) > tile_upper_slots(prec))
    {
        set_error("fused pass: more targets above the tile's low bits than the tile has slots");
        return EINVAL_;
    }
    if (gs.empty())
        return diag_apply(ctx, state, prec, n, diag, ndiag);
    return run_tile_pass(ctx, state, prec, n, gs, diag, ndiag);
}

} // namespace zxa

namespace zxa
{

int adjoint_pass_apply(Ctx &ctx, void *psi, void *lam, int prec, unsigned n, const AdjTarget *targets,
                       size_t ntargets, const DiagGate *diag, size_t ndiag, double *dgrads)
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
    if (!plan_tile(prec, n, gs, pl))
    {
        set_error("adjoint pass: more targets than the tile has slots");
        return EINVAL_;
    }
    TPass P;
    TTab G;
    bool use[3];
    fill_pass(prec, n, pl, P, G, use);
    TAdj A;
    std::memset(&A, 0, sizeof(A));
    std::vector<int> param(kAccs, -1);
    for (size_t i = 0; i < ntargets; i++)
    {
        const int slot = pl.slot[i];
        for (int k = 0; k < 8; k++)
        {
            G.g[slot].u[k] = targets[i].m[k];
            G.g[slot].uf[k] = float(targets[i].m[k]);
        }
        for (int o = 0; o < kObs; o++)
        {
            for (int k = 0; k < 8; k++)
            {
                A.o[slot].a[o][k] = targets[i].a[o][k];
                A.o[slot].af[o][k] = float(targets[i].a[o][k]);
            }
            param[slot * kObs + o] = targets[i].param[o];
        }
    }
    const double *ttab = nullptr;
    const double *tphase = nullptr;
    if (ndiag > 0)
        ZXA_TRY(setup_epilogue(ctx, prec, n, pl.ebit, diag, ndiag, P.ntiles, &ttab, &tphase));
    void *dtab = nullptr, *dadj = nullptr, *dpar = nullptr;
    ZXA_TRY(ctx.ring.stage(&G, sizeof(G), ctx.stream, &dtab));
    ZXA_TRY(ctx.ring.stage(&A, sizeof(A), ctx.stream, &dadj));
    ZXA_TRY(ctx.ring.stage(param.data(), param.size() * sizeof(int), ctx.stream, &dpar));
    // per-wave partials: at most (2 workgroups per CU) x CUs x 4 waves
    const size_t maxw = size_t(8) * size_t(ctx.num_cus > 0 ? ctx.num_cus : 256) * (kThreads / 64);
    ZXA_TRY(ctx.scratch_f.ensure(maxw * kAccs * sizeof(double)));
    double *partial = static_cast<double *>(ctx.scratch_f.ptr);
    const TTab *tab = static_cast<const TTab *>(dtab);
    const TAdj *adj = static_cast<const TAdj *>(dadj);
    const uint64_t bytes = 4ull * (uint64_t(1) << n) * elem_bytes(prec);
    int nwaves = int(maxw);
    ZXA_TRY(ctx.tic());
    int rc;
    if (c64)
        rc = ndiag ? launch_adj_l<true, true>(ctx, psi, lam, P, use[1], use[2], tab, adj, ttab, tphase, partial,
                                              &nwaves)
                   : launch_adj_l<true, false>(ctx, psi, lam, P, use[1], use[2], tab, adj, ttab, tphase, partial,
                                               &nwaves);
    else
        rc = ndiag ? launch_adj_l<false, true>(ctx, psi, lam, P, use[1], use[2], tab, adj, ttab, tphase, partial,
                                               &nwaves)
                   : launch_adj_l<false, false>(ctx, psi, lam, P, use[1], use[2], tab, adj, ttab, tphase, partial,
                                                &nwaves);
    ZXA_TRY(rc);
    ZXA_TRY(ctx.toc(F_ADJ, bytes));
    hipLaunchKernelGGL(adj_collect_kernel, dim3(1), dim3(64), 0, ctx.stream, partial, nwaves,
                       static_cast<const int *>(dpar), dgrads);
    ZXA_HIP(hipGetLastError());
    return ctx.ring.done(ctx.stream);
}

} // namespace zxa

namespace zxa
{

BatchLaunch::~BatchLaunch()
{
    // d_adjs and d_params point into the d_tabs allocation
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
    std::vector<int32_t> params(specs.size() * kAccs, -1);
    // epilogue tables: per item, per-thread table then per-tile factors
    std::vector<size_t> epi_off(specs.size(), 0), epi_tt(specs.size(), 0);
    size_t epi_bytes = 0;
    const int nb = c64 ? 5 : 4;
    const size_t ttab_doubles = size_t(kThreads) * 2 * (1 + nb);
    const size_t stride = 2 * (8 + (size_t(1) << nb));
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
        if (!plan_tile(prec, sp.n, gs, plans[i]))
        {
            set_error("batch: pass does not fit one tile");
            return EINVAL_;
        }
        TPass P;
        bool use[3];
        fill_pass(prec, sp.n, plans[i], P, tabs[i], use);
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
            partial_doubles += P.ntiles * kWavesPerTile * kAccs;
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
    b.adj_bytes = adj ? specs.size() * sizeof(TAdj) : 0;
    b.par_bytes = adj ? params.size() * sizeof(int32_t) : 0;
    char *dtabs = nullptr;
    ZXA_TRY(dmalloc(&dtabs, b.tab_bytes + b.adj_bytes + b.par_bytes));
    b.d_tabs = dtabs;
    b.d_adjs = adj ? dtabs + b.tab_bytes : nullptr;
    b.d_params = adj ? dtabs + b.tab_bytes + b.adj_bytes : nullptr;
    ZXA_HIP(hipHostMalloc(&b.h_tables, b.tab_bytes + b.adj_bytes + b.par_bytes, hipHostMallocDefault));
    std::memcpy(b.h_tables, tabs.data(), b.tab_bytes);
    if (adj)
    {
        std::memset(static_cast<char *>(b.h_tables) + b.tab_bytes, 0, b.adj_bytes);
        std::memcpy(static_cast<char *>(b.h_tables) + b.tab_bytes + b.adj_bytes, params.data(), b.par_bytes);
    }
    // the device tables start valid (identity gates, no parameters), so a
    // launch before the first set/upload is harmless
    ZXA_HIP(hipMemcpy(dtabs, b.h_tables, b.tab_bytes + b.adj_bytes + b.par_bytes, hipMemcpyHostToDevice));
    char *depi = nullptr;
    if (epi_bytes)
        ZXA_TRY(dmalloc(&depi, epi_bytes));
    b.d_epi = depi;
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
        {
            it.adj = reinterpret_cast<uint64_t>(dtabs + b.tab_bytes + i * sizeof(TAdj));
            it.partial = reinterpret_cast<uint64_t>(dpart + it.partial);
        }
        const BatchPassSpec &sp = specs[i];
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
        const unsigned grid = unsigned((it.ntiles + kThreads - 1) / kThreads);
        hipLaunchKernelGGL(tile_phase_kernel, dim3(grid), dim3(kThreads), pt.size() * sizeof(double), ctx.stream, dpt,
                           int(plans[i].ebit.size()), nh, nb2, poly.K, it.ntiles,
                           reinterpret_cast<double *>(depi + epi_off[i]));
        hipError_t e = hipGetLastError();
        if (e == hipSuccess)
            e = hipStreamSynchronize(ctx.stream);
        (void)hipFree(dpt);
        ZXA_HIP(e);
        it.ttab = reinterpret_cast<uint64_t>(depi + epi_tt[i]);
        it.tphase = reinterpret_cast<uint64_t>(depi + epi_off[i]);
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
            col[i].params =
                reinterpret_cast<uint64_t>(dtabs + b.tab_bytes + b.adj_bytes + i * kAccs * sizeof(int32_t));
            col[i].grads = reinterpret_cast<uint64_t>(specs[i].grads);
            col[i].nparams = int64_t(specs[i].nparams);
        }
        ZXA_TRY(dmalloc(&b.d_collect, col.size() * sizeof(BCollect)));
        ZXA_HIP(hipMemcpy(b.d_collect, col.data(), col.size() * sizeof(BCollect), hipMemcpyHostToDevice));
    }
    auto k = c64 ? (adj ? batch_tile_kernel<true, true> : batch_tile_kernel<true, false>)
                 : (adj ? batch_tile_kernel<false, true> : batch_tile_kernel<false, false>);
    b.grid = persistent_grid(ctx, k, b.total_tiles);
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

void batch_set_adjoint(BatchLaunch &b, size_t item, const AdjTarget *targets, size_t ntargets)
{
    TTab *tab = reinterpret_cast<TTab *>(b.h_tables) + item;
    TAdj *adj = reinterpret_cast<TAdj *>(static_cast<char *>(b.h_tables) + b.tab_bytes) + item;
    int32_t *par = reinterpret_cast<int32_t *>(static_cast<char *>(b.h_tables) + b.tab_bytes + b.adj_bytes) +
                   item * kAccs;
    for (size_t k = 0; k < ntargets && k < b.slots[item].size(); k++)
    {
        const int slot = b.slots[item][k];
        TGate &tg = tab->g[slot];
        for (int i = 0; i < 8; i++)
        {
            tg.u[i] = targets[k].m[i];
            tg.uf[i] = float(targets[k].m[i]);
        }
        for (int o = 0; o < kObs; o++)
        {
            for (int i = 0; i < 8; i++)
            {
                adj->o[slot].a[o][i] = targets[k].a[o][i];
                adj->o[slot].af[o][i] = float(targets[k].a[o][i]);
            }
            par[slot * kObs + o] = targets[k].param[o];
        }
    }
}

int batch_upload(Ctx &ctx, BatchLaunch &b)
{
    if (!b.nitems)
        return OK;
    return hip_check(hipMemcpyAsync(b.d_tabs, b.h_tables, b.tab_bytes + b.adj_bytes + b.par_bytes,
                                    hipMemcpyHostToDevice, ctx.stream),
                     "batch table upload");
}

int batch_launch(Ctx &ctx, BatchLaunch &b)
{
    if (!b.nitems)
        return OK;
    const bool c64 = b.prec == C64;
    auto k = c64 ? (b.adj ? batch_tile_kernel<true, true> : batch_tile_kernel<true, false>)
                 : (b.adj ? batch_tile_kernel<false, true> : batch_tile_kernel<false, false>);
    hipLaunchKernelGGL(k, dim3(b.grid), dim3(kThreads), 0, ctx.stream, static_cast<const BItem *>(b.d_items),
                       int(b.nitems), b.total_tiles);
    ZXA_HIP(hipGetLastError());
    if (b.adj)
    {
        hipLaunchKernelGGL(batch_collect_kernel, dim3(unsigned(b.nitems)), dim3(64), 0, ctx.stream,
                           static_cast<const BCollect *>(b.d_collect));
        ZXA_HIP(hipGetLastError());
    }
    return OK;
}

} // namespace zxa
