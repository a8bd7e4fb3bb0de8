/*
 * zxh_oracle.c -- CPU restatement of ZXH-Sim's CPU path (TEST INFRASTRUCTURE).
 *
 * This file is the oracle: tests, the smoke check and the benchmark's CPU
 * baseline leg are its only users.  The product never links or calls it.
 * Each function cites the reference code whose behaviour it restates
 * (paths relative to the reference checkout, src/zxh/...).
 *
 * Build: gcc -O3 -std=c11 -fopenmp -ffp-contract=off (see oracle/Makefile).
 * Contraction must stay off: the reference's x86-64 build has no FMA, so
 * every product and sum below is rounded separately, as there.
 */
#define _GNU_SOURCE
#include "zxh_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ZO_PI 3.14159265358979323846

/* ------------------------------------------------------------------------ */
/* complex helpers: (a+bi)(c+di) = (ac-bd) + (ad+bc)i, as GCC's std::complex */

typedef struct {
    double re, im;
} cplx;

static inline cplx c_mk(double re, double im)
{
    cplx r = {re, im};
    return r;
}
static inline cplx c_mul(cplx a, cplx b)
{
    return c_mk(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
static inline cplx c_add(cplx a, cplx b)
{
    return c_mk(a.re + b.re, a.im + b.im);
}
static inline cplx c_scale(cplx a, double s)
{
    return c_mk(a.re * s, a.im * s);
}
/* std::exp(complex(0, angle)) == (cos angle, sin angle) (zxh.cpp:31-34) */
static inline cplx c_phase(double angle)
{
    return c_mk(cos(angle), sin(angle));
}

/* ------------------------------------------------------------------------ */
/* selectors: packed = alpha | beta<<63, S(x) = parity(x & alpha) ^ beta    */
/* (selector.cpp:26-40, kernel_general.cpp:59-65)                           */

static inline uint64_t sel_pack(uint64_t alpha, int beta)
{
    return (alpha & ((UINT64_C(1) << 63) - 1)) | (beta ? (UINT64_C(1) << 63) : 0);
}
static inline uint64_t sel_alpha(uint64_t packed)
{
    return packed & ((UINT64_C(1) << 63) - 1);
}
static inline int sel_beta(uint64_t packed)
{
    return (int)((packed >> 63) & 1u);
}
static inline int sel_eval(uint64_t packed, uint64_t x)
{
    const int parity = __builtin_popcountll(x & sel_alpha(packed)) & 1;
    return parity != sel_beta(packed);
}

/* Selector evaluation where the reference calls selector_t::eval/eval_u64
 * (diagonal window, X, CX and CP kernels; kernel_omp.cpp:84-109, 155, 177,
 * 387).  Those are defined in their own translation unit (selector.cpp:26-38)
 * and the reference builds without interprocedural optimisation
 * (CMakeLists.txt:483), so every evaluation is a call.  The CPU-baseline build
 * (-DZO_REFERENCE_COST, oracle/liboracle_base.so) keeps that call
 * (zo_selector.c) so that its timing matches the reference's; the checker
 * build inlines it.  Both give identical results. */
#ifdef ZO_REFERENCE_COST
int zo_sel_eval_call(uint64_t packed, uint64_t x);
#define SEL_EVAL_CALL(p, x) zo_sel_eval_call((p), (x))
#else
#define SEL_EVAL_CALL(p, x) sel_eval((p), (x))
#endif

static inline uint64_t low_mask(size_t nbits)
{
    if (nbits == 0)
        return 0;
    if (nbits >= 64)
        return ~UINT64_C(0);
    return (UINT64_C(1) << nbits) - 1;
}

/* ------------------------------------------------------------------------ */
/* kernels (kernel_general.cpp / kernel_omp.cpp)                            */

/* pair mix for every i with S(i) = 0 and j = i ^ delta < n
 * (kernel_general.cpp:121-141, 173-202) */
static void u3_pairs(double *psi, size_t n, uint64_t delta, uint64_t packed, const double u[8])
{
    const cplx u00 = c_mk(u[0], u[1]), u01 = c_mk(u[2], u[3]);
    const cplx u10 = c_mk(u[4], u[5]), u11 = c_mk(u[6], u[7]);
    cplx *p = (cplx *)psi;
    const long long n64 = (long long)n;
#pragma omp parallel for schedule(static)
    for (long long ii = 0; ii < n64; ii++) {
        const uint64_t i = (uint64_t)ii;
        if (sel_eval(packed, i))
            continue;
        const uint64_t j = i ^ delta;
        if (j >= n)
            continue;
        const cplx even = p[i], odd = p[j];
        p[i] = c_add(c_mul(u00, even), c_mul(u01, odd));
        p[j] = c_add(c_mul(u10, even), c_mul(u11, odd));
    }
}

void zo_k_u3_batch(double *psi, size_t local_bits, const zo_u3desc *descs, size_t ndesc)
{
    const size_t n = (size_t)1 << local_bits;
    for (size_t d = 0; d < ndesc; d++)
        u3_pairs(psi, n, descs[d].delta, descs[d].packed_sel, descs[d].u);
}

/* kernel_general.cpp:143-156 */
void zo_k_x(double *psi, size_t local_bits, uint64_t delta, uint64_t packed_sel)
{
    const size_t n = (size_t)1 << local_bits;
    cplx *p = (cplx *)psi;
    for (size_t off = 0; off < n; off++) {
        if (SEL_EVAL_CALL(packed_sel, off))
            continue;
        const size_t peer = off ^ delta;
        if (peer >= n)
            continue;
        const cplx t = p[off];
        p[off] = p[peer];
        p[peer] = t;
    }
}

/* kernel_general.cpp:158-171 */
void zo_k_cx(double *psi, size_t local_bits, uint64_t delta, uint64_t packed_cq, uint64_t packed_q)
{
    const size_t n = (size_t)1 << local_bits;
    cplx *p = (cplx *)psi;
    for (size_t off = 0; off < n; off++) {
        if (SEL_EVAL_CALL(packed_q, off) || !SEL_EVAL_CALL(packed_cq, off))
            continue;
        const size_t peer = off ^ delta;
        if (peer >= n)
            continue;
        const cplx t = p[off];
        p[off] = p[peer];
        p[peer] = t;
    }
}

/* kernel_general.cpp:347-356 */
void zo_k_cp(double *psi, size_t local_bits, uint64_t packed_cq, uint64_t packed_q, double theta)
{
    const size_t n = (size_t)1 << local_bits;
    cplx *p = (cplx *)psi;
    const cplx ph = c_phase(theta);
    for (size_t off = 0; off < n; off++) {
        if (SEL_EVAL_CALL(packed_cq, off) && SEL_EVAL_CALL(packed_q, off))
            p[off] = c_mul(p[off], ph);
    }
}

/* diag window buckets (kernel_general.cpp:12-51, 278-345) */
enum { B_PL = 0, B_PH, B_PX, B_CPLL, B_CPHH, B_CPHL, B_COUNT };
typedef struct {
    uint64_t s0, s1;
    double theta;
} diag_ent;
typedef struct {
    diag_ent *e[B_COUNT];
    size_t n[B_COUNT], cap[B_COUNT];
} diag_window;

enum { CLS_LOW = 0, CLS_HIGH = 1, CLS_CROSS = 2 };
static int classify(uint64_t packed)
{
    const uint64_t a = sel_alpha(packed);
    if ((a >> 8) == 0)
        return CLS_LOW;
    if ((a & 0xffu) == 0)
        return CLS_HIGH;
    return CLS_CROSS;
}

static int dw_push(diag_window *w, int b, uint64_t s0, uint64_t s1, double th)
{
    if (w->n[b] == w->cap[b]) {
        size_t nc = w->cap[b] ? 2 * w->cap[b] : 16;
        diag_ent *ne = (diag_ent *)realloc(w->e[b], nc * sizeof(diag_ent));
        if (!ne)
            return ZO_ERR_ALLOC;
        w->e[b] = ne;
        w->cap[b] = nc;
    }
    w->e[b][w->n[b]].s0 = s0;
    w->e[b][w->n[b]].s1 = s1;
    w->e[b][w->n[b]].theta = th;
    w->n[b]++;
    return ZO_OK;
}
static void dw_reset(diag_window *w)
{
    for (int b = 0; b < B_COUNT; b++)
        w->n[b] = 0;
}
static void dw_free(diag_window *w)
{
    for (int b = 0; b < B_COUNT; b++) {
        free(w->e[b]);
        w->e[b] = NULL;
        w->n[b] = w->cap[b] = 0;
    }
}
static int dw_empty(const diag_window *w)
{
    for (int b = 0; b < B_COUNT; b++)
        if (w->n[b])
            return 0;
    return 1;
}
/* diag_pending_push_p (kernel_general.cpp:288-302) */
static int dw_push_p(diag_window *w, uint64_t sel, double th)
{
    static const int map[3] = {B_PL, B_PH, B_PX};
    return dw_push(w, map[classify(sel)], sel, 0, th);
}
/* diag_pending_try_push_cp (kernel_general.cpp:304-330); returns 1 if taken */
static int dw_try_push_cp(diag_window *w, uint64_t scq, uint64_t sq, double th)
{
    const int c0 = classify(scq), c1 = classify(sq);
    if (c0 == CLS_LOW && c1 == CLS_LOW)
        return dw_push(w, B_CPLL, scq, sq, th) == ZO_OK;
    if (c0 == CLS_HIGH && c1 == CLS_HIGH)
        return dw_push(w, B_CPHH, scq, sq, th) == ZO_OK;
    if (c0 == CLS_LOW && c1 == CLS_HIGH)
        return dw_push(w, B_CPHL, scq, sq, th) == ZO_OK;
    if (c0 == CLS_HIGH && c1 == CLS_LOW)
        return dw_push(w, B_CPHL, sq, scq, th) == ZO_OK;
    return 0;
}
/* apply_diag_window (kernel_general.cpp:72-113): theta summed in bucket order */
static void dw_apply(const diag_window *w, double *psi, size_t n)
{
    cplx *p = (cplx *)psi;
    const long long n64 = (long long)n;
#pragma omp parallel for schedule(static)
    for (long long ii = 0; ii < n64; ii++) {
        const uint64_t xr = (uint64_t)ii;
        const uint64_t lane = xr & 0xffu;
        const uint64_t slice = xr & ~(uint64_t)0xffu;
        double th = 0.0;
        for (size_t k = 0; k < w->n[B_PL]; k++)
            if (SEL_EVAL_CALL(w->e[B_PL][k].s0, lane))
                th += w->e[B_PL][k].theta;
        for (size_t k = 0; k < w->n[B_PH]; k++)
            if (SEL_EVAL_CALL(w->e[B_PH][k].s0, slice))
                th += w->e[B_PH][k].theta;
        for (size_t k = 0; k < w->n[B_PX]; k++)
            if (SEL_EVAL_CALL(w->e[B_PX][k].s0, xr))
                th += w->e[B_PX][k].theta;
        for (size_t k = 0; k < w->n[B_CPLL]; k++)
            if (SEL_EVAL_CALL(w->e[B_CPLL][k].s0, lane) && SEL_EVAL_CALL(w->e[B_CPLL][k].s1, lane))
                th += w->e[B_CPLL][k].theta;
        for (size_t k = 0; k < w->n[B_CPHH]; k++)
            if (SEL_EVAL_CALL(w->e[B_CPHH][k].s0, slice) && SEL_EVAL_CALL(w->e[B_CPHH][k].s1, slice))
                th += w->e[B_CPHH][k].theta;
        for (size_t k = 0; k < w->n[B_CPHL]; k++)
            if (SEL_EVAL_CALL(w->e[B_CPHL][k].s0, lane) && SEL_EVAL_CALL(w->e[B_CPHL][k].s1, slice))
                th += w->e[B_CPHL][k].theta;
        if (th != 0.0)
            p[ii] = c_mul(p[ii], c_phase(th));
    }
}

int zo_k_diag(double *psi, size_t local_bits, const uint64_t *sel0, const uint64_t *sel1, const double *theta,
              size_t ngates)
{
    diag_window w;
    memset(&w, 0, sizeof(w));
    int rc = ZO_OK;
    for (size_t g = 0; g < ngates && rc == ZO_OK; g++) {
        if (sel1[g] == ZO_NO_SEL)
            rc = dw_push_p(&w, sel0[g], theta[g]);
        else if (!dw_try_push_cp(&w, sel0[g], sel1[g], theta[g]))
            rc = ZO_ERR_ARG;
    }
    if (rc == ZO_OK && !dw_empty(&w))
        dw_apply(&w, psi, (size_t)1 << local_bits);
    dw_free(&w);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* U3 matrices and host fusion (zxh.cpp:101-172)                            */

void zo_u3_matrix(double theta, double lambda, double phi, double out[8])
{
    const double c = cos(theta / 2.0);
    const double s = sin(theta / 2.0);
    const cplx el = c_phase(lambda), ep = c_phase(phi), epl = c_phase(phi + lambda);
    const cplx u01 = c_scale(c_mk(-el.re, -el.im), s);
    const cplx u10 = c_scale(ep, s);
    const cplx u11 = c_scale(epl, c);
    out[0] = c;
    out[1] = 0.0;
    out[2] = u01.re;
    out[3] = u01.im;
    out[4] = u10.re;
    out[5] = u10.im;
    out[6] = u11.re;
    out[7] = u11.im;
}

/* kU3BatchEps is the float literal 1e-5f widened to double (zxh.cpp:29) */
static const double kU3Eps = (double)1e-5f;

typedef struct {
    int active;
    cplx u00, u01, u10, u11;
} u3acc;

static void acc_reset(u3acc *a)
{
    a->active = 0;
    a->u00 = c_mk(1.0, 0.0);
    a->u01 = c_mk(0.0, 0.0);
    a->u10 = c_mk(0.0, 0.0);
    a->u11 = c_mk(1.0, 0.0);
}
/* std::abs(a - b) <= eps, std::abs(complex) == hypot (zxh.cpp:115-118) */
static int close_val(cplx a, cplx b)
{
    return hypot(a.re - b.re, a.im - b.im) <= kU3Eps;
}
static int acc_is_identity(const u3acc *a)
{
    return close_val(a->u00, c_mk(1.0, 0.0)) && close_val(a->u01, c_mk(0.0, 0.0)) &&
           close_val(a->u10, c_mk(0.0, 0.0)) && close_val(a->u11, c_mk(1.0, 0.0));
}
/* left_multiply_u3 (zxh.cpp:144-156) */
static void acc_left_mul(u3acc *a, const double n[8])
{
    const cplx n00 = c_mk(n[0], n[1]), n01 = c_mk(n[2], n[3]), n10 = c_mk(n[4], n[5]), n11 = c_mk(n[6], n[7]);
    const cplx o00 = a->u00, o01 = a->u01, o10 = a->u10, o11 = a->u11;
    a->u00 = c_add(c_mul(n00, o00), c_mul(n01, o10));
    a->u01 = c_add(c_mul(n00, o01), c_mul(n01, o11));
    a->u10 = c_add(c_mul(n10, o00), c_mul(n11, o10));
    a->u11 = c_add(c_mul(n10, o01), c_mul(n11, o11));
    a->active = !acc_is_identity(a);
}

/* ------------------------------------------------------------------------ */
/* simulator state                                                          */

typedef struct {
    int type;
    size_t cq, q;
    double theta, lambda, phi;
} zo_gate;

struct zo_sim {
    size_t N;
    int disable_x, eager;
    zo_gate *gates;
    size_t ngates, gcap;
    int use_seed;
    uint64_t seed;
    cplx gphase;
    /* affine map: rows[q] bit c = A[q][c]; b[q] */
    uint64_t *rows;
    size_t m;
    uint8_t *b;
    /* state (single rank: block_start = 0) */
    size_t M, I;
    cplx *psi;
    size_t alloc_elems;
    size_t used_bits;
    size_t active_len;
    int executed;
    int fatal;
    size_t seg_min_i;
    /* measure cache */
    uint64_t *real_results;
    uint64_t *packed;
    size_t result_count;
};

zo_sim *zo_create(size_t n, int disable_x, int eager)
{
    zo_sim *s = (zo_sim *)calloc(1, sizeof(zo_sim));
    if (!s)
        return NULL;
    s->N = n;
    s->disable_x = disable_x;
    s->eager = eager;
    s->gphase = c_mk(1.0, 0.0);
    s->rows = (uint64_t *)calloc(n ? n : 1, sizeof(uint64_t));
    s->b = (uint8_t *)calloc(n ? n : 1, 1);
    s->seg_min_i = 18;
    s->alloc_elems = 1;
    s->psi = (cplx *)calloc(1, sizeof(cplx));
    s->psi[0] = c_mk(1.0, 0.0);
    s->active_len = 1;
    return s;
}

void zo_destroy(zo_sim *s)
{
    if (!s)
        return;
    free(s->gates);
    free(s->rows);
    free(s->b);
    free(s->psi);
    free(s->real_results);
    free(s->packed);
    free(s);
}

void zo_set_seed(zo_sim *s, uint64_t seed)
{
    s->use_seed = 1;
    s->seed = seed;
}
void zo_clear_seed(zo_sim *s)
{
    s->use_seed = 0;
    s->seed = 0;
}
void zo_clear_gates(zo_sim *s)
{
    s->ngates = 0;
    s->gphase = c_mk(1.0, 0.0);
}
void zo_set_seg_min_i(zo_sim *s, size_t min_i)
{
    s->seg_min_i = min_i ? min_i : 18;
}
size_t zo_num_qubits(const zo_sim *s)
{
    return s->N;
}
void zo_global_phase(const zo_sim *s, double *re, double *im)
{
    *re = s->gphase.re;
    *im = s->gphase.im;
}

static int push_gate(zo_sim *s, int type, size_t cq, size_t q, double th, double la, double ph)
{
    if (s->ngates == s->gcap) {
        size_t nc = s->gcap ? 2 * s->gcap : 64;
        zo_gate *ng = (zo_gate *)realloc(s->gates, nc * sizeof(zo_gate));
        if (!ng)
            return ZO_ERR_ALLOC;
        s->gates = ng;
        s->gcap = nc;
    }
    zo_gate g = {type, cq, q, th, la, ph};
    s->gates[s->ngates++] = g;
    return ZO_OK;
}

/* recording rules of ZXH::Barrier..Rx (zxh.cpp:234-310) */
int zo_add_gate(zo_sim *s, int type, size_t cq, size_t q, double theta, double lambda, double phi)
{
    const size_t N = s->N;
    switch (type) {
    case ZO_BARRIER:
        return push_gate(s, ZO_BARRIER, 0, 0, 0.0, 0.0, 0.0);
    case ZO_RZ:
    case ZO_P:
    case ZO_Z:
    case ZO_X:
    case ZO_H:
        if (q >= N)
            return ZO_ERR_RANGE;
        return push_gate(s, type, 0, q, type == ZO_RZ || type == ZO_P ? theta : 0.0, 0.0, 0.0);
    case ZO_U3:
        if (q >= N)
            return ZO_ERR_RANGE;
        return push_gate(s, ZO_U3, 0, q, theta, lambda, phi);
    case ZO_RX:
        if (q >= N)
            return ZO_ERR_RANGE;
        return push_gate(s, ZO_U3, 0, q, theta, ZO_PI / 2.0, -ZO_PI / 2.0);
    case ZO_CRZ: {
        if (cq >= N || q >= N)
            return ZO_ERR_RANGE;
        int rc = push_gate(s, ZO_P, 0, cq, -0.5 * theta, 0.0, 0.0);
        if (rc)
            return rc;
        return push_gate(s, ZO_CP, cq, q, theta, 0.0, 0.0);
    }
    case ZO_CP:
        if (cq >= N || q >= N)
            return ZO_ERR_RANGE;
        return push_gate(s, ZO_CP, cq, q, theta, 0.0, 0.0);
    case ZO_CX:
        if (cq >= N || q >= N)
            return ZO_ERR_RANGE;
        return push_gate(s, ZO_CX, cq, q, 0.0, 0.0, 0.0);
    default:
        return ZO_ERR_ARG;
    }
}

/* ---- GF(2) (bitmat.cpp) with rows stored as m-bit masks (m <= 63) ---- */

/* Solve A x = e_q.  A has full column rank (columns are only appended when
 * independent and rows only combined by invertible row_xor), so the solution
 * is unique and equals the reference's Gauss-Jordan result (bitmat.cpp:46-122). */
static int solve_eq(const uint64_t *rows, size_t N, size_t m, size_t q, uint64_t *x)
{
    if (m == 0) {
        *x = 0;
        return 0; /* rhs = e_q != 0 */
    }
    /* Gaussian elimination on [rows | rhs] */
    uint64_t *r = (uint64_t *)malloc(N * sizeof(uint64_t));
    uint8_t *rhs = (uint8_t *)calloc(N, 1);
    if (!r || !rhs) {
        free(r);
        free(rhs);
        return 0;
    }
    memcpy(r, rows, N * sizeof(uint64_t));
    rhs[q] = 1;
    size_t row = 0;
    int pivot_row[64];
    for (size_t c = 0; c < m; c++)
        pivot_row[c] = -1;
    for (size_t c = 0; c < m && row < N; c++) {
        size_t piv = row;
        while (piv < N && !((r[piv] >> c) & 1u))
            piv++;
        if (piv == N)
            continue;
        if (piv != row) {
            uint64_t t = r[piv];
            r[piv] = r[row];
            r[row] = t;
            uint8_t tb = rhs[piv];
            rhs[piv] = rhs[row];
            rhs[row] = tb;
        }
        for (size_t k = 0; k < N; k++) {
            if (k != row && ((r[k] >> c) & 1u)) {
                r[k] ^= r[row];
                rhs[k] ^= rhs[row];
            }
        }
        pivot_row[c] = (int)row;
        row++;
    }
    int ok = 1;
    for (size_t k = 0; k < N; k++)
        if (r[k] == 0 && rhs[k]) {
            ok = 0;
            break;
        }
    uint64_t sol = 0;
    if (ok)
        for (size_t c = 0; c < m; c++)
            if (pivot_row[c] >= 0 && rhs[pivot_row[c]])
                sol |= UINT64_C(1) << c;
    free(r);
    free(rhs);
    *x = sol;
    return ok;
}

/* calc_mem (zxh.cpp:648-681) */
static size_t calc_mem(const zo_sim *s)
{
    const size_t N = s->N;
    uint64_t *rows = (uint64_t *)calloc(N ? N : 1, sizeof(uint64_t));
    size_t m = 0;
    for (size_t gi = 0; gi < s->ngates; gi++) {
        const zo_gate *g = &s->gates[gi];
        int needs_solve = 0;
        if (g->type == ZO_X || g->type == ZO_CX) {
            if (!s->disable_x) {
                if (g->type == ZO_CX)
                    rows[g->q] ^= rows[g->cq];
            } else {
                needs_solve = 1;
            }
        } else if (g->type == ZO_H || g->type == ZO_U3) {
            needs_solve = 1;
        }
        if (needs_solve) {
            uint64_t d;
            if (!solve_eq(rows, N, m, g->q, &d)) {
                rows[g->q] |= UINT64_C(1) << m;
                m++;
            }
        }
    }
    free(rows);
    return m;
}

size_t zo_required_M(const zo_sim *s)
{
    return calc_mem(s);
}

/* choose_layout for one process (sv_common.cpp:87-100) */
static size_t choose_I(size_t M, size_t min_i)
{
    return M <= min_i ? M : min_i;
}

/* sv_t::resize + reset (sv_common.cpp:286-328) */
static int sv_resize_reset(zo_sim *s, size_t M)
{
    if (M > 40)
        return ZO_ERR_ALLOC;
    s->M = M;
    s->I = choose_I(M, s->seg_min_i);
    const size_t cap = (size_t)1 << M;
    if (s->alloc_elems < cap) {
        cplx *np = (cplx *)realloc(s->psi, cap * sizeof(cplx));
        if (!np)
            return ZO_ERR_ALLOC;
        s->psi = np;
        s->alloc_elems = cap;
    }
    s->used_bits = 0;
    memset(s->psi, 0, s->alloc_elems * sizeof(cplx));
    s->psi[0] = c_mk(1.0, 0.0);
    s->active_len = 1;
    return ZO_OK;
}

/* sv_t::expand (sv_common.cpp:330-349), one rank */
static void sv_expand(zo_sim *s)
{
    if (s->used_bits >= s->M) { /* the reference aborts: "expand exceeds max_bits" */
        s->fatal = ZO_ERR_STATE;
        return;
    }
    s->used_bits++;
    s->active_len = (size_t)1 << s->used_bits;
}

static size_t active_bits(const zo_sim *s)
{
    size_t len = s->active_len, bits = 0;
    while (len > 1) {
        len >>= 1;
        bits++;
    }
    return bits;
}

/* make_selector + make_local_selector with block_start = 0 (zxh.cpp:55-60, 683-691) */
static uint64_t local_selector(const zo_sim *s, size_t q)
{
    const uint64_t alpha = s->rows[q] & low_mask(active_bits(s));
    return sel_pack(alpha, s->b[q]);
}

/* solve_expand (zxh.cpp:701-719) */
static uint64_t solve_expand(zo_sim *s, size_t k)
{
    uint64_t delta;
    if (!solve_eq(s->rows, s->N, s->m, k, &delta)) {
        sv_expand(s);
        s->rows[k] |= UINT64_C(1) << s->m;
        s->m++;
        delta = UINT64_C(1) << (s->m - 1);
    }
    return delta;
}

/* flush_u3_batch (zxh.cpp:390-428): one pair sweep per active accumulator,
 * in qubit order 0..N-1 */
static void flush_u3_batch(zo_sim *s, u3acc *pend, size_t *pend_count)
{
    if (*pend_count == 0)
        return;
    const size_t lb = active_bits(s);
    for (size_t q = 0; q < s->N; q++) {
        if (!pend[q].active)
            continue;
        uint64_t d = 0;
        solve_eq(s->rows, s->N, s->m, q, &d);
        const double u[8] = {pend[q].u00.re, pend[q].u00.im, pend[q].u01.re, pend[q].u01.im,
                             pend[q].u10.re, pend[q].u10.im, pend[q].u11.re, pend[q].u11.im};
        u3_pairs((double *)s->psi, s->active_len, d & low_mask(lb), local_selector(s, q), u);
    }
    for (size_t q = 0; q < s->N; q++)
        acc_reset(&pend[q]);
    *pend_count = 0;
}

/* apply_u3, one rank (zxh.cpp:782-825): solve_expand, then one pair sweep */
static void apply_u3_eager(zo_sim *s, size_t q, double th, double la, double ph)
{
    const uint64_t d = solve_expand(s, q);
    const size_t lb = active_bits(s);
    double u[8];
    zo_u3_matrix(th, la, ph, u);
    u3_pairs((double *)s->psi, s->active_len, d & low_mask(lb), local_selector(s, q), u);
}

/* absorb_u3_gate (zxh.cpp:363-388) */
static void absorb_u3(u3acc *acc, size_t *pend_count, double th, double la, double ph)
{
    double u[8];
    zo_u3_matrix(th, la, ph, u);
    const int was = acc->active;
    if (!was) {
        acc_reset(acc);
        acc->active = 1;
    }
    acc_left_mul(acc, u);
    if (!was && acc->active)
        (*pend_count)++;
    else if (was && !acc->active)
        (*pend_count)--;
}

static void flush_diag(zo_sim *s, diag_window *w)
{
    if (!dw_empty(w)) {
        dw_apply(w, (double *)s->psi, s->active_len);
        dw_reset(w);
    }
}

/* ZXH::execute (zxh.cpp:312-553), one rank */
int zo_execute(zo_sim *s)
{
    const size_t N = s->N;
    const size_t target_M = s->eager ? N : calc_mem(s);
    int rc = sv_resize_reset(s, target_M);
    if (rc)
        return rc;
    s->gphase = c_mk(1.0, 0.0);
    memset(s->rows, 0, (N ? N : 1) * sizeof(uint64_t));
    memset(s->b, 0, N ? N : 1);
    s->m = 0;
    if (s->eager) { /* expand_all (zxh.cpp:721-729) */
        while (s->used_bits < N)
            sv_expand(s);
        for (size_t q = 0; q < N; q++)
            s->rows[q] = UINT64_C(1) << q;
        s->m = N;
    }
    s->executed = 1;
    s->fatal = 0;
    s->result_count = 0;

    diag_window w;
    memset(&w, 0, sizeof(w));
    u3acc *pend = (u3acc *)malloc((N ? N : 1) * sizeof(u3acc));
    if (!pend)
        return ZO_ERR_ALLOC;
    for (size_t q = 0; q < N; q++)
        acc_reset(&pend[q]);
    size_t pend_count = 0;

    size_t gi = 0;
    while (gi < s->ngates) {
        const zo_gate *g = &s->gates[gi];
        if (g->type == ZO_H || g->type == ZO_U3) {
            /* U3 cluster (zxh.cpp:440-470) */
            flush_diag(s, &w);
            while (gi < s->ngates && (s->gates[gi].type == ZO_H || s->gates[gi].type == ZO_U3)) {
                const zo_gate *ug = &s->gates[gi];
                const double th = ug->type == ZO_H ? ZO_PI / 2.0 : ug->theta;
                const double la = ug->type == ZO_H ? ZO_PI : ug->lambda;
                const double ph = ug->type == ZO_H ? 0.0 : ug->phi;
                uint64_t delta;
                if (!solve_eq(s->rows, N, s->m, ug->q, &delta)) {
                    flush_u3_batch(s, pend, &pend_count);
                    apply_u3_eager(s, ug->q, th, la, ph);
                } else {
                    absorb_u3(&pend[ug->q], &pend_count, th, la, ph);
                }
                gi++;
            }
            flush_u3_batch(s, pend, &pend_count);
            continue;
        }
        switch (g->type) {
        case ZO_BARRIER:
            flush_diag(s, &w);
            break;
        case ZO_Z:
        case ZO_RZ: { /* zxh.cpp:477-493 */
            const double th = g->type == ZO_Z ? ZO_PI : g->theta;
            dw_push_p(&w, local_selector(s, g->q), th);
            s->gphase = c_mul(s->gphase, c_phase(-0.5 * th));
            break;
        }
        case ZO_P:
            dw_push_p(&w, local_selector(s, g->q), g->theta);
            break;
        case ZO_X:
            if (s->disable_x) { /* apply_x_explicit (zxh.cpp:742-759) */
                flush_diag(s, &w);
                const uint64_t d = solve_expand(s, g->q);
                zo_k_x((double *)s->psi, active_bits(s), d, local_selector(s, g->q));
            } else { /* apply_x (zxh.cpp:731-734) */
                s->b[g->q] ^= 1u;
            }
            break;
        case ZO_CX:
            if (s->disable_x) { /* apply_cx_explicit (zxh.cpp:761-780) */
                flush_diag(s, &w);
                const uint64_t d = solve_expand(s, g->q);
                zo_k_cx((double *)s->psi, active_bits(s), d, local_selector(s, g->cq), local_selector(s, g->q));
            } else { /* apply_cx (zxh.cpp:736-740) */
                s->rows[g->q] ^= s->rows[g->cq];
                s->b[g->q] = (uint8_t)(s->b[g->q] != s->b[g->cq]);
            }
            break;
        case ZO_CP: { /* zxh.cpp:526-541 */
            const uint64_t scq = local_selector(s, g->cq), sq = local_selector(s, g->q);
            if (!dw_try_push_cp(&w, scq, sq, g->theta)) {
                flush_diag(s, &w);
                zo_k_cp((double *)s->psi, active_bits(s), scq, sq, g->theta);
            }
            break;
        }
        default:
            break;
        }
        gi++;
    }
    flush_diag(s, &w);
    dw_free(&w);
    free(pend);
    if (s->fatal)
        s->executed = 0;
    return s->fatal;
}

int zo_get_affine(const zo_sim *s, uint64_t *rows, uint8_t *b, size_t *m_out)
{
    memcpy(rows, s->rows, s->N * sizeof(uint64_t));
    memcpy(b, s->b, s->N);
    *m_out = s->m;
    return ZO_OK;
}

int zo_get_phys_state(const zo_sim *s, double *out, size_t cap_elems, size_t *m_out)
{
    if (!s->executed)
        return ZO_ERR_STATE;
    if (cap_elems < s->active_len)
        return ZO_ERR_ARG;
    memcpy(out, s->psi, s->active_len * sizeof(cplx));
    *m_out = active_bits(s);
    return ZO_OK;
}

/* get_state (zxh.cpp:579-636): virt = A*real ^ b, index = bit-reverse(virt) */
int zo_get_state(const zo_sim *s, double *out)
{
    if (!s->executed)
        return ZO_ERR_STATE;
    if (s->N > 40)
        return ZO_ERR_ARG;
    const size_t N = s->N;
    const size_t vsize = (size_t)1 << N;
    cplx *o = (cplx *)out;
    for (size_t i = 0; i < vsize; i++)
        o[i] = c_mk(0.0, 0.0);
    const size_t real_size = (size_t)1 << s->m;
    for (size_t real = 0; real < real_size; real++) {
        uint64_t virt = 0;
        for (size_t q = 0; q < N; q++) {
            const uint64_t bit = (uint64_t)((__builtin_popcountll(real & s->rows[q]) & 1) ^ s->b[q]);
            virt |= bit << q;
        }
        uint64_t rev = 0;
        for (size_t i = 0; i < N; i++)
            rev = (rev << 1) | ((virt >> i) & 1u);
        o[rev] = c_mul(s->psi[real], s->gphase);
    }
    return ZO_OK;
}

/* ------------------------------------------------------------------------ */
/* sampling                                                                 */

/* mix_seed (measure_general.cpp:16-22): SplitMix64 finaliser */
uint64_t zo_mix_seed(uint64_t base, uint64_t salt)
{
    uint64_t z = base + UINT64_C(0x9e3779b97f4a7c15) + salt;
    z = (z ^ (z >> 30)) * UINT64_C(0xbf58476d1ce4e5b9);
    z = (z ^ (z >> 27)) * UINT64_C(0x94d049bb133111eb);
    return z ^ (z >> 31);
}

/* mt19937_64 (Matsumoto & Nishimura 2004; C++11 [rand.predef] parameters) */
typedef struct {
    uint64_t mt[312];
    int idx;
} mt64;

static void mt64_seed(mt64 *g, uint64_t seed)
{
    g->mt[0] = seed;
    for (int i = 1; i < 312; i++)
        g->mt[i] = UINT64_C(6364136223846793005) * (g->mt[i - 1] ^ (g->mt[i - 1] >> 62)) + (uint64_t)i;
    g->idx = 312;
}

static uint64_t mt64_next(mt64 *g)
{
    static const uint64_t UM = UINT64_C(0xFFFFFFFF80000000), LM = UINT64_C(0x7FFFFFFF);
    static const uint64_t MATRIX_A = UINT64_C(0xB5026F5AA96619E9);
    if (g->idx >= 312) {
        for (int i = 0; i < 312; i++) {
            const uint64_t x = (g->mt[i] & UM) | (g->mt[(i + 1) % 312] & LM);
            uint64_t xa = x >> 1;
            if (x & 1u)
                xa ^= MATRIX_A;
            g->mt[i] = g->mt[(i + 156) % 312] ^ xa;
        }
        g->idx = 0;
    }
    uint64_t y = g->mt[g->idx++];
    y ^= (y >> 29) & UINT64_C(0x5555555555555555);
    y ^= (y << 17) & UINT64_C(0x71D67FFFEDA60000);
    y ^= (y << 37) & UINT64_C(0xFFF7EEE000000000);
    y ^= y >> 43;
    return y;
}

uint64_t zo_mt64_nth(uint64_t seed, size_t n)
{
    mt64 g;
    mt64_seed(&g, seed);
    uint64_t v = 0;
    for (size_t i = 0; i < n; i++)
        v = mt64_next(&g);
    return v;
}

/* uniform_real_distribution<double>(0,1) over mt19937_64 in libstdc++:
 * generate_canonical with one 64-bit draw: u = double(x) / 2^64, and a
 * result that rounds to 1.0 becomes nextafter(1, 0). */
static double canonical01(mt64 *g)
{
    const double x = (double)mt64_next(g);
    double u = x / 18446744073709551616.0;
    if (u >= 1.0)
        u = nextafter(1.0, 0.0);
    return u * (1.0 - 0.0) + 0.0;
}

/* threshold_stream_t (measure_common.cpp:14-71): sorted order statistics */
typedef struct {
    size_t rest;
    double begin, end, head;
    mt64 g;
} tstream;

static void ts_advance(tstream *t)
{
    double u = canonical01(&t->g);
    if (u < DBL_MIN)
        u = DBL_MIN;
    const double span = t->end - t->begin;
    const double factor = 1.0 - pow(u, 1.0 / (double)t->rest);
    t->head = t->begin + span * factor;
}
static void ts_init(tstream *t, size_t cnt, double begin, double end, uint64_t seed)
{
    t->rest = cnt;
    t->begin = begin;
    t->end = end;
    mt64_seed(&t->g, seed);
    if (t->rest)
        ts_advance(t);
}
static void ts_pop(tstream *t)
{
    t->begin = t->head;
    t->rest--;
    if (t->rest)
        ts_advance(t);
}

int zo_thresholds(uint64_t stream_seed, size_t cnt, double begin, double end, double *out)
{
    if (end < begin || (cnt && begin >= end))
        return ZO_ERR_ARG;
    tstream t;
    ts_init(&t, cnt, begin, end, stream_seed);
    for (size_t k = 0; k < cnt; k++) {
        out[k] = t.head;
        ts_pop(&t);
    }
    return ZO_OK;
}

static inline double norm2(const cplx *p, size_t i)
{
    return p[i].re * p[i].re + p[i].im * p[i].im;
}

/* merge_prob_sum_range (measure_general.cpp:39-49) */
static double merge_sum(const cplx *p, size_t begin, size_t end)
{
    if (begin >= end)
        return 0.0;
    if (end - begin == 1)
        return norm2(p, begin);
    const size_t mid = begin + (end - begin) / 2;
    return merge_sum(p, begin, mid) + merge_sum(p, mid, end);
}

double zo_prob_total(const double *psi, size_t m)
{
    return merge_sum((const cplx *)psi, 0, (size_t)1 << m);
}

/* exclusive_scan_merge (measure_general.cpp:51-86): Blelloch scan, padded */
static double blelloch_exclusive(const double *in, size_t count, double *work, size_t padded)
{
    for (size_t i = 0; i < padded; i++)
        work[i] = i < count ? in[i] : 0.0;
    for (size_t stride = 1; stride < padded; stride <<= 1) {
        const size_t step = stride << 1;
        for (size_t i = 0; i < padded; i += step)
            work[i + step - 1] += work[i + stride - 1];
    }
    const double total = work[padded - 1];
    work[padded - 1] = 0.0;
    for (size_t stride = padded >> 1; stride >= 1; stride >>= 1) {
        const size_t step = stride << 1;
        for (size_t i = 0; i < padded; i += step) {
            const double left = work[i + stride - 1];
            work[i + stride - 1] = work[i + step - 1];
            work[i + step - 1] += left;
        }
        if (stride == 1)
            break;
    }
    return total;
}

/* segment_prefix_state_t (measure_streaming_impl.h:46-95): binary counter */
typedef struct {
    double sums[64];
    int active[64];
} seg_prefix;

static double sp_prefix(const seg_prefix *sp)
{
    double acc = 0.0;
    for (int level = 63; level >= 0; level--)
        if (sp->active[level])
            acc += sp->sums[level];
    return acc;
}
static void sp_push(seg_prefix *sp, double seg_sum)
{
    double carry = seg_sum;
    for (int level = 0; level < 64; level++) {
        if (!sp->active[level]) {
            sp->active[level] = 1;
            sp->sums[level] = carry;
            return;
        }
        carry = sp->sums[level] + carry;
        sp->active[level] = 0;
        sp->sums[level] = 0.0;
    }
}

/* std::upper_bound with libstdc++'s probe sequence, then the clamp of
 * resolve_in_segment (measure_streaming_impl.h:15-26) */
static size_t upper_bound_clamped(const double *cdf, size_t n, double tau)
{
    size_t first = 0, len = n;
    while (len > 0) {
        const size_t half = len >> 1;
        const size_t mid = first + half;
        if (tau < cdf[mid])
            len = half;
        else {
            first = mid + 1;
            len = len - half - 1;
        }
    }
    return first == n ? n - 1 : first;
}

/* threshold source: the reference's stream, or a caller-given sorted array */
typedef struct {
    tstream ts;
    const double *arr;
    size_t pos, cnt;
} tsource;
static size_t src_rest(const tsource *s)
{
    return s->arr ? s->cnt - s->pos : s->ts.rest;
}
static double src_head(const tsource *s)
{
    return s->arr ? s->arr[s->pos] : s->ts.head;
}
static void src_pop(tsource *s)
{
    if (s->arr)
        s->pos++;
    else
        ts_pop(&s->ts);
}

/* sample_local_streaming (measure_streaming_impl.h:97-147) on 2^m amplitudes */
static int sample_streaming_src(const cplx *p, size_t m, size_t I, tsource *src, size_t cnt, uint64_t *results);

static int sample_streaming(const cplx *p, size_t m, size_t I, double local_prob, uint64_t stream_seed, size_t cnt,
                            uint64_t *results)
{
    if (cnt == 0)
        return ZO_OK;
    if (local_prob <= 0.0)
        return ZO_ERR_ZERO_PROB;
    tsource src;
    memset(&src, 0, sizeof(src));
    ts_init(&src.ts, cnt, 0.0, local_prob, stream_seed);
    return sample_streaming_src(p, m, I, &src, cnt, results);
}

int zo_sample_thresholds(const double *psi, size_t m, size_t seg_bits, const double *thr, size_t cnt, uint64_t *out)
{
    const size_t I = m < seg_bits ? m : seg_bits;
    tsource src;
    memset(&src, 0, sizeof(src));
    src.arr = thr;
    src.cnt = cnt;
    return sample_streaming_src((const cplx *)psi, m, I, &src, cnt, out);
}

static int sample_streaming_src(const cplx *p, size_t m, size_t I, tsource *src, size_t cnt, uint64_t *results)
{
    if (cnt == 0)
        return ZO_OK;
    const size_t n = (size_t)1 << m;
    const size_t seg_len = (size_t)1 << I;
    double *cdf = (double *)malloc(seg_len * sizeof(double));
    double *work = (double *)malloc(seg_len * sizeof(double));
    double *pin = (double *)malloc(seg_len * sizeof(double));
    if (!cdf || !work || !pin) {
        free(cdf);
        free(work);
        free(pin);
        return ZO_ERR_ALLOC;
    }
    seg_prefix sp;
    memset(&sp, 0, sizeof(sp));
    size_t pcnt = 0;
    size_t seg = 0;
    double cur_prefix = sp_prefix(&sp);
    while (seg < n && src_rest(src) != 0) {
        /* segment_cdf_t::pre_cdf (measure_general.cpp:182-203) */
        const size_t valid = seg_len < n - seg ? seg_len : n - seg;
        size_t padded = 1;
        while (padded < valid)
            padded <<= 1;
        for (size_t i = 0; i < valid; i++)
            pin[i] = norm2(p, seg + i);
        const double total = blelloch_exclusive(pin, valid, work, padded);
        for (size_t i = 0; i < valid; i++)
            cdf[i] = cur_prefix + work[i] + pin[i];
        const double tail = cur_prefix + total;
        sp_push(&sp, total);
        /* consume_stream_in_segment */
        while (src_rest(src) != 0 && src_head(src) < tail) {
            if (pcnt >= cnt) {
                free(cdf);
                free(work);
                free(pin);
                return ZO_ERR_UNRESOLVED;
            }
            results[pcnt++] = seg + upper_bound_clamped(cdf, valid, src_head(src));
            src_pop(src);
        }
        seg += seg_len;
        cur_prefix = sp_prefix(&sp);
    }
    free(cdf);
    free(work);
    free(pin);
    if (src_rest(src) != 0 || pcnt != cnt)
        return ZO_ERR_UNRESOLVED;
    return ZO_OK;
}

int zo_sample_state(const double *psi, size_t m, size_t seg_bits, uint64_t stream_seed, size_t cnt, uint64_t *out)
{
    const cplx *p = (const cplx *)psi;
    const size_t I = m < seg_bits ? m : seg_bits;
    const double total = merge_sum(p, 0, (size_t)1 << m);
    return sample_streaming(p, m, I, total, stream_seed, cnt, out);
}

/* measure (measure_common.cpp:73-92) */
int zo_measure(zo_sim *s, size_t cnt)
{
    if (!s->executed)
        return ZO_ERR_STATE;
    if (cnt == 0) {
        s->result_count = 0;
        return ZO_OK;
    }
    const size_t m = active_bits(s);
    const double local_prob = merge_sum(s->psi, 0, s->active_len);
    if (local_prob <= 0.0)
        return ZO_ERR_ZERO_PROB;
    uint64_t base = s->seed;
    if (!s->use_seed) {
        base = (uint64_t)rand() ^ ((uint64_t)rand() << 31);
    }
    const uint64_t stream_seed = zo_mix_seed(base, 0);
    uint64_t *rr = (uint64_t *)realloc(s->real_results, cnt * sizeof(uint64_t));
    const size_t wc = (s->N + 63) / 64;
    uint64_t *pk = (uint64_t *)realloc(s->packed, cnt * (wc ? wc : 1) * sizeof(uint64_t));
    if (!rr || !pk)
        return ZO_ERR_ALLOC;
    s->real_results = rr;
    s->packed = pk;
    int rc = sample_streaming(s->psi, m, s->I < m ? s->I : m, local_prob, stream_seed, cnt, rr);
    if (rc)
        return rc;
    /* pack_virtual_results_host (measure_result_pack.h:24-44) */
    memset(pk, 0, cnt * wc * sizeof(uint64_t));
    for (size_t i = 0; i < cnt; i++) {
        for (size_t q = 0; q < s->N; q++) {
            const int bit = (__builtin_popcountll(rr[i] & s->rows[q]) & 1) ^ s->b[q];
            if (bit)
                pk[i * wc + (q >> 6)] |= UINT64_C(1) << (q & 63);
        }
    }
    s->result_count = cnt;
    return ZO_OK;
}

size_t zo_measured_count(const zo_sim *s)
{
    return s->result_count;
}

int zo_get_results(const zo_sim *s, uint64_t *packed, size_t cnt)
{
    if (cnt > s->result_count)
        return ZO_ERR_ARG;
    const size_t wc = (s->N + 63) / 64;
    memcpy(packed, s->packed, cnt * wc * sizeof(uint64_t));
    return ZO_OK;
}

int zo_get_real_results(const zo_sim *s, uint64_t *real, size_t cnt)
{
    if (cnt > s->result_count)
        return ZO_ERR_ARG;
    memcpy(real, s->real_results, cnt * sizeof(uint64_t));
    return ZO_OK;
}
