/*
 * CPU oracle for the XAMG solve path -- TEST INFRASTRUCTURE ONLY (see oracle.h).
 * Build: make -C oracle   (gcc -O2 -fopenmp -ffp-contract=off)
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ---- rand.h:33-56: xorshift32 then an LCG step in 32-bit two's complement
 * (the reference relies on signed int wrap-around; restated in u32). ---- */
static uint32_t xorshift32(uint32_t i) {
    uint32_t s = i + 1u;
    s ^= (s << 13);
    s ^= (s >> 17);
    s ^= (s << 5);
    return s;
}

int32_t orc_pseudo_rand_i(uint32_t i) {
    uint32_t g = xorshift32(i);
    g = 1664525u * g + 1013904223u;
    return (int32_t)g;
}

/* rand.h:64-71: pseudo_rand<double>(uint64_t i) truncates i to 32 bits */
double orc_pseudo_rand_d(uint64_t i) {
    double val = (double)orc_pseudo_rand_i((uint32_t)i);
    return 0.5 * (1.0 + val / 2147483647.0);
}

void orc_fill_rhs(double *b, int64_t n, int nv, uint32_t seed) {
    for (int64_t i = 0; i < n; i++)
        for (int l = 0; l < nv; l++) {
            uint32_t k = (uint32_t)((uint64_t)seed * (1ull << 27) + (uint64_t)i * nv + l);
            b[i * nv + l] = 2.0 * (orc_pseudo_rand_d(k) - 0.5);
        }
}

/* misc.h:47-56: size/nblocks rounded up to the 64-byte element multiple; a size
 * already on the multiple still gains a full extra multiple. */
uint64_t orc_split_range(uint64_t size, int nblocks, int elem_size) {
    uint32_t factor = 64u / (uint32_t)elem_size;
    uint64_t core_size = size / (uint64_t)nblocks;
    core_size += factor - (core_size % factor);
    return core_size;
}

/* part.h:115-126 / vector.inl:207-211 */
void orc_core_block(uint64_t size, int nblocks, int k, int elem_size, uint64_t *off,
                    uint64_t *len) {
    uint64_t cs = orc_split_range(size, nblocks, elem_size);
    uint64_t o = cs * (uint64_t)k;
    if (o > size)
        o = size;
    uint64_t l = cs < size - o ? cs : size - o;
    *off = o;
    *len = l;
}

/* blas.inl:431-484 bxor over the 64-bit patterns, per RHS */
void orc_bxor_d(const double *x, int64_t n, int nv, uint64_t *res) {
    for (int l = 0; l < nv; l++)
        res[l] = 0;
    const uint64_t *p = (const uint64_t *)x;
    for (int64_t i = 0; i < n; i++)
        for (int l = 0; l < nv; l++)
            res[l] ^= p[i * nv + l];
}

/* chebyshev.inl:67-125: bounds [(1.1*max - min)*f + min, 1.1*max], closed-form
 * coefficients of the order-k polynomial, in double. */
void orc_cheb_coeffs(double min_eig, double max_eig, int order, double fraction, double *c) {
    double den;
    double upper = max_eig * 1.1;
    double lower = (upper - min_eig) * fraction + min_eig;
    double theta = (upper + lower) / 2.;
    double delta = (upper - lower) / 2.;
    switch (order) {
    case 1:
        c[0] = 1. / theta;
        break;
    case 2:
        den = delta * delta - 2 * theta * theta;
        c[0] = -4 * theta / den;
        c[1] = 2 / den;
        break;
    case 3:
        den = 3 * (delta * delta) * theta - 4 * (theta * theta * theta);
        c[0] = (3 * delta * delta - 12 * theta * theta) / den;
        c[1] = 12 * theta / den;
        c[2] = -4 / den;
        break;
    case 4:
        den = pow(delta, 4) - 8 * delta * delta * theta * theta + 8 * pow(theta, 4);
        c[0] = (32 * pow(theta, 3) - 16 * delta * delta * theta) / den;
        c[1] = (8 * delta * delta - 48 * theta * theta) / den;
        c[2] = 32 * theta / den;
        c[3] = -8 / den;
        break;
    default:
        break;
    }
}

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ---- float instantiation ---- */
#define F float
#define SFX(x) x##_f
#define ORC_IS_DOUBLE 0
#define orc_conv_t_f conv_t_f
#include "oracle_tmpl.h"
#undef F
#undef SFX
#undef ORC_IS_DOUBLE

typedef struct rt_f orc_rt_f;

/* ---- double instantiation ---- */
#define F double
#define SFX(x) x##_d
#define ORC_IS_DOUBLE 1
#define orc_conv_t_d conv_t_d
#include "oracle_tmpl.h"
#undef F
#undef SFX
#undef ORC_IS_DOUBLE

struct orc_ctx {
    int prec;
    rt_d *d;
    rt_f *f;
};

void *orc_setup(const orc_node *node, const orc_csr *A, int nv, int single) {
    struct orc_ctx *c = (struct orc_ctx *)calloc(1, sizeof(*c));
    c->prec = single;
    if (single)
        c->f = rt_build_f(node, A, nv, 0);
    else
        c->d = rt_build_d(node, A, nv, 0);
    return c;
}

static void fill_result_d(const conv_t_d *cv, int nv, orc_result *res) {
    res->iters = cv->iter;
    res->sub_iters = cv->sub_iter;
    res->rr_checks = cv->rr_checks;
    res->rr_stopped = cv->rr_stopped;
    for (int l = 0; l < nv; l++) {
        res->converged[l] = cv->status[l];
        res->res0[l] = cv->res0[l];
        res->res[l] = cv->res[l];
    }
}

static void fill_result_f(const conv_t_f *cv, int nv, orc_result *res) {
    res->iters = cv->iter;
    res->sub_iters = cv->sub_iter;
    res->rr_checks = cv->rr_checks;
    res->rr_stopped = cv->rr_stopped;
    for (int l = 0; l < nv; l++) {
        res->converged[l] = cv->status[l];
        res->res0[l] = cv->res0[l];
        res->res[l] = cv->res[l];
    }
}

int orc_run_d(void *ctx, double *x, int xzero, const double *b, orc_result *res) {
    struct orc_ctx *c = (struct orc_ctx *)ctx;
    if (res)
        res->nhist = 0;
    c->d->conv.rec = res;
    solve_d(c->d, x, xzero, b, NULL);
    if (res)
        fill_result_d(&c->d->conv, c->d->nv, res);
    return 0;
}

int orc_run_f(void *ctx, float *x, int xzero, const float *b, orc_result *res) {
    struct orc_ctx *c = (struct orc_ctx *)ctx;
    if (res)
        res->nhist = 0;
    c->f->conv.rec = res;
    solve_f(c->f, x, xzero, b, NULL);
    if (res)
        fill_result_f(&c->f->conv, c->f->nv, res);
    return 0;
}

int orc_solve_d(const orc_node *node, const orc_csr *A, double *x, const double *b, int nv,
                orc_result *res) {
    void *c = orc_setup(node, A, nv, 0);
    return orc_run_d(c, x, 0, b, res);
}

int orc_solve_f(const orc_node *node, const orc_csr *A, float *x, const float *b, int nv,
                orc_result *res) {
    void *c = orc_setup(node, A, nv, 1);
    return orc_run_f(c, x, 0, b, res);
}
