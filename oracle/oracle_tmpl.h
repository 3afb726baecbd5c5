/*
 * Precision-generic body of the oracle (TEST INFRASTRUCTURE ONLY).  Included twice
 * from oracle.c: with F=float/SFX=_f and F=double/SFX=_d.  Each routine restates the
 * reference routine it names, in the same operation order.
 */

/* ------------------------------------------------------------------------- */
/* BLAS2 (csr_matvec.inl)                                                     */
/* ------------------------------------------------------------------------- */

/* csr_Ax_y, csr_matvec.inl:148-260 (loop :213-231): y[i,:]=0; y[i,:] += x[col,:]*val */
void SFX(orc_spmv)(const orc_csr *A, const F *x, F *y, int nv) {
    const F *val = (const F *)A->val;
    const uint32_t *row = A->row, *col = A->col;
    int64_t n = A->nrows;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        F acc[ORC_MAX_NV];
        for (int l = 0; l < nv; l++)
            acc[l] = 0.0;
        for (uint32_t j = row[i]; j < row[i + 1]; j++) {
            const F *xp = x + (int64_t)col[j] * nv;
            for (int l = 0; l < nv; l++)
                acc[l] += xp[l] * val[j];
        }
        for (int l = 0; l < nv; l++)
            y[i * nv + l] = acc[l];
    }
}

/* csr_Axpy, csr_matvec.inl:32-144 (loop :106-120): accumulates directly into y */
void SFX(orc_spmv_axpy)(const orc_csr *A, const F *x, F *y, int nv) {
    const F *val = (const F *)A->val;
    const uint32_t *row = A->row, *col = A->col;
    int64_t n = A->nrows;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        for (uint32_t j = row[i]; j < row[i + 1]; j++) {
            const F *xp = x + (int64_t)col[j] * nv;
            for (int l = 0; l < nv; l++)
                y[i * nv + l] += xp[l] * val[j];
        }
    }
}

/* dense_matvec.inl:100-117 Ax_y on a row-major dense block */
static void SFX(dense_mv)(const F *a, int64_t n, const F *x, F *y, int nv) {
    for (int64_t i = 0; i < n; i++) {
        for (int l = 0; l < nv; l++)
            y[i * nv + l] = 0.0;
        for (int64_t j = 0; j < n; j++)
            for (int l = 0; l < nv; l++)
                y[i * nv + l] += a[i * n + j] * x[j * nv + l];
    }
}

/* ------------------------------------------------------------------------- */
/* BLAS1 (blas.inl / blas_ext.inl)                                           */
/* ------------------------------------------------------------------------- */

/* blas.inl:521-589 dot; chunks follow split_range (misc.h:47-56) and are summed in
 * chunk order like the shared-memory allreduce (shm_allreduce.h:123-134). */
void SFX(orc_dot)(const F *x, const F *y, int64_t n, int nv, int nchunks, F *res) {
    if (nchunks < 1)
        nchunks = 1;
    F part[256][ORC_MAX_NV];
    if (nchunks > 256)
        nchunks = 256;
#pragma omp parallel for schedule(static)
    for (int c = 0; c < nchunks; c++) {
        uint64_t off, len;
        orc_core_block((uint64_t)n, nchunks, c, (int)sizeof(double), &off, &len);
        F acc[ORC_MAX_NV];
        for (int l = 0; l < nv; l++)
            acc[l] = 0.0;
        for (uint64_t i = off; i < off + len; i++)
            for (int l = 0; l < nv; l++)
                acc[l] += x[i * nv + l] * y[i * nv + l];
        for (int l = 0; l < nv; l++)
            part[c][l] = acc[l];
    }
    for (int l = 0; l < nv; l++) {
        F s = part[0][l];
        for (int c = 1; c < nchunks; c++)
            s += part[c][l];
        res[l] = s;
    }
}

/* csr_matrix.h:323-355 inv_diag = 1.0/a_ii (computed in double, stored as F) */
void SFX(orc_inv_diag)(const orc_csr *A, F *d) {
    const F *val = (const F *)A->val;
    for (int64_t i = 0; i < A->nrows; i++)
        for (uint32_t j = A->row[i]; j < A->row[i + 1]; j++)
            if ((uint32_t)i == A->col[j])
                d[i] = 1.0 / val[j];
}

/* csr_matrix.h:357-389 inv_sqrt_diag = sqrt(1.0/a_ii) */
void SFX(orc_inv_sqrt_diag)(const orc_csr *A, F *d) {
    const F *val = (const F *)A->val;
    for (int64_t i = 0; i < A->nrows; i++)
        for (uint32_t j = A->row[i]; j < A->row[i + 1]; j++)
            if ((uint32_t)i == A->col[j])
                d[i] = sqrt(1.0 / val[j]);
}

/* inverse.inl:31-104: Gauss-Jordan on [A | I] without pivoting, in precision F */
void SFX(orc_dense_inverse)(const F *a, int64_t n, F *ainv) {
    int64_t nc = 2 * n;
    F *e = (F *)calloc((size_t)(n * nc), sizeof(F));
    for (int64_t i = 0; i < n; i++) {
        e[i * nc + i + n] = 1.0;
        for (int64_t j = 0; j < n; j++)
            e[i * nc + j] = a[i * n + j];
    }
    for (int64_t i = 0; i < n; i++) {
        F *ri = e + i * nc;
        for (int64_t k = nc - 1; k > i; --k)
            ri[k] /= ri[i];
        ri[i] = 1.0;
        for (int64_t j = i + 1; j < n; j++) {
            F *rj = e + j * nc;
            for (int64_t k = nc - 1; k > i; --k)
                rj[k] -= rj[i] * ri[k];
            rj[i] = 0.0;
        }
    }
    for (int64_t i = 0; i < n; i++) {
        int64_t r = n - 1 - i;
        F *ri = e + r * nc;
        for (int64_t j = 0; j < r; j++) {
            F *rj = e + j * nc;
            for (int64_t k = nc / 2; k < nc; k++)
                rj[k] -= rj[r] * ri[k];
            rj[r] = 0.0;
        }
    }
    for (int64_t i = 0; i < n; i++)
        for (int64_t j = 0; j < n; j++)
            ainv[i * n + j] = e[i * nc + n + j];
    free(e);
}

/* ------------------------------------------------------------------------- */
/* Convergence (convergence.h:49-303, stats.h:50-320)                         */
/* ------------------------------------------------------------------------- */

typedef struct {
    int conv_check;
    double abs_tol, rel_tol;
    int status[ORC_MAX_NV];
    double res0[ORC_MAX_NV], res[ORC_MAX_NV];
    int iter, sub_iter;
    int nv;
    orc_result *rec;
    /* residual replacement (convergence_details RR, residual_control block) */
    int rr;
    double rr_stepping, rr_deviation; /* float32_t in convergence.h:61-62 */
    double rr_tol;
    int rr_checks, rr_stopped;
} SFX(conv_t);

/* convergence.h:173-195 init(conv_vector) / init() */
static void SFX(conv_init)(SFX(conv_t) * c, const F *iconv) {
    c->iter = c->sub_iter = 0;
    for (int l = 0; l < c->nv; l++) {
        c->res0[l] = c->res[l] = 0.0;
        c->status[l] = iconv ? (fabs(1.0 - (double)iconv[l]) > 1e-10) : 0;
    }
    c->rr_tol = c->rr_stepping; /* convergence.h:184-186, 193-195 */
    c->rr_checks = c->rr_stopped = 0;
}

static void SFX(conv_record)(SFX(conv_t) * c) {
    orc_result *r = c->rec;
    if (!r || !r->hist || r->nhist >= r->hist_cap)
        return;
    for (int l = 0; l < c->nv; l++)
        r->hist[(int64_t)r->nhist * c->nv + l] = c->status[l] ? -1.0 : c->res[l];
    r->nhist++;
}

/* convergence.h:223-239 check(): per RHS compare in precision F */
static int SFX(conv_check_all)(SFX(conv_t) * c) {
    if (!c->conv_check)
        return 0;
    int flag = 1;
    for (int l = 0; l < c->nv; l++) {
        if (c->status[l])
            continue;
        double tol = fmax(c->abs_tol, c->rel_tol * sqrt(c->res0[l]));
        F abs_res = sqrt((F)c->res[l]);
        if (abs_res > (F)tol)
            flag = 0;
        else
            c->status[l] = 1;
    }
    return flag;
}

/* stats.h:165-181 internal_is_converged + convergence.h:241-262 is_converged(res, conv) */
static int SFX(conv_is_converged)(SFX(conv_t) * c, const F *rnorm, F *iconv) {
    if (!c->conv_check)
        return 0;
    for (int l = 0; l < c->nv; l++)
        if (!c->status[l])
            c->res[l] = (double)rnorm[l];
    SFX(conv_record)(c);
    int flag = SFX(conv_check_all)(c);
    if (!flag && iconv)
        for (int l = 0; l < c->nv; l++)
            if (c->status[l])
                iconv[l] = 0.0;
    return flag;
}

static void SFX(conv_set_initial)(SFX(conv_t) * c, const F *rnorm) {
    for (int l = 0; l < c->nv; l++)
        c->res0[l] = (double)rnorm[l];
}

/* ------------------------------------------------------------------------- */
/* Runtime tree                                                               */
/* ------------------------------------------------------------------------- */

typedef struct SFX(rt) SFX(rt);
struct SFX(rt) {
    const orc_node *node;
    orc_csr A; /* level matrix, values of type F */
    int64_t n;
    int nv;
    F *inv_diag, *inv_sqrt_diag;
    F *ainv; /* Direct */
    F cheb[4];
    F *buf[8];
    int nbuf;
    F *rrbuf; /* residual replacement: true residual (stats.h:58-59 rr_temp_vector) */
    SFX(conv_t) conv;
    SFX(rt) * precond;
    /* MultiGrid */
    int nlev;
    SFX(rt) * pre[ORC_MAX_LEVELS], *post[ORC_MAX_LEVELS], *coarse;
    F *temp[ORC_MAX_LEVELS], *bH[ORC_MAX_LEVELS], *xH[ORC_MAX_LEVELS];
    orc_csr Al[ORC_MAX_LEVELS], Pl[ORC_MAX_LEVELS], Rl[ORC_MAX_LEVELS];
#if ORC_IS_DOUBLE
    /* mixed MG (mg_reduced_precision): levels >= 1 in float */
    orc_rt_f *pre32[ORC_MAX_LEVELS], *post32[ORC_MAX_LEVELS], *coarse32;
    float *temp32_0;
    float *temp32[ORC_MAX_LEVELS], *bH32[ORC_MAX_LEVELS], *xH32[ORC_MAX_LEVELS];
    /* IterativeRefinement with reduced inner solver */
    orc_rt_f *inner32;
    float *ix32, *ib32;
    float *A32val;
    orc_csr A32;
#endif
};

static F *SFX(vnew)(int64_t n, int nv) {
    return (F *)calloc((size_t)(n > 0 ? n : 1) * nv, sizeof(F));
}

SFX(rt) * SFX(rt_build)(const orc_node *node, const orc_csr *A, int nv, int level);

/* Ax_y with the reference's zero-vector shortcut (blas2.h:305-326). */
static void SFX(ax_y)(const orc_csr *A, const F *x, int xzero, F *y, int nv) {
    if (xzero) {
        memset(y, 0, sizeof(F) * A->nrows * nv);
        return;
    }
    SFX(orc_spmv)(A, x, y, nv);
}

/* solver.h:183-190 get_residual: r = Ax; r = 1*b + (-1)*r */
static void SFX(residual)(const orc_csr *A, const F *x, int xzero, const F *b, F *r, int nv) {
    int64_t n = A->nrows;
    if (xzero) {
        for (int64_t i = 0; i < n * nv; i++)
            r[i] = (F)1.0 * b[i];
        return;
    }
    SFX(orc_spmv)(A, x, r, nv);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n * nv; i++)
        r[i] = (F)1.0 * b[i] + (F)(-1.0) * r[i];
}

static void SFX(norm2)(SFX(rt) * rt, const F *r, F *res) {
    SFX(orc_dot)(r, r, rt->n, rt->nv, rt->node->nchunks, res);
}

static void SFX(solve)(SFX(rt) * rt, F *x, int xzero, const F *b, F *iconv);

/* stats.h:151-160 rr_is_converged with convergence.h:269-302 rr_check and
 * rr_deviation_check.  Once the largest relative recurrence residual sqrt(res/res0)
 * over every RHS (a converged one at its last value) has fallen to rr_tol, rr_tol
 * becomes that value times RR_stepping, the true residual b - A x is computed
 * (solver.h:183-190) and its squared norm taken; a recurrence/true ratio of the norms
 * outside [1/RR_deviation, RR_deviation] for any RHS ends the solve: the check reports
 * convergence and the RHS statuses stay as they are. */
static int SFX(rr_is_converged)(SFX(rt) * rt, const F *x, int xzero, const F *b) {
    SFX(conv_t) *c = &rt->conv;
    if (!c->rr || !c->conv_check)
        return 0;
    double maxrel = 0.0;
    for (int l = 0; l < c->nv; l++) {
        const double ratio = sqrt(c->res[l] / c->res0[l]);
        maxrel = maxrel < ratio ? ratio : maxrel; /* std::max(maxrel, ratio) */
    }
    if (maxrel > c->rr_tol)
        return 0;
    c->rr_tol = maxrel * c->rr_stepping;
    c->rr_checks++;
    F tn[ORC_MAX_NV];
    SFX(residual)(&rt->A, x, xzero, b, rt->rrbuf, rt->nv);
    SFX(norm2)(rt, rt->rrbuf, tn);
    for (int l = 0; l < c->nv; l++) {
        const double ratio = sqrt(c->res[l] / (double)tn[l]);
        if (ratio > c->rr_deviation || ratio < 1.0 / c->rr_deviation) {
            c->rr_stopped = 1;
            return 1;
        }
    }
    return 0;
}

/* stats.h:165-181 internal_is_converged: the check, then residual replacement */
static int SFX(conv_is_converged_rr)(SFX(rt) * rt, const F *rnorm, F *iconv, const F *x, int xzero,
                                     const F *b) {
    int flag = SFX(conv_is_converged)(&rt->conv, rnorm, iconv);
    if (!flag)
        flag = SFX(rr_is_converged)(rt, x, xzero, b);
    return flag;
}

/* ---- BiCGStab bicgstab.inl:43-197 / PBiCGStab pbicgstab.inl:47-207 ---- */
static void SFX(solve_bicgstab)(SFX(rt) * rt, F *x, int xzero, const F *b, F *iconv_in,
                                int precond) {
    int64_t n = rt->n, N;
    int nv = rt->nv;
    N = n * nv;
    const orc_csr *A = &rt->A;
    F iconv[ORC_MAX_NV];
    for (int l = 0; l < nv; l++)
        iconv[l] = iconv_in ? iconv_in[l] : (F)1.0;
    F rho[ORC_MAX_NV], res[ORC_MAX_NV], s1[ORC_MAX_NV], s2[ORC_MAX_NV], s3[ORC_MAX_NV],
        s4[ORC_MAX_NV], alpha[ORC_MAX_NV], beta[ORC_MAX_NV], omega[ORC_MAX_NV], ac[ORC_MAX_NV],
        oc[ORC_MAX_NV];
    F *r0 = rt->buf[0], *r = rt->buf[1], *p = rt->buf[2], *v = rt->buf[3], *s = rt->buf[4],
      *t = rt->buf[5];
    F *p_ = precond ? rt->buf[6] : p;
    F *s_ = precond ? rt->buf[7] : s;
    SFX(conv_t) *c = &rt->conv;
    SFX(conv_init)(c, iconv);
    /* stats.h:195-206 is_converged_initial(r0, rho) */
    SFX(residual)(A, x, xzero, b, r0, nv);
    SFX(norm2)(rt, r0, rho);
    if (c->conv_check) {
        SFX(conv_set_initial)(c, rho);
        if (SFX(conv_is_converged_rr)(rt, rho, iconv, x, xzero, b))
            return;
    }
    memcpy(r, r0, sizeof(F) * N);
    memcpy(p, r0, sizeof(F) * N);
    /* a zero-flagged x (set_const(x, 0)) makes axpbypcz skip the old x (blas_ext.inl:117-120) */
    if (xzero)
        memset(x, 0, sizeof(F) * N);
    int max_iters = rt->node->max_iters;
    for (int last = 0; !last;) {
        int it = ++c->iter;
        if (it >= max_iters)
            last = 1;
        if (precond)
            SFX(solve)(rt->precond, p_, 1, p, NULL); /* set_const(p_,0); precond->solve */
        SFX(orc_spmv)(A, p_, v, nv);
        SFX(orc_dot)(v, r0, n, nv, rt->node->nchunks, s1);
        for (int l = 0; l < nv; l++)
            alpha[l] = rho[l] / s1[l];
        for (int l = 0; l < nv; l++)
            alpha[l] *= (F)-1.0;
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; i++)
            for (int l = 0; l < nv; l++)
                s[i * nv + l] = (F)1.0 * r[i * nv + l] + alpha[l] * v[i * nv + l];
        for (int l = 0; l < nv; l++)
            alpha[l] *= (F)-1.0;
        if (precond)
            SFX(solve)(rt->precond, s_, 1, s, NULL);
        SFX(orc_spmv)(A, s_, t, nv);
        SFX(orc_dot)(t, s, n, nv, rt->node->nchunks, s2);
        SFX(orc_dot)(t, t, n, nv, rt->node->nchunks, s3);
        SFX(orc_dot)(s, s, n, nv, rt->node->nchunks, s4);
        for (int l = 0; l < nv; l++)
            omega[l] = s2[l] / s3[l];
        for (int l = 0; l < nv; l++) {
            ac[l] = iconv[l] * alpha[l];
            oc[l] = iconv[l] * omega[l];
        }
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; i++)
            for (int l = 0; l < nv; l++)
                x[i * nv + l] = ac[l] * p_[i * nv + l] + oc[l] * s_[i * nv + l] +
                                (F)1.0 * x[i * nv + l];
        for (int l = 0; l < nv; l++) {
            res[l] = omega[l] * s2[l];
            res[l] = (F)1.0 * s4[l] + (F)(-1.0) * res[l];
        }
        if (SFX(conv_is_converged_rr)(rt, res, iconv, x, 0, b))
            return;
        for (int l = 0; l < nv; l++)
            omega[l] *= (F)-1.0;
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; i++)
            for (int l = 0; l < nv; l++)
                r[i * nv + l] = (F)1.0 * s[i * nv + l] + omega[l] * t[i * nv + l];
        for (int l = 0; l < nv; l++)
            omega[l] *= (F)-1.0;
        SFX(orc_dot)(r, r0, n, nv, rt->node->nchunks, s1);
        for (int l = 0; l < nv; l++) {
            s2[l] = s1[l] / rho[l];
            s3[l] = alpha[l] / omega[l];
            beta[l] = s2[l] * s3[l];
            rho[l] = s1[l];
            s1[l] = omega[l] * beta[l];
            s1[l] *= (F)-1.0;
        }
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; i++)
            for (int l = 0; l < nv; l++)
                p[i * nv + l] = (F)1.0 * r[i * nv + l] + s1[l] * v[i * nv + l] +
                                beta[l] * p[i * nv + l];
    }
}

/* ---- PCG pcg.inl:47-157 ---- */
static void SFX(solve_pcg)(SFX(rt) * rt, F *x, int xzero, const F *b, F *iconv_in) {
    int64_t n = rt->n;
    int nv = rt->nv;
    const orc_csr *A = &rt->A;
    F iconv[ORC_MAX_NV], rn[ORC_MAX_NV], s1[ORC_MAX_NV], s2[ORC_MAX_NV], s3[ORC_MAX_NV],
        alpha[ORC_MAX_NV], beta[ORC_MAX_NV], ac[ORC_MAX_NV];
    for (int l = 0; l < nv; l++)
        iconv[l] = iconv_in ? iconv_in[l] : (F)1.0;
    F *r = rt->buf[0], *z = rt->buf[1], *p = rt->buf[2], *Ap = rt->buf[3];
    SFX(conv_t) *c = &rt->conv;
    SFX(conv_init)(c, iconv);
    /* stats.init(r, iconv); is_converged_initial(r) (stats.h:208-219) */
    SFX(residual)(A, x, xzero, b, r, nv);
    if (c->conv_check) {
        SFX(norm2)(rt, r, rn);
        SFX(conv_set_initial)(c, rn);
        if (SFX(conv_is_converged_rr)(rt, rn, iconv, x, xzero, b))
            return;
    }
    /* z = M^{-1} r from a zero-flagged z; p = z */
    SFX(solve)(rt->precond, z, 1, r, NULL);
    memcpy(p, z, sizeof(F) * n * nv);
    if (xzero)
        memset(x, 0, sizeof(F) * n * nv);
    int max_iters = rt->node->max_iters;
    for (int last = 0; !last;) {
        int it = ++c->iter;
        if (it >= max_iters)
            last = 1;
        SFX(orc_spmv)(A, p, Ap, nv);
        SFX(orc_dot)(r, z, n, nv, rt->node->nchunks, s1);
        SFX(orc_dot)(p, Ap, n, nv, rt->node->nchunks, s2);
        for (int l = 0; l < nv; l++) {
            alpha[l] = s1[l] / s2[l];
            ac[l] = iconv[l] * alpha[l];
        }
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; i++)
            for (int l = 0; l < nv; l++)
                x[i * nv + l] = ac[l] * p[i * nv + l] + (F)1.0 * x[i * nv + l];
        for (int l = 0; l < nv; l++)
            alpha[l] = (F)(-1.0) * alpha[l];
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; i++)
            for (int l = 0; l < nv; l++)
                r[i * nv + l] = alpha[l] * Ap[i * nv + l] + (F)1.0 * r[i * nv + l];
        SFX(solve)(rt->precond, z, 1, r, NULL);
        SFX(orc_dot)(r, z, n, nv, rt->node->nchunks, s3);
        for (int l = 0; l < nv; l++)
            beta[l] = s3[l] / s1[l];
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; i++)
            for (int l = 0; l < nv; l++)
                p[i * nv + l] = (F)1.0 * z[i * nv + l] + beta[l] * p[i * nv + l];
        SFX(orc_dot)(r, r, n, nv, rt->node->nchunks, rn);
        if (SFX(conv_is_converged_rr)(rt, rn, iconv, x, 0, b))
            return;
    }
}

/* ---- RBiCGStab rbicgstab.inl:47-205 ---- */
static void SFX(solve_rbicgstab)(SFX(rt) * rt, F *x, int xzero, const F *b, F *iconv_in) {
    int64_t n = rt->n;
    int nv = rt->nv;
    const orc_csr *A = &rt->A;
    F iconv[ORC_MAX_NV], rho[ORC_MAX_NV], res[ORC_MAX_NV], theta[ORC_MAX_NV], phi[ORC_MAX_NV],
        psi[ORC_MAX_NV], delta[ORC_MAX_NV], alpha[ORC_MAX_NV], beta[ORC_MAX_NV], omega[ORC_MAX_NV],
        eta[ORC_MAX_NV], temp[ORC_MAX_NV], ac[ORC_MAX_NV], oc[ORC_MAX_NV];
    for (int l = 0; l < nv; l++)
        iconv[l] = iconv_in ? iconv_in[l] : (F)1.0;
    F *r0 = rt->buf[0], *r = rt->buf[1], *z = rt->buf[2], *v_ = rt->buf[3], *v = rt->buf[4],
      *s = rt->buf[5], *t_ = rt->buf[6], *t = rt->buf[7];
    SFX(conv_t) *c = &rt->conv;
    SFX(conv_init)(c, iconv);
    /* 1-2: r0 = b - A x0, rho = (r0, r0) (stats.h:195-206) */
    SFX(residual)(A, x, xzero, b, r0, nv);
    SFX(norm2)(rt, r0, rho);
    if (c->conv_check) {
        SFX(conv_set_initial)(c, rho);
        if (SFX(conv_is_converged_rr)(rt, rho, iconv, x, xzero, b))
            return;
    }
    memcpy(r, r0, sizeof(F) * n * nv);
    SFX(solve)(rt->precond, z, 1, r0, NULL); /* 3 */
    memcpy(v_, z, sizeof(F) * n * nv);      /* 4 */
    if (xzero)
        memset(x, 0, sizeof(F) * n * nv);
    int max_iters = rt->node->max_iters;
    for (int last = 0; !last;) {
        int it = ++c->iter;
        if (it >= max_iters)
            last = 1;
        SFX(orc_spmv)(A, v_, v, nv);                          /* 6 */
        SFX(orc_dot)(v, r0, n, nv, rt->node->nchunks, delta); /* 7 */
        SFX(solve)(rt->precond, s, 1, v, NULL);               /* 8 */
        for (int l = 0; l < nv; l++)                          /* 9 */
            alpha[l] = rho[l] / delta[l];
        for (int l = 0; l < nv; l++) /* 10: t_ = z - alpha s */
            alpha[l] = (F)(-1.0) * alpha[l];
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; i++)
            for (int l = 0; l < nv; l++)
                t_[i * nv + l] = (F)1.0 * z[i * nv + l] + alpha[l] * s[i * nv + l];
        for (int l = 0; l < nv; l++)
            alpha[l] = (F)(-1.0) * alpha[l];
        SFX(orc_spmv)(A, t_, t, nv); /* 11 */
        for (int l = 0; l < nv; l++) /* 12: r = r - alpha v */
            alpha[l] = (F)(-1.0) * alpha[l];
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; i++)
            for (int l = 0; l < nv; l++)
                r[i * nv + l] = alpha[l] * v[i * nv + l] + (F)1.0 * r[i * nv + l];
        for (int l = 0; l < nv; l++)
            alpha[l] = (F)(-1.0) * alpha[l];
        SFX(orc_dot)(t, r, n, nv, rt->node->nchunks, theta); /* 13 */
        SFX(orc_dot)(t, t, n, nv, rt->node->nchunks, phi);
        SFX(orc_dot)(t, r0, n, nv, rt->node->nchunks, psi);
        SFX(orc_dot)(r, r, n, nv, rt->node->nchunks, eta);
        SFX(solve)(rt->precond, z, 1, t, NULL); /* 14 */
        for (int l = 0; l < nv; l++)            /* 15 */
            omega[l] = theta[l] / phi[l];
        for (int l = 0; l < nv; l++) /* 16: r = r - omega t */
            omega[l] = (F)(-1.0) * omega[l];
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; i++)
            for (int l = 0; l < nv; l++)
                r[i * nv + l] = omega[l] * t[i * nv + l] + (F)1.0 * r[i * nv + l];
        for (int l = 0; l < nv; l++)
            omega[l] = (F)(-1.0) * omega[l];
        for (int l = 0; l < nv; l++) { /* 17 */
            res[l] = omega[l] * theta[l];
            res[l] = (F)1.0 * eta[l] + (F)(-1.0) * res[l];
        }
        for (int l = 0; l < nv; l++) { /* 23: x += alpha v_ + omega t_ */
            ac[l] = iconv[l] * alpha[l];
            oc[l] = iconv[l] * omega[l];
        }
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; i++)
            for (int l = 0; l < nv; l++)
                x[i * nv + l] = ac[l] * v_[i * nv + l] + oc[l] * t_[i * nv + l] +
                                (F)1.0 * x[i * nv + l];
        if (SFX(conv_is_converged_rr)(rt, res, iconv, x, 0, b))
            return;
        for (int l = 0; l < nv; l++) { /* 21-22 */
            rho[l] = omega[l] * psi[l];
            rho[l] = (F)(-1.0) * rho[l];
            temp[l] = delta[l] * omega[l];
            beta[l] = rho[l] / temp[l];
        }
        for (int l = 0; l < nv; l++) /* 24: z = t_ - omega z (omega stays negated) */
            omega[l] = (F)(-1.0) * omega[l];
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; i++)
            for (int l = 0; l < nv; l++)
                z[i * nv + l] = (F)1.0 * t_[i * nv + l] + omega[l] * z[i * nv + l];
        for (int l = 0; l < nv; l++) { /* 25: v_ = z + beta v_ - beta omega s */
            delta[l] = omega[l] * beta[l];
            omega[l] = (F)(-1.0) * omega[l];
        }
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; i++)
            for (int l = 0; l < nv; l++)
                v_[i * nv + l] = (F)1.0 * z[i * nv + l] + delta[l] * s[i * nv + l] +
                                 beta[l] * v_[i * nv + l];
    }
}

/* ---- Jacobi jacobi.inl:51-102 + merged_jacobi blas_merged_jacobi.inl:32-90 ---- */
static void SFX(solve_jacobi)(SFX(rt) * rt, F *x, int xzero, const F *b, F *iconv_in) {
    int64_t n = rt->n;
    int nv = rt->nv;
    const orc_csr *A = &rt->A;
    F iconv[ORC_MAX_NV], relax[ORC_MAX_NV], rn[ORC_MAX_NV];
    for (int l = 0; l < nv; l++) {
        iconv[l] = iconv_in ? iconv_in[l] : (F)1.0;
        relax[l] = (F)rt->node->relax;
    }
    F *r = rt->buf[0];
    const F *d = rt->inv_diag;
    SFX(conv_t) *c = &rt->conv;
    SFX(conv_init)(c, iconv);
    int max_iters = rt->node->max_iters;
    if (c->conv_check) {
        SFX(residual)(A, x, xzero, b, r, nv);
        SFX(norm2)(rt, r, rn);
        SFX(conv_set_initial)(c, rn);
        if (SFX(conv_is_converged)(c, rn, iconv))
            return;
        if (xzero)
            memset(x, 0, sizeof(F) * n * nv);
        for (int last = 0; !last;) {
            int it = ++c->iter;
            if (it >= max_iters)
                last = 1;
            for (int64_t i = 0; i < n; i++)
                for (int l = 0; l < nv; l++)
                    r[i * nv + l] *= d[i];
            for (int l = 0; l < nv; l++)
                relax[l] *= iconv[l];
            for (int64_t i = 0; i < n; i++)
                for (int l = 0; l < nv; l++)
                    x[i * nv + l] = relax[l] * r[i * nv + l] + (F)1.0 * x[i * nv + l];
            SFX(residual)(A, x, 0, b, r, nv);
            SFX(norm2)(rt, r, rn);
            if (SFX(conv_is_converged)(c, rn, iconv))
                return;
        }
    } else {
        int xz = xzero;
        for (int last = 0; !last;) {
            int it = ++c->iter;
            if (it >= max_iters)
                last = 1;
            if (xz) {
                /* Ax_y of a zero vector leaves r zero: x = relax*inv_diag*b */
#pragma omp parallel for schedule(static)
                for (int64_t i = 0; i < n; i++)
                    for (int l = 0; l < nv; l++)
                        x[i * nv + l] = relax[l] * d[i] * b[i * nv + l];
                xz = 0;
            } else {
                SFX(orc_spmv)(A, x, r, nv);
#pragma omp parallel for schedule(static)
                for (int64_t i = 0; i < n; i++)
                    for (int l = 0; l < nv; l++)
                        x[i * nv + l] += relax[l] * d[i] * (b[i * nv + l] - r[i * nv + l]);
            }
        }
    }
}

/* ---- Chebyshev chebyshev.inl:128-183 ---- */
static void SFX(solve_chebyshev)(SFX(rt) * rt, F *x, int xzero, const F *b) {
    int64_t n = rt->n;
    int nv = rt->nv;
    const orc_csr *A = &rt->A;
    F *r = rt->buf[0], *t1 = rt->buf[1], *t2 = rt->buf[2];
    const F *d = rt->inv_sqrt_diag;
    int k = rt->node->poly_order;
    SFX(conv_t) *c = &rt->conv;
    F rn[ORC_MAX_NV];
    SFX(conv_init)(c, NULL);
    SFX(residual)(A, x, xzero, b, r, nv);
    if (c->conv_check) {
        SFX(norm2)(rt, r, rn);
        SFX(conv_set_initial)(c, rn);
        if (SFX(conv_is_converged)(c, rn, NULL))
            return;
    }
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++)
        for (int l = 0; l < nv; l++) {
            r[i * nv + l] *= d[i];
            t1[i * nv + l] = rt->cheb[k - 1] * r[i * nv + l];
        }
    for (int q = k - 2; q >= 0; --q) {
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; i++)
            for (int l = 0; l < nv; l++)
                t1[i * nv + l] *= d[i];
        SFX(orc_spmv)(A, t1, t2, nv);
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; i++)
            for (int l = 0; l < nv; l++) {
                t2[i * nv + l] *= d[i];
                t1[i * nv + l] = rt->cheb[q] * r[i * nv + l] + (F)1.0 * t2[i * nv + l];
            }
    }
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++)
        for (int l = 0; l < nv; l++) {
            t1[i * nv + l] *= d[i];
            if (xzero)
                x[i * nv + l] = (F)1.0 * t1[i * nv + l];
            else
                x[i * nv + l] = (F)1.0 * t1[i * nv + l] + (F)1.0 * x[i * nv + l];
        }
    if (c->conv_check) {
        SFX(residual)(A, x, 0, b, r, nv);
        SFX(norm2)(rt, r, rn);
        SFX(conv_is_converged)(c, rn, NULL);
    }
}

/* ---- HSGS hsgs.inl:52-91 + blas2::SGS blas2.h:184-281 + csr_SGS csr_gs.inl:32-134.
 * The rank's rows are split into gs_blocks core blocks (part.h:115-126 with
 * split_range); each block relaxes its diagonal part sequentially with the
 * off-block couplings frozen at the sweep's start (Jacobi between blocks). ---- */
static void SFX(sgs_sweep)(SFX(rt) * rt, const F *b, F *x, F *t, const F *relax) {
    const orc_csr *A = &rt->A;
    const F *val = (const F *)A->val;
    const uint32_t *row = A->row, *col = A->col;
    int64_t n = rt->n;
    int nv = rt->nv;
    int P = rt->node->gs_blocks > 0 ? rt->node->gs_blocks : 1;
    const F *d = rt->inv_diag;
    /* t = 0; t += A_offd x (offd entries in ascending column order); t = 1*b + (-1)*t */
#pragma omp parallel for schedule(static)
    for (int k = 0; k < P; k++) {
        uint64_t off, len;
        orc_core_block((uint64_t)n, P, k, (int)sizeof(double), &off, &len);
        for (uint64_t i = off; i < off + len; i++) {
            F acc[ORC_MAX_NV];
            for (int l = 0; l < nv; l++)
                acc[l] = 0.0;
            for (uint32_t j = row[i]; j < row[i + 1]; j++) {
                uint64_t cj = col[j];
                if (cj >= off && cj < off + len)
                    continue;
                for (int l = 0; l < nv; l++)
                    acc[l] += x[cj * nv + l] * val[j];
            }
            for (int l = 0; l < nv; l++)
                t[i * nv + l] = (F)1.0 * b[i * nv + l] + (F)(-1.0) * acc[l];
        }
    }
#pragma omp parallel for schedule(static)
    for (int k = 0; k < P; k++) {
        uint64_t off, len;
        orc_core_block((uint64_t)n, P, k, (int)sizeof(double), &off, &len);
        F sum[ORC_MAX_NV];
        for (uint64_t ii = 0; ii < 2 * len; ii++) {
            uint64_t i = ii < len ? off + ii : off + (2 * len - 1 - ii);
            for (int l = 0; l < nv; l++)
                sum[l] = t[i * nv + l];
            for (uint32_t j = row[i]; j < row[i + 1]; j++) {
                uint64_t cj = col[j];
                if (cj < off || cj >= off + len)
                    continue;
                for (int l = 0; l < nv; l++)
                    sum[l] -= x[cj * nv + l] * val[j];
            }
            for (int l = 0; l < nv; l++) {
                F coeff = relax[l] * d[i];
                x[i * nv + l] += coeff * sum[l];
            }
        }
    }
}

static void SFX(solve_hsgs)(SFX(rt) * rt, F *x, int xzero, const F *b, F *iconv_in) {
    int64_t n = rt->n;
    int nv = rt->nv;
    F iconv[ORC_MAX_NV], relax[ORC_MAX_NV], rn[ORC_MAX_NV];
    for (int l = 0; l < nv; l++) {
        iconv[l] = iconv_in ? iconv_in[l] : (F)1.0;
        relax[l] = (F)rt->node->relax;
    }
    F *r = rt->buf[0], *t = rt->buf[1];
    SFX(conv_t) *c = &rt->conv;
    SFX(conv_init)(c, iconv);
    if (c->conv_check) {
        SFX(residual)(&rt->A, x, xzero, b, r, nv);
        SFX(norm2)(rt, r, rn);
        SFX(conv_set_initial)(c, rn);
        if (SFX(conv_is_converged)(c, rn, iconv))
            return;
    }
    if (xzero)
        memset(x, 0, sizeof(F) * n * nv);
    int max_iters = rt->node->max_iters;
    for (int last = 0; !last;) {
        int it = ++c->iter;
        if (it >= max_iters)
            last = 1;
        for (int l = 0; l < nv; l++)
            relax[l] *= iconv[l];
        SFX(sgs_sweep)(rt, b, x, t, relax);
        if (c->conv_check) {
            SFX(residual)(&rt->A, x, 0, b, r, nv);
            SFX(norm2)(rt, r, rn);
            if (SFX(conv_is_converged)(c, rn, iconv))
                return;
        }
    }
}

/* ---- MultiGrid cycle mg.inl:186-280, uniform precision F ---- */
static void SFX(mg_cycle)(SFX(rt) * rt, int level, F *x, int xzero, const F *b, int cycle) {
    int nv = rt->nv;
    const orc_csr *A = &rt->Al[level];
    int64_t n = A->nrows;
    F *temp = rt->temp[level];
    F *bH = rt->bH[level + 1], *xH = rt->xH[level + 1];
    if (cycle != 1) {
        SFX(solve)(rt->pre[level], x, xzero, b, NULL);
        xzero = 0;
    }
    SFX(ax_y)(A, x, xzero, temp, nv);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n * nv; i++)
        temp[i] = (F)1.0 * b[i] + (F)(-1.0) * temp[i];
    SFX(orc_spmv)(&rt->Rl[level], temp, bH, nv);
    if (level + 1 == rt->nlev - 1) {
        SFX(solve)(rt->coarse, xH, 1, bH, NULL);
    } else {
        SFX(mg_cycle)(rt, level + 1, xH, 1, bH, cycle);
        if (cycle == 2)
            SFX(mg_cycle)(rt, level + 1, xH, 0, bH, 2);
        else if (cycle == 1)
            SFX(mg_cycle)(rt, level + 1, xH, 0, bH, 0);
    }
    if (xzero && cycle == 1)
        memset(x, 0, sizeof(F) * n * nv);
    SFX(orc_spmv_axpy)(&rt->Pl[level], xH, x, nv);
    SFX(solve)(rt->post[level], x, 0, b, NULL);
}

#if ORC_IS_DOUBLE
/* MG_cycle<double,float> at level 0 (mg.inl:212-226,246-256) */
static void mg_cycle_df(rt_d *rt, double *x, int xzero, const double *b, int cycle) {
    int nv = rt->nv;
    const orc_csr *A = &rt->Al[0];
    int64_t n = A->nrows, nH = rt->Al[1].nrows;
    double *temp = rt->temp[0];
    float *t32 = rt->temp32_0;
    float *bH = rt->bH32[1], *xH = rt->xH32[1];
    orc_rt_f *sub = rt->inner32; /* float view holding levels >= 1 */
    if (cycle != 1) {
        solve_d(rt->pre[0], x, xzero, b, NULL);
        xzero = 0;
    }
    ax_y_d(A, x, xzero, temp, nv);
    for (int64_t i = 0; i < n * nv; i++)
        temp[i] = 1.0 * b[i] + (-1.0) * temp[i];
    for (int64_t i = 0; i < n * nv; i++)
        t32[i] = (float)temp[i];
    orc_spmv_f(&rt->Rl[0], t32, bH, nv);
    (void)nH;
    if (rt->nlev == 2) {
        solve_f(sub->coarse, xH, 1, bH, NULL);
    } else {
        mg_cycle_f(sub, 1, xH, 1, bH, cycle);
        if (cycle == 2)
            mg_cycle_f(sub, 1, xH, 0, bH, 2);
        else if (cycle == 1)
            mg_cycle_f(sub, 1, xH, 0, bH, 0);
    }
    orc_spmv_f(&rt->Pl[0], xH, t32, nv); /* P_0 (float) x_H */
    if (xzero && cycle == 1)
        memset(x, 0, sizeof(double) * n * nv);
    for (int64_t i = 0; i < n * nv; i++)
        x[i] = 1.0 * (double)t32[i] + 1.0 * x[i];
    solve_d(rt->post[0], x, 0, b, NULL);
}
#endif

static void SFX(solve_mg)(SFX(rt) * rt, F *x, int xzero, const F *b, F *iconv_in) {
    int nv = rt->nv;
    F iconv[ORC_MAX_NV], rn[ORC_MAX_NV];
    for (int l = 0; l < nv; l++)
        iconv[l] = iconv_in ? iconv_in[l] : (F)1.0;
    F *r = rt->buf[0];
    SFX(conv_t) *c = &rt->conv;
    SFX(conv_init)(c, iconv);
    if (c->conv_check) {
        SFX(residual)(&rt->A, x, xzero, b, r, nv);
        SFX(norm2)(rt, r, rn);
        SFX(conv_set_initial)(c, rn);
        if (SFX(conv_is_converged)(c, rn, iconv))
            return;
    }
    int max_iters = rt->node->max_iters;
    int xz = xzero;
    for (int last = 0; !last;) {
        int it = ++c->iter;
        if (it >= max_iters)
            last = 1;
#if ORC_IS_DOUBLE
        if (rt->node->hier->reduced)
            mg_cycle_df(rt, x, xz, b, rt->node->mg_cycle);
        else
#endif
            SFX(mg_cycle)(rt, 0, x, xz, b, rt->node->mg_cycle);
        xz = 0;
        if (c->conv_check) {
            SFX(residual)(&rt->A, x, 0, b, r, nv);
            SFX(norm2)(rt, r, rn);
            if (SFX(conv_is_converged)(c, rn, iconv))
                return;
        }
    }
}

#if ORC_IS_DOUBLE
/* ---- IterativeRefinement ir.inl:107-194 ----
 * The inner solver keeps the constant abs/rel tolerance of its own "solver" list:
 * ir.inl:144 set_tolerance(per-RHS vector) returns early because the inner
 * convergence object already holds a constant tolerance (convergence.h:132-142). */
static void solve_ir(rt_d *rt, double *x, int xzero, const double *b) {
    int64_t n = rt->n;
    int nv = rt->nv;
    double *r = rt->buf[0], *xc = rt->buf[1];
    double rn[ORC_MAX_NV];
    orc_conv_t_d *c = &rt->conv;
    conv_init_d(c, NULL);
    residual_d(&rt->A, x, xzero, b, r, nv);
    if (c->conv_check) {
        norm2_d(rt, r, rn);
        conv_set_initial_d(c, rn);
        if (conv_is_converged_d(c, rn, NULL))
            return;
    }
    if (xzero)
        memset(x, 0, sizeof(double) * n * nv);
    int max_iters = rt->node->max_iters;
    for (int last = 0; !last;) {
        int it = ++c->iter;
        if (it >= max_iters)
            last = 1;
        if (rt->node->reduced) {
            for (int64_t i = 0; i < n * nv; i++)
                rt->ib32[i] = (float)r[i];
            solve_f(rt->inner32, rt->ix32, 1, rt->ib32, NULL);
            for (int64_t i = 0; i < n * nv; i++)
                xc[i] = (double)rt->ix32[i];
            c->sub_iter += rt->inner32->conv.iter;
            c->rr_checks += rt->inner32->conv.rr_checks; /* the inner solves' RR record */
            c->rr_stopped += rt->inner32->conv.rr_stopped;
        } else {
            solve_d(rt->precond, xc, 1, r, NULL);
            c->sub_iter += rt->precond->conv.iter;
            c->rr_checks += rt->precond->conv.rr_checks;
            c->rr_stopped += rt->precond->conv.rr_stopped;
        }
        for (int64_t i = 0; i < n * nv; i++)
            x[i] = 1.0 * xc[i] + 1.0 * x[i];
        residual_d(&rt->A, x, 0, b, r, nv);
        if (c->conv_check) {
            norm2_d(rt, r, rn);
            if (conv_is_converged_d(c, rn, NULL))
                return;
        }
    }
}
#endif

static void SFX(solve)(SFX(rt) * rt, F *x, int xzero, const F *b, F *iconv) {
    int64_t N = rt->n * rt->nv;
    switch (rt->node->method) {
    case ORC_IDENTITY:
        memcpy(x, b, sizeof(F) * N);
        break;
    case ORC_BICGSTAB:
        SFX(solve_bicgstab)(rt, x, xzero, b, iconv, 0);
        break;
    case ORC_PBICGSTAB:
        SFX(solve_bicgstab)(rt, x, xzero, b, iconv, 1);
        break;
    case ORC_PCG:
        SFX(solve_pcg)(rt, x, xzero, b, iconv);
        break;
    case ORC_RBICGSTAB:
        SFX(solve_rbicgstab)(rt, x, xzero, b, iconv);
        break;
    case ORC_JACOBI:
        SFX(solve_jacobi)(rt, x, xzero, b, iconv);
        break;
    case ORC_CHEBYSHEV:
        SFX(solve_chebyshev)(rt, x, xzero, b);
        break;
    case ORC_HSGS:
        SFX(solve_hsgs)(rt, x, xzero, b, iconv);
        break;
    case ORC_MULTIGRID:
        SFX(solve_mg)(rt, x, xzero, b, iconv);
        break;
    case ORC_DIRECT:
        SFX(dense_mv)(rt->ainv, rt->n, b, x, rt->nv);
        break;
#if ORC_IS_DOUBLE
    case ORC_IR:
        solve_ir(rt, x, xzero, b);
        break;
#endif
    default:
        break;
    }
}

static void SFX(csr_to_dense)(const orc_csr *A, F *a) {
    const F *val = (const F *)A->val;
    int64_t n = A->nrows;
    memset(a, 0, sizeof(F) * n * n);
    for (int64_t i = 0; i < n; i++)
        for (uint32_t j = A->row[i]; j < A->row[i + 1]; j++)
            a[i * n + A->col[j]] = val[j];
}

static int SFX(nbufs)(int method) {
    switch (method) {
    case ORC_BICGSTAB:
        return 6;
    case ORC_PBICGSTAB:
        return 8;
    case ORC_PCG:
        return 4;
    case ORC_RBICGSTAB:
        return 8;
    case ORC_JACOBI:
        return 1;
    case ORC_CHEBYSHEV:
        return 3;
    case ORC_HSGS:
        return 2;
    case ORC_MULTIGRID:
        return 1;
    case ORC_IR:
        return 2;
    default:
        return 0;
    }
}

/* Setup of one node on matrix A (solver.h:147-161 init_base, plus the per-method
 * init(): jacobi.inl:36-43, chebyshev.inl:36-44, hsgs.inl:36-44, direct.inl:37-44,
 * mg.inl:117-184, ir.inl:61-93). */
SFX(rt) * SFX(rt_build)(const orc_node *node, const orc_csr *A, int nv, int level) {
    SFX(rt) *rt = (SFX(rt) *)calloc(1, sizeof(SFX(rt)));
    rt->node = node;
    rt->A = *A;
    rt->n = A->nrows;
    rt->nv = nv;
    rt->conv.nv = nv;
    rt->conv.conv_check = node->conv_check;
    rt->conv.abs_tol = node->abs_tol;
    rt->conv.rel_tol = node->rel_tol;
    rt->conv.rr = node->rr;
    rt->conv.rr_stepping = (double)(float)node->rr_stepping;
    rt->conv.rr_deviation = (double)(float)node->rr_deviation;
    if (node->rr)
        rt->rrbuf = SFX(vnew)(A->nrows, nv);
    rt->nbuf = SFX(nbufs)(node->method);
    for (int k = 0; k < rt->nbuf; k++)
        rt->buf[k] = SFX(vnew)(rt->n, nv);
    switch (node->method) {
    case ORC_PBICGSTAB:
    case ORC_PCG:
    case ORC_RBICGSTAB:
        rt->precond = SFX(rt_build)(node->precond, A, nv, level);
        break;
    case ORC_JACOBI:
    case ORC_HSGS:
        rt->inv_diag = SFX(vnew)(rt->n, 1);
        SFX(orc_inv_diag)(A, rt->inv_diag);
        break;
    case ORC_CHEBYSHEV: {
        rt->inv_sqrt_diag = SFX(vnew)(rt->n, 1);
        SFX(orc_inv_sqrt_diag)(A, rt->inv_sqrt_diag);
        double cc[4];
        orc_cheb_coeffs(node->min_eig, node->max_eig, node->poly_order, node->spec_frac, cc);
        for (int q = 0; q < node->poly_order; q++)
            rt->cheb[q] = (F)cc[q];
        break;
    }
    case ORC_DIRECT: {
        F *a = SFX(vnew)(rt->n * rt->n, 1);
        SFX(csr_to_dense)(A, a);
        rt->ainv = SFX(vnew)(rt->n * rt->n, 1);
        SFX(orc_dense_inverse)(a, rt->n, rt->ainv);
        free(a);
        break;
    }
    case ORC_MULTIGRID: {
        const orc_hierarchy *h = node->hier;
        rt->nlev = h->nlevels;
#if ORC_IS_DOUBLE
        if (h->reduced) {
            /* level 0 in double, levels >= 1 + all transfers in float */
            rt->Al[0] = h->A[0];
            orc_node *sm_pre = node->pre, *sm_post = node->post;
            orc_node pre0 = *sm_pre, post0 = *sm_post;
            pre0.min_eig = h->min_eig[0];
            pre0.max_eig = h->max_eig[0];
            post0.min_eig = h->min_eig[0];
            post0.max_eig = h->max_eig[0];
            orc_node *p0 = (orc_node *)malloc(sizeof(orc_node));
            orc_node *q0 = (orc_node *)malloc(sizeof(orc_node));
            *p0 = pre0;
            *q0 = post0;
            rt->pre[0] = rt_build_d(p0, &h->A[0], nv, 0);
            rt->post[0] = rt_build_d(q0, &h->A[0], nv, 0);
            rt->temp[0] = vnew_d(h->A[0].nrows, nv);
            rt->temp32_0 = vnew_f(h->A[0].nrows, nv);
            rt->Rl[0] = h->R32[0];
            rt->Pl[0] = h->P32[0];
            /* float sub-hierarchy for levels >= 1 */
            orc_rt_f *sub = (orc_rt_f *)calloc(1, sizeof(orc_rt_f));
            sub->node = node;
            sub->nv = nv;
            sub->nlev = h->nlevels;
            for (int l = 1; l < h->nlevels; l++) {
                sub->Al[l] = h->A32[l];
                sub->bH[l] = vnew_f(h->A32[l].nrows, nv);
                sub->xH[l] = vnew_f(h->A32[l].nrows, nv);
                sub->temp[l] = vnew_f(h->A32[l].nrows, nv);
                if (l < h->nlevels - 1) {
                    sub->Pl[l] = h->P32[l];
                    sub->Rl[l] = h->R32[l];
                    orc_node *pl = (orc_node *)malloc(sizeof(orc_node));
                    orc_node *ql = (orc_node *)malloc(sizeof(orc_node));
                    *pl = *sm_pre;
                    *ql = *sm_post;
                    pl->min_eig = ql->min_eig = h->min_eig[l];
                    pl->max_eig = ql->max_eig = h->max_eig[l];
                    sub->pre[l] = rt_build_f(pl, &h->A32[l], nv, l);
                    sub->post[l] = rt_build_f(ql, &h->A32[l], nv, l);
                }
            }
            sub->coarse = rt_build_f(node->coarse, &h->A32[h->nlevels - 1], nv, h->nlevels - 1);
            rt->bH32[1] = sub->bH[1];
            rt->xH32[1] = sub->xH[1];
            rt->inner32 = sub;
            break;
        }
#endif
        for (int l = 0; l < h->nlevels; l++) {
#if ORC_IS_DOUBLE
            rt->Al[l] = h->A[l];
            if (l < h->nlevels - 1) {
                rt->Pl[l] = h->P[l];
                rt->Rl[l] = h->R[l];
            }
#else
            rt->Al[l] = h->A32[l];
            if (l < h->nlevels - 1) {
                rt->Pl[l] = h->P32[l];
                rt->Rl[l] = h->R32[l];
            }
#endif
            rt->temp[l] = SFX(vnew)(rt->Al[l].nrows, nv);
            rt->bH[l] = SFX(vnew)(rt->Al[l].nrows, nv);
            rt->xH[l] = SFX(vnew)(rt->Al[l].nrows, nv);
            if (l < h->nlevels - 1) {
                orc_node *pl = (orc_node *)malloc(sizeof(orc_node));
                orc_node *ql = (orc_node *)malloc(sizeof(orc_node));
                *pl = *node->pre;
                *ql = *node->post;
                pl->min_eig = ql->min_eig = h->min_eig[l];
                pl->max_eig = ql->max_eig = h->max_eig[l];
                rt->pre[l] = SFX(rt_build)(pl, &rt->Al[l], nv, l);
                rt->post[l] = SFX(rt_build)(ql, &rt->Al[l], nv, l);
            }
        }
        rt->coarse = SFX(rt_build)(node->coarse, &rt->Al[h->nlevels - 1], nv, h->nlevels - 1);
        break;
    }
#if ORC_IS_DOUBLE
    case ORC_IR:
        if (node->reduced) {
            /* matrix::convert (convert.inl) FP64 -> FP32 values, same indices */
            rt->A32val = (float *)malloc(sizeof(float) * (A->nnz > 0 ? A->nnz : 1));
            for (int64_t j = 0; j < A->nnz; j++)
                rt->A32val[j] = (float)((const double *)A->val)[j];
            rt->A32 = *A;
            rt->A32.val = rt->A32val;
            rt->inner32 = rt_build_f(node->precond, &rt->A32, nv, 0);
            rt->ix32 = vnew_f(rt->n, nv);
            rt->ib32 = vnew_f(rt->n, nv);
        } else {
            rt->precond = rt_build_d(node->precond, A, nv, 0);
        }
        break;
#endif
    default:
        break;
    }
    return rt;
}
