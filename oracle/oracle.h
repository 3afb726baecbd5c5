/*
 * CPU oracle for the XAMG solve path -- TEST INFRASTRUCTURE ONLY.
 *
 * This is a plain-C restatement of the reference algorithms (MatveevAlexy/xamg-test,
 * header-only C++ under src/xamg/) used as the checker for the HIP implementation
 * and as the CPU baseline that bench.py times beside it.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; the product
 * library never links or calls it.
 *
 * Every function names the reference file:line it follows.  Operation order inside
 * each loop is the reference's (no FMA contraction: build with -ffp-contract=off),
 * so element-wise results and single-chunk reductions are bit-identical to the
 * reference compiled for one MPI rank with one core.
 *
 * Parity pins (what checks each function; see tests/test_oracle_*.py):
 *   orc_pseudo_rand_*        -- closed form of rand.h + golden stream values
 *   orc_spmv_*, orc_dot_*    -- integer-valued stencils (exact arithmetic), scipy,
 *                               mtx_Paus fixtures (norms from out.yml)
 *   orc_solve (BiCGStab/PBiCGStab/Jacobi/Chebyshev/HSGS/MG/Direct/IR)
 *                            -- dense LU solution on small systems, out.yml X_norm,
 *                               closed-form Chebyshev polynomial, brute-force GS,
 *                               textbook BiCGStab re-derivation in numpy
 */
#ifndef XAMG_ORACLE_H
#define XAMG_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAX_NV 16

/* CSR block with u32 indices, as csr_matrix<F,u32,u32,u32,u32> (csr_matrix.h:47-559). */
typedef struct {
    int64_t nrows, ncols, nnz;
    const uint32_t *row;
    const uint32_t *col;
    const void *val; /* double* or float* depending on the routine suffix */
} orc_csr;

/* ---- rand.h:33-73 ---- */
int32_t orc_pseudo_rand_i(uint32_t i);
double orc_pseudo_rand_d(uint64_t i);
/* b[i*nv+l] = 2*(pseudo_rand<double>(k) - 0.5), k = (u32)(seed*2^27 + i*nv + l) */
void orc_fill_rhs(double *b, int64_t n, int nv, uint32_t seed);

/* ---- misc.h:47-56 split_range + vector.inl:198-236 core ranges ---- */
uint64_t orc_split_range(uint64_t size, int nblocks, int elem_size);
void orc_core_block(uint64_t size, int nblocks, int k, int elem_size, uint64_t *off, uint64_t *len);

/* ---- csr_matvec.inl:148-260 (Ax_y) and :32-144 (Axpy) ---- */
void orc_spmv_d(const orc_csr *A, const double *x, double *y, int nv);
void orc_spmv_f(const orc_csr *A, const float *x, float *y, int nv);
void orc_spmv_axpy_d(const orc_csr *A, const double *x, double *y, int nv);
void orc_spmv_axpy_f(const orc_csr *A, const float *x, float *y, int nv);

/* ---- blas.inl:521-589 dot: per-chunk sequential sums, chunks summed in order
 *      (shm_allreduce.h:123-134).  nchunks=1 is a single-core rank. ---- */
void orc_dot_d(const double *x, const double *y, int64_t n, int nv, int nchunks, double *res);
void orc_dot_f(const float *x, const float *y, int64_t n, int nv, int nchunks, float *res);

/* ---- blas.inl:431-484 bxor_global ---- */
void orc_bxor_d(const double *x, int64_t n, int nv, uint64_t *res);

/* ---- csr_matrix.h:323-389 inv_diag / inv_sqrt_diag ---- */
void orc_inv_diag_d(const orc_csr *A, double *d);
void orc_inv_sqrt_diag_d(const orc_csr *A, double *d);
void orc_inv_diag_f(const orc_csr *A, float *d);
void orc_inv_sqrt_diag_f(const orc_csr *A, float *d);

/* ---- inverse.inl:31-104 Gauss-Jordan (no pivoting) of a dense row-major matrix ---- */
void orc_dense_inverse_d(const double *a, int64_t n, double *ainv);
void orc_dense_inverse_f(const float *a, int64_t n, float *ainv);

/* ---- chebyshev.inl:67-125 coefficients ---- */
void orc_cheb_coeffs(double min_eig, double max_eig, int order, double fraction, double *c);

/* ---------------------------------------------------------------------------
 * Solver hierarchy.  A node mirrors one base_solver (solver.h:77-254) with its
 * parameter list (xamg_params_details.h:171-289 defaults are applied by the
 * caller).  MG levels carry the host-built hierarchy (hypre_base.inl:340-406).
 * ------------------------------------------------------------------------- */
enum {
    ORC_IDENTITY = 0,
    ORC_BICGSTAB = 1,
    ORC_PBICGSTAB = 2,
    ORC_JACOBI = 3,
    ORC_CHEBYSHEV = 4,
    ORC_HSGS = 5,
    ORC_MULTIGRID = 6,
    ORC_DIRECT = 7,
    ORC_IR = 8,
    ORC_PCG = 9,
    ORC_RBICGSTAB = 10
};

#define ORC_MAX_LEVELS 32

typedef struct orc_node orc_node;

typedef struct {
    int nlevels;
    /* level l system matrix (double and, for reduced levels, float values) */
    orc_csr A[ORC_MAX_LEVELS], A32[ORC_MAX_LEVELS];
    /* transfer operators for l -> l+1 (P: n_l x n_{l+1}, R = P^T) */
    orc_csr P[ORC_MAX_LEVELS], R[ORC_MAX_LEVELS], P32[ORC_MAX_LEVELS], R32[ORC_MAX_LEVELS];
    double min_eig[ORC_MAX_LEVELS], max_eig[ORC_MAX_LEVELS];
    int reduced; /* mg_reduced_precision: levels >= 1 and all P/R in float */
} orc_hierarchy;

struct orc_node {
    int method;
    int max_iters;
    double abs_tol, rel_tol;
    int conv_check;   /* convergence_details check-L2 vs nocheck */
    double relax;     /* Jacobi / HSGS relax_factor */
    int poly_order;   /* Chebyshev */
    double spec_frac; /* Chebyshev spectrum_fraction */
    double min_eig, max_eig;
    int gs_blocks;     /* HSGS: number of core blocks (ncores of the modelled run) */
    int mg_cycle;      /* 0 V, 1 F, 2 W */
    int reduced;       /* IR: subsolver_reduced_precision */
    orc_node *precond; /* PBiCGStab preconditioner / IR inner solver */
    orc_node *pre, *post, *coarse; /* MultiGrid smoothers (same for all levels) and coarse solver */
    const orc_hierarchy *hier;
    int nchunks; /* dot-product chunking (cores of the modelled rank) */
    /* residual replacement: convergence_details RR with the residual_control block's
     * RR_stepping / RR_deviation (xamg_params_details.h:104-105, 184-187) */
    int rr;
    double rr_stepping, rr_deviation;
};

/* Convergence record of the top solver (stats.h:120-148, convergence.h:49-303). */
typedef struct {
    int iters;
    int sub_iters;
    int converged[ORC_MAX_NV];
    double res0[ORC_MAX_NV];
    double res[ORC_MAX_NV];
    /* residual history: hist[k*nv + l] = res after check k (k=0 initial), -1 when
     * the RHS was already converged; nhist entries written */
    double *hist;
    int hist_cap;
    int nhist;
    /* residual replacement: true residuals computed, 1 if a deviation ended the solve */
    int rr_checks;
    int rr_stopped;
} orc_result;

/* Solve A x = b in FP64 (x in/out, [n][nv] interleaved). */
int orc_solve_d(const orc_node *node, const orc_csr *A, double *x, const double *b, int nv,
                orc_result *res);
/* Same hierarchy in FP32 (A->val is float*). */
int orc_solve_f(const orc_node *node, const orc_csr *A, float *x, const float *b, int nv,
                orc_result *res);

/* Setup once (buffers, diagonals, coefficients, dense inverse), then run; the run is
 * what the CPU baseline times, like the reference's wall clock around solve()
 * (examples/test/tst_solver.h:106-113). single=1 builds the FP32 tree. */
void *orc_setup(const orc_node *node, const orc_csr *A, int nv, int single);
int orc_run_d(void *ctx, double *x, int xzero, const double *b, orc_result *res);
int orc_run_f(void *ctx, float *x, int xzero, const float *b, orc_result *res);

/* Number of OpenMP threads the oracle uses for its streaming loops. */
int orc_num_threads(void);

#ifdef __cplusplus
}
#endif

#endif
