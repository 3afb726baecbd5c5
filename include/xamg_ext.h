/*
 * Extensions of the XAMG C API for the MI355X build.  These entry points have no
 * counterpart in the reference's api/c/xamg_c_api.h; they expose what a GPU caller
 * (and the parity tests and benchmark of this repository) needs beyond the
 * reference's interface: error text, workload generators, direct access to the
 * SpMV / dot kernels on C-API objects, per-RHS residual records and timing helpers.
 * All return 0 on success unless stated otherwise.
 */
#ifndef XAMG_EXT_H_
#define XAMG_EXT_H_

#include "xamg_c_api.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Message of the last failed call (empty string if none). */
const char *xamgx_last_error(void);

/* Device name / CU count / HBM size of the GPU in use, as text. */
int xamgx_device_info(char *buf, int len);

/* ---- workload generators (host) ---- */
/* 3D 7-point Dirichlet Laplacian: diag 6, -1 per existing neighbour, lexicographic
 * order i + nx*(j + ny*k), ascending columns.  nnz = 7n - 2(nx ny + ny nz + nx nz). */
int64_t xamgx_gen_poisson7_nnz(int nx, int ny, int nz);
int xamgx_gen_poisson7(int nx, int ny, int nz, int *row, int *col, double *val);
/* 3D 27-point finite-volume anisotropic diffusion, coefficients k=(kx,ky,kz) times a
 * per-cell log-normal(0,1) multiplier (harmonic face average), Dirichlet walls. */
int64_t xamgx_gen_aniso27_nnz(int nx, int ny, int nz);
int xamgx_gen_aniso27(int nx, int ny, int nz, double kx, double ky, double kz, uint32_t seed,
                      int *row, int *col, double *val);
/* 2D 5-point variable-coefficient diffusion on an s x s grid, log-normal(0,1) cell
 * coefficients (harmonic face average), Dirichlet walls. */
int64_t xamgx_gen_var5_nnz(int s);
int xamgx_gen_var5(int s, uint32_t seed, int *row, int *col, double *val);
/* b[i*nv+l] = 2*(pseudo_rand(k) - 0.5), k = (u32)(seed*2^27 + i*nv + l) (rand.h:33-73) */
int xamgx_fill_rhs(double *b, int64_t n, int nv, uint32_t seed);

/* ---- kernels on C-API objects ---- */
/* y = A x (mode 0, csr_Ax_y) or y += A x (mode 1, csr_Axpy) on host arrays of nrows*nv
 * elements; prec 0: FP64 values and vectors, 1: FP32 values and vectors,
 * 2: BF16 values with FP32 vectors.  x/y are double* for prec 0, float* otherwise. */
int xamgx_spmv_host(XAMG_matrix_id mid, int prec, int mode, int nv, const void *x, void *y);
/* res[l] = sum_i x[i,l]*y[i,l] on two device vectors (blas.inl:521-589) */
int xamgx_dot(XAMG_vector_id x, XAMG_vector_id y, int nv, double *res);
/* The same dot product, allocating nothing: no hipMalloc, hipFree or allocating library
 * routine.  No counterpart in the reference; it is for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that forbid it).  Its
 * per-workgroup partial sums and result live in a fixed device array of 256.1 KiB per
 * GPU (2048 workgroup slots x 16 RHS), held until the process ends.  Needs XAMG_init
 * first; rejects a reduction grid above 2048 workgroups (xamgx_set_engine grid_red) and
 * nv outside {1, 2, 4, 8, 16}.  With the same grid its result is bit-identical to
 * xamgx_dot's.  Calls on one GPU share the array: calls that overlap in time race. */
int xamgx_dot_noalloc(XAMG_vector_id x, XAMG_vector_id y, int nv, double *res);

/* ---- measurement ---- */
/* Average device time (ms) of one y = A x launch over `iters` back-to-back launches
 * after 10 warm-ups, timed with HIP events on the launch stream. */
int xamgx_bench_spmv(XAMG_matrix_id mid, int prec, int nv, int iters, double *avg_ms);
/* Sustained copy bandwidth (GB/s, read+write bytes) of a streaming copy kernel over
 * `bytes` bytes, best of `iters` launches (each timed from the kernel's start to its end). */
int xamgx_bench_copy(uint64_t bytes, int iters, double *gbs);
/* STREAM triad a = b + s c on FP64 arrays of `bytes` bytes each: GB/s of the 3 x bytes
 * moved, best of `iters` launches. */
int xamgx_bench_triad(uint64_t bytes, int iters, double *gbs);
/* Average device time (ms) of one launch of an SpMV epilogue (0 y=Ax, 1 y+=Ax, 2 residual,
 * 3 Jacobi step, 4 Chebyshev step, 5 last Chebyshev step, 6 FP64 residual stored as FP32,
 * 7 and 8 Chebyshev step and last step on the scaled operator after the first step, 9 first
 * Chebyshev step on the scaled operator, 10, 11 and 12 first, middle and last Chebyshev
 * step on the unscaled operator storing their output multiplied by the row scaling)
 * with nd fused dot products (0, 1 or 3 as the solvers use them) on operand buffers of
 * the matrix's shape.  prec 0 FP64, 1 FP32, 2 BF16 values with FP32 vectors, 3 scaled
 * FP32.  cfg = {RHS per lane, workgroups, XCD mapping, LDS staging} or NULL for the
 * autotuned choice.  cold = 1 times each launch alone after a 512 MiB copy has evicted
 * the caches (as inside a solve); 0 times iters launches back to back. */
int xamgx_bench_spmv_epi(XAMG_matrix_id mid, int prec, int nv, int epi, int nd, int cold, int iters,
                         const int *cfg, double *avg_ms);
/* Per-launch device timing of the solve path: HIP events on the launch stream around
 * every SpMV and element-wise dispatch (a dispatch with fused dot products includes its
 * finish kernel), grouped by launch class with the algorithmic HBM bytes of each launch.
 * on = 1 clears the record and starts recording, 0 stops.  Launches captured into a
 * graph are not recorded.  Has no counterpart in the reference. */
int xamgx_kernel_timing(int on);
/* Number of launch classes recorded (waits for the device). */
int xamgx_kernel_stats_count(int *count);
/* Class i: name (NUL-terminated, at most len bytes), launches, total device ms, total
 * algorithmic bytes.  NULL skips an output.  Launches that returned at their flags
 * without doing their work (queued behind a converged solver, or gated off), taken to
 * be those that finished faster than their algorithmic bytes at 30 TB/s, are left out
 * of all three. */
int xamgx_kernel_stats_get(int i, char *name, int len, int64_t *calls, double *ms, double *bytes);
/* Device time (ms) of the last solve (first kernel to the convergence readback). */
int xamgx_solver_last_ms(XAMG_solver_id sid, double *ms);

/* ---- solver records ---- */
/* Residual history of the top solver: rows of nv squared residual norms, -1 for an
 * RHS already converged (print.inl:119-142 table without the sqrt). */
int xamgx_solver_get_history(XAMG_solver_id sid, double *out, int cap_rows, int *nrows);
/* Squared initial / final residual norms per RHS (convergence.h res0/res). */
int xamgx_solver_get_residuals(XAMG_solver_id sid, double *res0, double *res);
/* Residual replacement record of the last solve (convergence_details RR,
 * convergence.h:269-302): true residuals computed at the RR_stepping checkpoints, and
 * whether a recurrence/true deviation beyond RR_deviation ended the solve.  For
 * IterativeRefinement, summed over its inner solves (stopped = inner solves ended so).
 * Only the Krylov solvers compute them; elsewhere the checked residual is the true one.
 * NULL skips an output. */
int xamgx_solver_get_rr_info(XAMG_solver_id sid, int *checks, int *stopped);

/* ---- configuration ---- */
/* Batch of iterations between host polls of the device halt flag, HIP-graph replay
 * (0/1), grid sizes of streaming / reduction kernels, XCD-aware row mapping (0/1).
 * A negative value keeps the current setting. */
int xamgx_set_engine(int batch, int use_graphs, int grid_spmv, int grid_ew, int grid_red,
                     int xcd_remap);
/* Right-hand sides per lane of the SpMV kernel for nv and element size es (8: FP64
 * vectors, 4: FP32 vectors); 0 restores the default.  Used by the autotuner. */
int xamgx_set_spmv_vpt(int nv, int es, int vpt);
/* Staging mode of every SpMV launch: -1 as autotuned (default), 0 direct loads, 1 each
 * row tile's column indices and values staged through LDS, 2 as 1 with the streamed
 * operands loaded non-temporally, 3 double-buffered staging by direct global-to-LDS
 * loads (BF16 values: as 1), 4 and 5 as 2 and 3 with the column indices read as 16-bit
 * offsets from the row (matrices whose entries all lie within 32767 of the diagonal;
 * otherwise as 2 and 3).  Results are identical in every mode. */
int xamgx_set_spmv_stage(int mode);
/* Spectrum bounds of D^{-1/2} A D^{-1/2} used by Chebyshev (the reference takes
 * them from hypre's setup, hypre_base.inl:369-375). */
int xamgx_matrix_set_eigs(XAMG_matrix_id mid, double lo, double hi);
int xamgx_matrix_get_eigs(XAMG_matrix_id mid, double *lo, double *hi);

/* ---- multigrid hierarchies ----
 * The reference builds its MultiGrid levels with hypre BoomerAMG and imports A_l,
 * P_l, R_l = P_l^T and spectrum estimates (hypre_base.inl:340-406).  Here a host
 * classical AMG setup (Ruge-Stueben coarsening, classical or direct interpolation,
 * Galerkin products) driven by the same MultiGrid keys builds them, or a caller
 * imports its own levels.  Level l >= 1 holds A_l; P_l (l < levels-1) interpolates
 * level l+1 to level l; level 0 is the caller's matrix.  Indices are int like
 * XAMG_csr_matrix_fill.  Setup runs on the host and needs no GPU. */
typedef size_t XAMG_hierarchy_id;
/* Setup from a constructed matrix and the MultiGrid list of a parameter dictionary
 * (mg_strength_threshold, mg_max_row_sum, mg_coarse_matrix_size, mg_max_levels,
 * mg_interpolation_type 0 classical / 3 direct, mg_trunc_factor, mg_Pmax_elements). */
int xamgx_hierarchy_build(XAMG_matrix_id mid, XAMG_param_id pid, XAMG_hierarchy_id *hid);
/* Same setup from host CSR arrays (nrows x nrows). */
int xamgx_hierarchy_build_csr(XAMG_param_id pid, uint64_t nrows, const int *row, const int *col,
                              const double *val, XAMG_hierarchy_id *hid);
/* Empty hierarchy, then one call per coarse level: P (nfine x ncoarse) and A_coarse. */
int xamgx_hierarchy_create(XAMG_hierarchy_id *hid);
int xamgx_hierarchy_add_level(XAMG_hierarchy_id hid, uint64_t nfine, uint64_t ncoarse,
                              const int *prow, const int *pcol, const double *pval,
                              const int *arow, const int *acol, const double *aval);
int xamgx_hierarchy_num_levels(XAMG_hierarchy_id hid, int *nlevels);
/* sizes = {rows of level l, nnz(A_l) (0 for l = 0), columns of P_l, nnz(P_l)} */
int xamgx_hierarchy_level_size(XAMG_hierarchy_id hid, int level, uint64_t sizes[4]);
/* Copies A_l (l >= 1) and P_l (l < levels-1) out; a NULL pointer skips that array. */
int xamgx_hierarchy_level_get(XAMG_hierarchy_id hid, int level, int *arow, int *acol, double *aval,
                              int *prow, int *pcol, double *pval);
/* Spectrum bounds of D^{-1/2} A_l D^{-1/2} (l >= 1) for Chebyshev smoothers; get
 * estimates them by Lanczos when unset (the estimate a solver would make). */
int xamgx_hierarchy_set_eigs(XAMG_hierarchy_id hid, int level, double lo, double hi);
int xamgx_hierarchy_get_eigs(XAMG_hierarchy_id hid, int level, double *lo, double *hi);
/* MultiGrid solvers created on the matrix afterwards use this hierarchy (uploaded to
 * HBM now) instead of running the setup. */
int xamgx_matrix_set_hierarchy(XAMG_matrix_id mid, XAMG_hierarchy_id hid);
int xamgx_hierarchy_destroy(XAMG_hierarchy_id hid);

/* ---- system files and result records ----
 * Binary system file of the reference's test drivers (mtx_gen.cpp:116-147 writes it,
 * reader.inl:47-177 reads it): u32 nrows, u32 nnz, u32 row[nrows+1], u32 col[nnz],
 * f64 val[nnz], then optionally f64 b[nrows] and f64 x[nrows].  A file without b reads
 * as b = 1 and one without x as x = 0 (the reference reader's fallbacks). */
int xamgx_read_system_info(const char *path, uint64_t *nrows, uint64_t *nnz, int *has_b, int *has_x);
/* Arrays of nrows+1, nnz, nnz, nrows, nrows entries; NULL skips one. */
int xamgx_read_system(const char *path, int *row, int *col, double *val, double *b, double *x);
/* b and x may be NULL (x only when b is given too). */
int xamgx_write_system(const char *path, uint64_t nrows, const int *row, const int *col,
                       const double *val, const double *b, const double *x);
/* YAML record of the last solve in the layout of the reference's test driver
 * (tst_solver.h:155-179, tst_output.h:32-91): info (nrows, nonzeros, NV, nprocs,
 * config), solver (X_initial_norm when x0_norm is given, X_norm, Y_norm, iters,
 * converged, abs_residual, rel_residual, initial_residual), timing (solver seconds),
 * topo. */
int xamgx_solver_write_record(XAMG_solver_id sid, XAMG_vector_id x, XAMG_vector_id b,
                              const double *x0_norm, double seconds, const char *config,
                              const char *path);

/* ---- batches of independent systems ----
 * Many independent systems of uneven sizes, each with its own nv (1..4), solved
 * together: the reference solves them one by one (one solver per system, NV fixed at
 * compile time); here one launch per iteration step covers all of them and systems
 * drop out as they converge.  The matrix is the concatenated block-diagonal CSR with
 * global indices (int, as XAMG_csr_matrix_fill); sys_rows[k] is the first row of
 * system k (nsys + 1 entries, sys_rows[0] = 0).  Vectors concatenate each system's
 * [n_k][nv_k] row-interleaved block in system order; pass them as device vectors of
 * sum_k n_k nv_k rows and nv = 1. */
typedef size_t XAMG_batch_id;
int xamgx_batch_create(int nsys, const int *sys_rows, const int *nv, const int *row, const int *col,
                       const double *val, XAMG_batch_id *bid);
/* Solver from a parameter dictionary: "solver" BiCGStab or PBiCGStab with a Jacobi
 * "preconditioner" (max_iters sweeps, relax_factor) in FP64 (pbicgstab.inl semantics),
 * or under a "meta_solver" IterativeRefinement with subsolver_reduced_precision: FP32
 * values and recurrences, FP64 solution and residual replacement, convergence on the
 * FP64 residual (replace_step, default 1e-3 here; recurrence_operator fp64 (default) or
 * fp32 for the recurrence products; preconditioner_operator fp32 (default) or fp64 for
 * the Jacobi sweeps). */
int xamgx_batch_setup(XAMG_batch_id bid, XAMG_param_id pid);
int xamgx_batch_solve(XAMG_batch_id bid, XAMG_vector_id x, XAMG_vector_id b);
/* Per system: Krylov iterations and FP64 residual checks (mixed precision); per RHS in
 * vector order: convergence flags.  NULL skips an output. */
int xamgx_batch_get_info(XAMG_batch_id bid, int *iters, int *checks, int *converged);
/* FP64 relative residuals ||b - A x|| / ||b|| per RHS, in vector order. */
int xamgx_batch_residuals(XAMG_batch_id bid, XAMG_vector_id x, XAMG_vector_id b, double *relres);
/* Device time (ms) of the last batch solve. */
int xamgx_batch_last_ms(XAMG_batch_id bid, double *ms);
/* Workgroups of the grid, rows per tile, iterations per host poll, hipGraph replay (0/1);
 * values <= 0 (< 0 for use_graph) keep the current setting.  The tile is the block over
 * which dot products are summed first (default 1024 rows), so it fixes their order.  A
 * grid > 0 also switches the autotuner off (one tile per workgroup step, one element per
 * SpMV lane). */
int xamgx_batch_set_launch(int grid, int tile_rows, int batch, int use_graph);
/* Autotuner of the launch shape (on by default): at setup, and again whenever the tile
 * size changes, it times workgroups {2048, ..., 32768} x tiles per workgroup step
 * {1, 2, 4} x elements per SpMV lane {1, 2} on the batch itself and keeps the fastest.
 * None of these changes a result.  With on = 0, tps and ept (> 0) fix the last two. */
int xamgx_batch_set_tuning(int on, int tps, int ept);
/* Launch shape of the batch's solver: workgroups, tiles per step, elements per SpMV lane,
 * and 1 if the autotuner chose it. */
int xamgx_batch_get_launch(XAMG_batch_id bid, int *grid, int *tps, int *ept, int *tuned);
int xamgx_batch_destroy(XAMG_batch_id bid);
/* Workload generator: nsys 2D 5-point variable-coefficient systems, side
 * round(smin (smax/smin)^U_k) and nv_k = 1 + floor(maxnv V_k) with U_k, V_k the
 * pseudo_rand values at seed*2^27 + 2k, + 2k + 1; system k's coefficients are those of
 * xamgx_gen_var5(side_k, seed + 7919 (k + 1)). */
int xamgx_gen_batch_var5_layout(int nsys, uint32_t seed, int smin, int smax, int maxnv, int *side,
                                int *nv);
int64_t xamgx_gen_batch_var5_nnz(int nsys, const int *side);
int xamgx_gen_batch_var5(int nsys, const int *side, uint32_t seed, int *sys_rows, int *row, int *col,
                         double *val);

#ifdef __cplusplus
}
#endif

#endif /* XAMG_EXT_H_ */
