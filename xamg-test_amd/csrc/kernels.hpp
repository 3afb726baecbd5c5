// Kernel launch interface of the MI355X XAMG solve path (host side).
//
// Every streaming kernel walks a contiguous row range per workgroup, with the
// workgroup -> range map made XCD-aware (MI355X dispatches workgroups round-robin
// over 8 XCDs, each with its own L2): the workgroups of one XCD cover one contiguous
// eighth of the rows, so the gathered neighbours of a stencil stay in that XCD's L2.
//
// Reductions are deterministic: per-workgroup partials in fixed lane/wave order, then a
// one-workgroup finish kernel sums the partials in workgroup order and applies the
// solver's scalar update ("final op") on the device.  (A last-workgroup ticket scheme
// would need an agent-scope release per workgroup, i.e. an L2 write-back per workgroup
// on a multi-XCD GPU; the kernel boundary provides the visibility once.)
#pragma once

#include "common.hpp"

#include <vector>

namespace xamg {

// SpMV epilogues.  acc = sum_j x[col_j] * a_ij in ascending j (csr_matvec.inl:213-231).
enum Epi : int {
    EPI_AXY = 0,      // y = acc                                     (csr_Ax_y)
    EPI_AXPY = 1,     // y += acc                                    (csr_Axpy)
    EPI_RESID = 2,    // y = 1*b + (-1)*acc                           (solver.h:183-190)
    EPI_JACOBI = 3,   // y = xin + (w*d_i)*(b - acc)                  (blas_merged_jacobi.inl:80-85)
    EPI_CHEB = 4,     // y = c*(b*d_i) + 1*(acc*d_i), x pre-scaled by d_col (chebyshev.inl:172-178)
    EPI_CHEB_LAST = 5, // as EPI_CHEB then y = 1*(y*d_i) [+ 1*xold]  (chebyshev.inl:180-181)
    EPI_RESID_MX = 6,  // FP64 r = b - acc stored as FP32 (yo, yo2, yo3); dot in FP64
    // Horner steps after the first on the scaled operator D^{-1/2} A D^{-1/2}: x is used
    // as gathered (no column scaling), so the kernels carry no scaling code and need far
    // fewer registers (more waves in flight)
    EPI_CHEB_S = 7,      // y = c*(b*d_i) + 1*acc
    EPI_CHEB_LAST_S = 8, // as EPI_CHEB_S then y = 1*(y*d_i) [+ 1*xold]
    // first Horner step on the scaled operator (not also the last): x is the right-hand
    // side b, gathered as c_{k-1} (b_j d_j), then as EPI_CHEB_S
    EPI_CHEB_SF = 9,
    // Horner steps on the unscaled operator whose output is the next step's input already
    // multiplied by d (the scal that opens the next step, or the final one, chebyshev.inl:
    // 172-180, done per row in the epilogue): the next step gathers x as it is, with no
    // per-nonzero gather of the column scaling.  fl(t1_j * d_j) is the same product
    // whichever kernel forms it, so the iterates are bit-identical to the reference's.
    EPI_CHEB_FU = 10,    // first step: x = (c_{k-1}*(b_j*d_j))*d_j gathered from b;
                         // y = 1*((c*(b*d_i) + 1*(acc*d_i))*d_i) [+ 1*xold]
    EPI_CHEB_U = 11,     // middle steps, x pre-scaled: y as EPI_CHEB_FU
    EPI_CHEB_LAST_U = 12 // last step (its own launch class), x pre-scaled: y as EPI_CHEB_FU
};

// Epilogue families of the Chebyshev steps.
constexpr bool epi_cheb(int e) {
    return e == EPI_CHEB || e == EPI_CHEB_LAST || e == EPI_CHEB_S || e == EPI_CHEB_LAST_S ||
           e == EPI_CHEB_SF || e == EPI_CHEB_FU || e == EPI_CHEB_U || e == EPI_CHEB_LAST_U;
}
// epilogues that end with y = 1*(y*d_i) [+ 1*xold]
constexpr bool epi_cheb_last(int e) {
    return e == EPI_CHEB_LAST || e == EPI_CHEB_LAST_S || e == EPI_CHEB_FU || e == EPI_CHEB_U ||
           e == EPI_CHEB_LAST_U;
}
// unscaled-operator steps with compile-time row scaling t2 = acc*d_i
constexpr bool epi_cheb_unscaled(int e) {
    return e == EPI_CHEB_FU || e == EPI_CHEB_U || e == EPI_CHEB_LAST_U;
}
constexpr bool epi_cheb_scaled(int e) {
    return e == EPI_CHEB_S || e == EPI_CHEB_LAST_S || e == EPI_CHEB_SF;
}

// Elementwise ops on [n][nv] vectors with per-RHS coefficients (nullptr = 1).
enum Ew : int {
    EW_FILL = 0,       // z = c0
    EW_AXPBY_Z = 1,    // z = a*x + b*y                               (blas_ext.inl:30-102)
    EW_AXPBYPCZ = 2,   // z = a*x + b*y + c*z                         (blas_ext.inl:106-189)
    EW_JACOBI0 = 3,    // z = (w*d_i)*b  (zero initial guess)          (blas_merged_jacobi.inl:66-72)
    EW_CHEB0 = 4,      // z = c*(b*d_i)                               (chebyshev.inl:169-170)
    EW_CHEB0_LAST = 5, // z = 1*((c*(b*d_i))*d_i) [+1*zold]           (order-1 polynomial)
    EW_COPY3 = 6,      // z = 1*x, y2 = z, y3 = z (initial residual r0, r, p from b)
    EW_ADD_CONV = 7,   // z(T) = 1*(T)x32 + 1*z                       (ir.inl:156-158)
    EW_CONVERT = 8,    // z(T) = (T)x(other precision)                (convert.inl:31-83)
    EW_SCAL_D = 9,     // z = z * d_i  (row scaling)                  (blas.inl:173-183)
    EW_JAC_UPD = 10,   // z = (a*c)*(y*d_i) + 1*z  (check-mode Jacobi, jacobi.inl:80-84)
    EW_MX_XUPD = 11,   // z(FP64) = ao*(T)xo + bo*(T)yo + 1*z, FP32 directions/coefficients
    EW_DOTS = 12,      // no store; dots (x, w), (x, x), (w, w)       (pcg.inl:125-126, 145)
    EW_RB_R = 13,      // z = a*x + b*y; dots (w,z), (w,w), (w,w2), (z,z) (rbicgstab.inl:113-134)
};

// Scalar update applied by the last workgroup after a fused reduction.
enum Fin : int {
    FIN_NONE = 0,   // only store the reduced sums
    FIN_INIT = 1,   // rho = res0 = ||r0||^2, tolerance, initial check   (stats.h:195-206)
    FIN_ALPHA = 2,  // iter++, alpha = rho/(v,r0)                          (pbicgstab.inl:114-131)
    FIN_OMEGA = 3,  // omega, masked alpha/omega, res = theta - omega*phi, check (:146-171)
    FIN_BETA = 4,   // beta, rho update, -(omega*beta)                     (:176-199)
    FIN_NORM_INIT = 5, // res0 = ||r||^2, tolerance, check (Jacobi/MG/IR/Cheb check mode)
    FIN_NORM = 6,      // iter++ then check ||r||^2 against the tolerance
    FIN_NORM_NOINC = 7, // check ||r||^2 without counting an iteration
    // continued refinement: FP32 Krylov recurrences with FP64 residual replacement
    FIN_MX_INIT = 8,  // FP64 ||r0||^2: res0, tolerance, initial check; FP32 rho (aux)
    FIN_MX_OMEGA = 9, // as FIN_OMEGA, but a recurrence residual below the tolerance or
                      // below delta^2 * (last FP64 residual) requests an FP64 check
    FIN_MX_REPL = 10, // FP64 true residual: per-RHS convergence, masks (aux), restart ref
    FIN_MX_RST = 11,  // FP64 true residual of a restart (r0 = p = r): as FIN_MX_REPL, and
                      // rho = (r, r); halts at once when every RHS has converged
    FIN_PCG_RZ = 12,    // rho = (r, z)                                       (pcg.inl:96-99)
    FIN_PCG_ALPHA = 13, // alpha = rho / (p, Ap), masked alpha, -alpha         (pcg.inl:112-128)
    FIN_PCG_BETA = 14,  // beta = (r, z) / rho, rho = (r, z), check on (r, r)  (pcg.inl:130-152)
    FIN_RB_ALPHA = 15,  // delta = (v, r0), alpha = rho / delta, -alpha        (rbicgstab.inl:104-115)
    FIN_RB_OMEGA = 16,  // omega, residual, check, rho, beta, -omega beta      (rbicgstab.inl:136-200)
    FIN_RR = 17         // true residual at a residual-replacement checkpoint: stop the solve
                        // when a recurrence/true norm ratio leaves [1/dev, dev] (convergence.h:286-302)
};

template <typename F>
struct FinArgs {
    KrylovScalars<F> *ks = nullptr;
    ConvState *cs = nullptr;
    double *hist = nullptr; // [hist_cap][nv]
    int fin = FIN_NONE;
    int conv_check = 0;
    double abs_tol = 0, rel_tol = 0;
    F *iconv = nullptr; // mask updated on partial convergence (nullptr: no mask)
    void *aux = nullptr;         // FIN_MX_*: the FP32 KrylovScalars of the iteration
    double delta2 = 0;           // FIN_MX_OMEGA: replacement step (squared)
    double gap2 = 0;             // FIN_MX_REPL: restart when true > gap2 * recurrence residual^2
                                 // (0: only on breakdown or no progress)
    int *clear_on_fin = nullptr; // flag reset to 0 by the final op
    int *clear_on_fin2 = nullptr;
    // FIN_BETA with the dots (r, r0), (r, r), (r0, r0): ask for a restart of the
    // recurrence when |(r, r0)| < theta ||r|| ||r0|| (bi-orthogonality lost)
    double theta = 0;
    // residual replacement: RR_stepping and RR_deviation (float32_t in the reference,
    // convergence.h:61-62, stored here already rounded to float); stepping 0 is off
    double rr_stepping = 0, rr_deviation = 0;
};

struct RedArgs {
    ReduceWs ws;
    const int *halt = nullptr; // skip the launch's work once set (loop early exit)
    const int *gate = nullptr; // run only while *gate != 0 (conditional steps)
    int set_halt_if_done = 0;  // kernel sets *halt when the solver's done flag is up
    ConvState *cs = nullptr;
};

// The flags a kernel tests before any work: the loop's halt flag, the gate of a
// conditional step, the solver's done flag.  All three are loaded before the first test,
// so they cost one round trip to memory together (the previous kernel wrote them, often
// from another XCD) instead of one each; an absent flag reads a constant.
struct StartFlags {
    int halt, gate, done;
};
static __device__ const int kStartFlagConst[2] = {0, 1};

__device__ __forceinline__ StartFlags start_flags(const RedArgs &r) {
    const int *hp = r.halt != nullptr ? r.halt : &kStartFlagConst[0];
    const int *gp = r.gate != nullptr ? r.gate : &kStartFlagConst[1];
    const int *dp =
        (r.set_halt_if_done && r.cs != nullptr) ? &r.cs->done : &kStartFlagConst[0];
    return StartFlags{*hp, *gp, *dp};
}

template <typename T>
struct SpmvArgs {
    const uint32_t *row;
    const uint32_t *col;
    const int16_t *col16 = nullptr; // column - row (banded matrices), staging modes 4 and 5
    const void *val; // V-typed
    int64_t n;
    const T *x;
    T *y;
    // epilogue inputs
    const T *b = nullptr;    // RESID/JACOBI/CHEB
    const T *xin = nullptr;  // JACOBI previous iterate; CHEB_LAST previous x
    const T *d = nullptr;    // row scaling (inv_diag / inv_sqrt_diag)
    const T *dcol = nullptr; // CHEB: column pre-scaling of x
    T coef = 0;              // CHEB coefficient c_q
    int x_is_rhs = 0;        // CHEB: x is the (unscaled) RHS of the first Horner step
    T xcoef = 0;             // CHEB: leading coefficient c_{k-1} applied to that RHS
    int prescaled = 0;       // CHEB: values hold D^{-1/2} A D^{-1/2} (no column / row scaling)
    const T *wcoef = nullptr; // JACOBI per-RHS relax factors
    const T *w = nullptr;    // second operand of the fused dot(s)
    T *y2 = nullptr, *y3 = nullptr; // extra copies of y (initial residual: r, p)
    // EPI_RESID_MX: FP32 copies of the FP64 residual (y is not written)
    float *yo = nullptr, *yo2 = nullptr, *yo3 = nullptr;
    // launch configuration chosen by the autotuner (0 / -1: global defaults); cfg_mode 1
    // stages each row tile's column indices and values through LDS, 2 also loads the
    // streamed operands non-temporally, 3 double-buffers the staging with direct
    // global-to-LDS loads; 4 and 5 are 2 and 3 with 16-bit column deltas (col16); 6
    // spreads each tile's nonzeros over the lanes (long rows); 7 and 8 are 2 and 4 with
    // 16-byte staging loads, twice the staging capacity and twice the nonzero chunk
    int cfg_vpt = 0, cfg_grid = 0, cfg_remap = -1, cfg_mode = 0;
};

template <typename T>
struct EwArgs {
    int64_t n; // rows
    T *z;
    const T *x = nullptr, *y = nullptr;
    const T *a = nullptr, *b = nullptr, *c = nullptr; // per-RHS coefficients, nullptr = 1
    const T *d = nullptr;                             // per-row scaling
    T c0 = 0;                                         // scalar (FILL, CHEB0 coefficient)
    const void *xo = nullptr;                         // other-precision input (CONVERT/ADD_CONV)
    const T *w = nullptr;                             // fused dot: dot(z, w)
    T *y2 = nullptr, *y3 = nullptr;                   // COPY3 extra outputs
    const void *yo = nullptr;                         // second other-precision input
    const float *ao = nullptr, *bo = nullptr;         // FP32 coefficients (EW_MX_XUPD)
    const int *keep_if = nullptr; // AXPBY_Z / AXPBYPCZ: leave z unchanged while *keep_if != 0
    // while *copy_if != 0: AXPBY_Z also stores z into y2 and dots z with itself;
    // AXPBYPCZ sets z = x (restart of a Krylov recurrence from its residual)
    const int *copy_if = nullptr;
    // AXPBYPCZ / MX_XUPD: leave z unchanged for an RHS whose coefficients a and b are
    // both exactly 0; AXPBY_Z: whose coefficient a is exactly 0 (the masked solution
    // update of a converged RHS keeps its solution even when that RHS's directions have
    // broken down to NaN)
    int freeze_zero = 0;
    const T *w2 = nullptr; // EW_RB_R: third operand of the dots
};

struct LaunchCfg {
    int grid_spmv = 2048;
    int grid_ew = 2048;
    int grid_red = 2048;
    int xcd_remap = 1;
    int spmv_stage = -1; // SpMV staging mode for every launch: -1 as tuned, else 0..8
};

LaunchCfg &launch_cfg();

// Nonzeros of one row tile that SpMV staging modes 1, 2 and 4 (7 and 8) hold in LDS.
constexpr int kSpmvStageCap = 2048;
constexpr int kSpmvStageCapVec = 4096;

// RHS per lane of the SpMV kernel for (nv, element size); 0 selects the default
// (1 for nv=1, 2 for nv=2, 4 otherwise).  Compiled widths: nv=2 {1,2}, nv=4 {2,4},
// nv=8 {2,4,8}, nv=16 {4,8,16}.
int &spmv_vpt_choice(int nv, int es);

// V value type: double, float or __hip_bfloat16; T vector type: double or float.
template <typename V, typename T>
void launch_spmv(int epi, int ndots, int nv, const SpmvArgs<T> &a, const RedArgs &r,
                 const FinArgs<T> &f, hipStream_t s);

template <typename T>
void launch_ew(int op, int ndots, int nv, const EwArgs<T> &a, const RedArgs &r,
               const FinArgs<T> &f, hipStream_t s);

// The LDS-staged SpMV instantiations (spmv_staged.hip: cfg_mode 1; spmv_stage2.hip:
// cfg_mode 2, streamed operands loaded non-temporally; spmv_stage3.hip: cfg_mode 3,
// double-buffered staging by direct global-to-LDS loads; spmv_stage4.hip and
// spmv_stage5.hip: modes 2 and 3 reading 16-bit column deltas).
template <typename V, typename T>
void launch_spmv_staged(int epi, int ndots, int nv, const SpmvArgs<T> &a, const RedArgs &r,
                        const FinArgs<T> &f, hipStream_t s);
template <typename V, typename T>
void launch_spmv_stage2(int epi, int ndots, int nv, const SpmvArgs<T> &a, const RedArgs &r,
                        const FinArgs<T> &f, hipStream_t s);
template <typename V, typename T>
void launch_spmv_stage3(int epi, int ndots, int nv, const SpmvArgs<T> &a, const RedArgs &r,
                        const FinArgs<T> &f, hipStream_t s);
template <typename V, typename T>
void launch_spmv_stage4(int epi, int ndots, int nv, const SpmvArgs<T> &a, const RedArgs &r,
                        const FinArgs<T> &f, hipStream_t s);
template <typename V, typename T>
void launch_spmv_stage5(int epi, int ndots, int nv, const SpmvArgs<T> &a, const RedArgs &r,
                        const FinArgs<T> &f, hipStream_t s);
// Staging modes 7 and 8 (spmv_stage7.hip): modes 2 and 4 with 16-byte staging loads, twice
// the staging capacity and twice the nonzero chunk.
template <typename V, typename T>
void launch_spmv_stage7(int epi, int ndots, int nv, const SpmvArgs<T> &a, const RedArgs &r,
                        const FinArgs<T> &f, hipStream_t s);
template <typename V, typename T>
void launch_spmv_stage8(int epi, int ndots, int nv, const SpmvArgs<T> &a, const RedArgs &r,
                        const FinArgs<T> &f, hipStream_t s);
// Staging mode 6 (spmv_nzp.hip): long rows, nonzeros spread over the lanes and the row sums
// taken from LDS in column order.
template <typename V, typename T>
void launch_spmv_nzp(int epi, int ndots, int nv, const SpmvArgs<T> &a, const RedArgs &r,
                     const FinArgs<T> &f, hipStream_t s);

// Final pass of a fused reduction over G workgroup partials (nd dot products per RHS),
// applying the solver's scalar update.
template <typename T, int NV>
void launch_finish(int nd, const RedArgs &r, const FinArgs<T> &f, int G, hipStream_t s);

// Dense GEMV y = A x for the coarse-grid Direct solver (dense_matvec.inl:100-117);
// `at` holds A column-major.
template <typename TA, typename T>
void launch_dense_mv(const TA *at, int64_t n, const T *x, T *y, int nv, const int *halt,
                     hipStream_t s);

// Core blocks of the block-hybrid Gauss-Seidel sweep of one matrix: the block size of
// split_range (misc.h:47-56) for nblocks blocks, the number of blocks, and the largest
// number of in-block entries of a row.  Built on the host from the CSR.
// When every row has at most kHsgsRow in-block entries, those entries are also kept
// as rs slots per row in device memory (rs = 4 when every row has at most 4, else
// kHsgsRow; ec: column - block start, ev: value in the sweep's precision, in the row's
// column order), which the LDS sweep copies per block; a row's unused slots hold value 0
// and column bsize, a row of zeros in the sweep's LDS copy of x, so x_c * a = +0 and
// sum - (+0) = sum for every sum: every slot is added, with no per-row count or branch.
// Otherwise (rs = 0, or when a block's copies exceed kHsgsLdsMax bytes) the sweep runs
// from global memory.
struct HsgsPlan {
    int64_t bsize = 0;
    int nblocks = 0;
    int64_t max_row_in = 0;
    int rs = 0;
    const uint16_t *ec = nullptr;
    const void *ev = nullptr;
};
constexpr int kHsgsRow = 8;
constexpr size_t kHsgsLdsMax = 64 * 1024;
// rows ahead of the relaxation at which the LDS sweep loads a lane's t values (even)
constexpr int kHsgsTRing = 16;
HsgsPlan hsgs_plan(const uint32_t *hrow, const uint32_t *hcol, int64_t n, int nblocks);
// Host copies of the row slots of plan p (values from hval, rounded to T).
template <typename T>
void hsgs_slots(const HsgsPlan &p, const uint32_t *hrow, const uint32_t *hcol,
                const double *hval, int64_t n, std::vector<uint16_t> &ec, std::vector<T> &ev);

// Block-hybrid symmetric Gauss-Seidel sweep (csr_gs.inl:32-134 + blas2.h:225-263) in two
// launches: t = b - A_offd x_old (the off-block residual, t is scratch of n * nv
// elements), then forward and backward relaxation of x inside each block.
template <typename V, typename T>
void launch_hsgs_offd(const uint32_t *row, const uint32_t *col, const V *val, int64_t n, int nv,
                      const T *b, const T *x, T *t, const HsgsPlan &p, const int *halt,
                      hipStream_t s);
template <typename V, typename T>
void launch_hsgs_sweep(const uint32_t *row, const uint32_t *col, const V *val,
                       const T *inv_diag, int64_t n, int nv, T *x, const T *t, const T *relax,
                       const T *iconv, const HsgsPlan &p, const int *halt, hipStream_t s);
// true when the in-block sweep of plan p runs from LDS (row slots built, copies fit)
bool hsgs_uses_lds(const HsgsPlan &p, int nv, int es);

template <typename T>
void launch_state_init(KrylovScalars<T> *ks, ConvState *cs, int nv, int hist_cap, hipStream_t s);

} // namespace xamg

namespace xamg {
void launch_copy(const void *src, void *dst, int64_t bytes, hipStream_t s);
// a = b + sc * c over `bytes` bytes of FP64 per array
void launch_triad(void *a, const void *b, const void *c, double sc, int64_t bytes, hipStream_t s);
}
