// Shared host/device definitions for the MI355X XAMG solve path.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>
#include <hip/hip_ext.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>

namespace xamg {

constexpr int kMaxNV = 16;
constexpr int kBlock = 256;    // threads per workgroup for streaming kernels
constexpr int kNumXcd = 8;     // MI355X: 8 XCDs, workgroups dispatched round-robin
constexpr int kMaxDots = 4;    // fused reductions per kernel (RBiCGStab: theta, phi, psi, eta)

struct HipError : std::runtime_error {
    explicit HipError(const std::string &s) : std::runtime_error(s) {}
};

#define XAMG_HIP_CHECK(expr)                                                                      \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess) {                                                                    \
            char b_[512];                                                                          \
            snprintf(b_, sizeof(b_), "HIP error %s at %s:%d: %s", hipGetErrorName(e_), __FILE__,   \
                     __LINE__, #expr);                                                             \
            throw ::xamg::HipError(b_);                                                            \
        }                                                                                          \
    } while (0)

#define XAMG_REQUIRE(cond, msg)                                                                    \
    do {                                                                                           \
        if (!(cond))                                                                               \
            throw std::runtime_error(std::string("xamg: ") + (msg));                              \
    } while (0)

// F32S: FP32 values of D^{-1/2} A D^{-1/2} (the symmetrically scaled operator whose
// polynomial the Chebyshev smoother evaluates), formed in FP64 and rounded once.
enum class Prec : int { F64 = 0, F32 = 1, BF16 = 2, F32S = 3 };

inline size_t prec_size(Prec p) { return p == Prec::F64 ? 8 : (p == Prec::BF16 ? 2 : 4); }

// Device-resident convergence record of one solver instance (convergence.h:49-303).
// Lives in HBM; kernels update it, the host reads it back only at batch ends.
struct ConvState {
    double res0[kMaxNV];
    double res[kMaxNV];
    double tol[kMaxNV];
    int status[kMaxNV];
    int iter;      // iterations counted like stats.increment_iters_counter
    int sub_iter;  // inner iterations (IterativeRefinement)
    int done;      // all RHS converged at the last check
    int halt;      // loop kernels return immediately once set
    int nhist;     // residual-history rows written
    int hist_cap;
    // continued refinement (FP32 Krylov, FP64 residual replacement)
    int replace;            // set when the FP64 true residual must be recomputed
    int replaced;           // set by that recomputation for the following r update
    int restart;            // FP32 recurrence broke down: restart it from the FP64 residual
    int rstnow;             // (r, r0) lost: restart at the end of this iteration
    double maxres[kMaxNV];  // squared residual at the last FP64 check
    // residual replacement (convergence_details RR, convergence.h:269-302)
    double rr_tol;   // next checkpoint of the largest relative recurrence residual
    int rr_pending;  // set by a check that reached rr_tol: compute the true residual
    int rr_checks;   // true residuals computed in this solve
    int rr_stopped;  // a recurrence/true deviation ended this solve
    int rr_pad;
};

// Krylov scalars of one BiCGStab/PBiCGStab instance, precision of the iteration.
template <typename F>
struct KrylovScalars {
    F rho[kMaxNV], alpha[kMaxNV], nalpha[kMaxNV], omega[kMaxNV], nomega[kMaxNV];
    F beta[kMaxNV], nbo[kMaxNV], ac[kMaxNV], oc[kMaxNV], iconv[kMaxNV];
    F s1[kMaxNV];
};

// Reduction workspace shared by every fused-dot kernel on a stream: per-block
// partial sums (pair-major) plus a ticket counter for the last-block final pass.
struct ReduceWs {
    double *partials;  // [kMaxDots*kMaxNV][max_blocks]
    unsigned *counter; // zero between kernels (the last block resets it)
    double *result;    // [kMaxDots*kMaxNV], written by the last block
    int max_blocks;
};

// Events of the library's per-launch timer: while `stop` is set, the launches below carry
// HIP events that their dispatches stamp at the kernel's start (the first launch) and
// end (every launch re-stamps `stop`, so it ends up at the last kernel's end), so a class
// is timed from its first kernel's start to its last kernel's end, without the gaps
// between dispatches that events recorded on the stream around them would include.
struct LaunchEvents {
    hipEvent_t start = nullptr, stop = nullptr;
};
LaunchEvents &launch_events();

} // namespace xamg

// hipLaunchKernelGGL, or with the timer's events when it is armed
#define XAMG_LAUNCH(K, G, B, SH, S, ...)                                                           \
    do {                                                                                           \
        ::xamg::LaunchEvents &le_ = ::xamg::launch_events();                                       \
        if (le_.stop != nullptr) {                                                                 \
            hipExtLaunchKernelGGL(K, G, B, SH, S, le_.start, le_.stop, 0, __VA_ARGS__);            \
            le_.start = nullptr;                                                                   \
        } else {                                                                                   \
            hipLaunchKernelGGL(K, G, B, SH, S, __VA_ARGS__);                                       \
        }                                                                                          \
    } while (0)
