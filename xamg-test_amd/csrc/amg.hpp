// Host classical algebraic multigrid setup.  It produces the level operators that the
// reference imports from hypre's BoomerAMG (hypre_base.inl:340-406: A_l, P_l and
// R_l = P_l^T, plus per-level spectrum estimates of D^{-1/2} A_l D^{-1/2}); the cycle
// that uses them runs on the GPU.  Everything here is plain host code: no HIP calls, so
// it can be exercised without a device.
#pragma once

#include <cstdint>
#include <limits>
#include <string>
#include <vector>

namespace xamg {

class ParamList;

namespace amg {

struct Csr {
    int64_t n = 0, ncols = 0;
    std::vector<uint32_t> row, col;
    std::vector<double> val;
    int64_t nnz() const { return (int64_t)col.size(); }
};

// Setup options; names follow the MultiGrid keys of xamg_params_details.h:99-118.
struct Options {
    double strength = 0.5;    // mg_strength_threshold
    double max_row_sum = 0.9; // mg_max_row_sum
    double trunc = 0.25;      // mg_trunc_factor
    int pmax = 4;             // mg_Pmax_elements (0: no limit)
    int max_levels = 25;      // mg_max_levels (levels including the finest)
    int coarse_size = 100;    // mg_coarse_matrix_size
    int interp = 0;           // mg_interpolation_type: 0 classical, 3 direct
    int coarsen = 6;          // mg_coarsening_type (every value runs Ruge-Stueben here)
    int agg_levels = 0;       // mg_agg_num_levels (not built: treated as 0)
};
Options options_from(const ParamList &mg);

struct Hierarchy {
    // A[0] is left empty (the caller owns the fine matrix); A[l], l >= 1, are the
    // Galerkin operators; P[l] interpolates level l+1 to level l.
    std::vector<Csr> A, P;
    std::vector<double> min_eig, max_eig; // NaN until estimated or set
    int levels() const { return (int)A.size(); }
};

// Strong dependencies (classical measure, hypre_BoomerAMGCreateS semantics): entry j of
// row i is strong if -s_i a_ij > theta * max_k(-s_i a_ik), s_i = sign(a_ii); a row whose
// |sum_k a_ik| exceeds max_row_sum * |a_ii| keeps none.  Returns one flag per nonzero.
std::vector<uint8_t> strength(const Csr &A, double theta, double max_row_sum);

// Ruge-Stueben first pass: returns +1 for C points and -1 for F points.
std::vector<int8_t> rs_coarsen(const Csr &A, const std::vector<uint8_t> &strong);

// Interpolation from the C points (classical with the sign-modified distribution, or
// direct), truncated by trunc_factor and Pmax with the row sums kept.
Csr interpolation(const Csr &A, const std::vector<uint8_t> &strong, const std::vector<int8_t> &cf,
                  const Options &o);

Csr transpose(const Csr &A);
Csr multiply(const Csr &A, const Csr &B);

// Full setup: levels until the coarse size or level limit is reached.
Hierarchy setup(const Csr &A0, const Options &o);

// Extreme eigenvalues of D^{-1/2} A D^{-1/2} by Lanczos (steps iterations) and Sturm
// bisection on the tridiagonal matrix.  Deterministic for a given matrix.
void lanczos_bounds(int64_t n, const uint32_t *row, const uint32_t *col, const double *val,
                    int steps, double *lo, double *hi);

// Host worker threads used by the setup (XAMG_HOST_THREADS, default up to 16).
int host_threads();

} // namespace amg
} // namespace xamg
