// Batched solves of many independent systems of uneven sizes, each with its own number
// of right-hand sides (1..4), on one GPU.  The systems are concatenated into one
// block-diagonal CSR; vectors concatenate each system's [n_k][nv_k] row-interleaved
// block.  The reference solves such a collection one system at a time, with the
// PBiCGStab / Jacobi / IterativeRefinement semantics of pbicgstab.inl, jacobi.inl and
// ir.inl; here every step of the iteration is one launch over all systems:
//  - the rows are cut into tiles that never cross a system boundary; a fixed grid of
//    workgroups strides over the tiles (persistent launch) and skips the tiles of
//    systems that have finished, so the work per iteration shrinks as systems converge;
//  - dot products are reduced per tile (fixed order) and then per system, in tile order,
//    by a finish kernel that also runs each system's scalar update and convergence test;
//  - the host enqueues iterations in batches, replays them as a captured hipGraph, and
//    polls one counter of finished systems per batch.
#pragma once

#include "params.hpp"
#include "runtime.hpp"

#include <memory>
#include <vector>

namespace xamg {

// Field indices of the per-(system, RHS) scalar arrays (structure of arrays, 4 RHS
// slots per system).
enum BScalar {
    BS_RHO = 0,
    BS_ALPHA,
    BS_NALPHA,
    BS_OMEGA,
    BS_NOMEGA,
    BS_BETA,
    BS_NBO,
    BS_AC,
    BS_OC,
    BS_ICONV,
    BS_RELAX,
    BS_NF
};
// Per-system integer flags.
enum BSys {
    BY_ITER = 0,
    BY_SUB,
    BY_DONE,
    BY_HALT,
    BY_REPLACE,
    BY_REPLACED,
    BY_RESTART,
    BY_RSTNOW,
    BY_NF
};
// Per-(system, RHS) convergence doubles.
enum BConv { BC_RES0 = 0, BC_RES, BC_TOL, BC_MAXRES, BC_NF };

constexpr int kBatchMaxNV = 4;
constexpr int kBatchDots = 3;

// Device view of the batch layout.
struct BLayout {
    int nsys = 0, ntiles = 0;
    const uint32_t *srow = nullptr; // first global row of each system [nsys+1]
    const int *nv = nullptr;        // RHS per system [nsys]
    const int64_t *voff = nullptr;  // first vector element of each system [nsys+1]
    const int *tsys = nullptr;      // tile -> system [ntiles]
    const uint32_t *trow = nullptr; // tile -> first global row [ntiles+1] (tiles are contiguous)
    const int *stile = nullptr;     // system -> first tile [nsys+1]
    const uint32_t *row = nullptr, *col = nullptr;
    int max_row_nnz = 0; // longest row of the batch (fixes the SpMV's nonzero chunk)
    int tps = 1;         // tiles a workgroup takes per step of the grid stride
    // launch shape (host side): workgroups, elements a SpMV lane keeps in flight
    int grid = 8192, ept = 1;
};

// Device view of the per-system solver state.
template <typename F>
struct BState {
    F *sc = nullptr;          // [BS_NF][nsys*4]
    double *cv = nullptr;     // [BC_NF][nsys*4]
    int *status = nullptr;    // [nsys*4]
    int *sy = nullptr;        // [BY_NF][nsys]
    int *ndone = nullptr;     // systems finished
    double *part = nullptr;   // per-tile partial sums [ntiles][3][4]
    double *hist = nullptr;   // unused (kept for layout symmetry)
};

class BatchKrylovBase;

// A batch of systems resident in HBM.
class Batch {
  public:
    int nsys = 0;
    int64_t nrows = 0, nnz = 0, nelem = 0;
    std::vector<uint32_t> h_srow, hrow, hcol;
    std::vector<int> h_nv;
    std::vector<int64_t> h_voff;
    std::vector<double> hval;
    int max_row_nnz = 0;
    DBuf srow, nv, voff, row, col, val64, val32, dinv64, dinv32;
    // tiling (rebuilt when the tile size changes)
    int tile_rows = 0, ntiles = 0;
    DBuf tsys, trow, stile;
    int gen = 0; // incremented whenever the tile arrays are reallocated
    std::vector<int> h_stile;
    std::unique_ptr<BatchKrylovBase> solver;
    double last_ms = 0;

    void build(int nsys, const int *sys_rows, const int *nvs, const int *row, const int *col,
               const double *val);
    void set_tiles(int rows);
    BLayout layout() const;
    template <typename T>
    const T *inv_diag();
    const void *values(bool fp32);
};

class BatchKrylovBase {
  public:
    virtual ~BatchKrylovBase() {}
    virtual void solve(double *x, const double *b) = 0;
    // per-system iterations / inner iterations, per-RHS convergence flags (nelem order
    // of RHS: system-major)
    virtual void info(int *iters, int *sub_iters, int *converged) = 0;
    // launch shape in use: workgroups, tiles per workgroup step, elements per lane of the
    // SpMV, and whether the autotuner chose it
    virtual void launch(int *grid, int *tps, int *ept, int *tuned) const = 0;
};

// Builds the batch solver from a parameter dictionary: "solver" BiCGStab / PBiCGStab
// (preconditioner Jacobi) in FP64, or under a "meta_solver" IterativeRefinement with
// subsolver_reduced_precision: FP32 recurrences with FP64 residual replacement.
std::unique_ptr<BatchKrylovBase> make_batch_solver(Batch &B, const GlobalParams &gp);

// FP64 relative residuals ||b - A x|| / ||b|| per RHS (system-major order).
void batch_residuals(Batch &B, const double *x, const double *b, double *relres);

// Launch configuration of the batch kernels: workgroups of the grid, rows per tile,
// iterations per host poll, hipGraph replay.  A tile is the block over which the dot
// products are summed first, so tile_rows fixes their order; the workgroups, the tiles per
// workgroup step and the elements a SpMV lane keeps in flight do not change any result
// and are timed by the autotuner at setup unless `tune` is 0 (then grid, tps and ept
// below are used).
struct BatchCfg {
    int grid = 8192;
    int tile_rows = 1024;
    int batch = 8;
    int use_graph = 1;
    int tune = 1;
    int tps = 1;
    int ept = 1;
};
BatchCfg &batch_cfg();

} // namespace xamg
