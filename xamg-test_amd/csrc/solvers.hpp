// Solver hierarchy of the MI355X XAMG path (src/solvers rebuilt: the host enqueues
// HIP kernels, every scalar of the iteration stays in HBM, and the loops exit on a
// device-side halt flag that the host polls once per batch of iterations).
#pragma once

#include "amg.hpp"
#include "params.hpp"
#include "runtime.hpp"

#include <functional>
#include <memory>
#include <string>
#include <vector>

namespace xamg {

struct SolveStats {
    int iters = 0, sub_iters = 0;
    int rr_checks = 0, rr_stopped = 0; // residual replacement: true residuals, stops
    int converged[kMaxNV] = {0};
    double res0[kMaxNV] = {0}, res[kMaxNV] = {0};
    std::vector<double> hist; // [nhist][nv]
    int nhist = 0;
    double seconds = 0;
};

// MG hierarchy levels (hypre_base.inl:340-406 layout: A_l, P_l, R_l = P_l^T).
struct MgLevel {
    DevMatrix *A = nullptr; // level matrix (level 0 is the user's matrix)
    std::unique_ptr<DevMatrix> Aown, P, R;
};

struct MgHierarchy {
    std::vector<MgLevel> levels;
};

class SolverBase {
  public:
    virtual ~SolverBase() {}
    virtual void stats(SolveStats &st, bool with_hist) = 0;
    virtual void renew(const GlobalParams &gp) = 0;
};

template <typename T>
class Solver : public SolverBase {
  public:
    DevMatrix *A = nullptr;
    ParamList prm;
    std::string role;
    int nv = 1;
    Prec vprec;        // matrix value precision used by this solver's SpMVs
    int level = 0;     // MG level this solver works on
    const GlobalParams *gp = nullptr;

    int conv_check = 0;
    double abs_tol = 0, rel_tol = 0;
    int max_iters = 1;
    // residual replacement (convergence_details RR): on for this solver, whose kind
    // computes true residuals at the checkpoints (the Krylov solvers: elsewhere the
    // checked residual is the true one), and the residual_control keys
    int rr = 0;
    bool rr_capable = false;
    double rr_step = 0.1, rr_dev = 2.0;
    int hist_cap = 0;
    DBuf state; // ConvState + KrylovScalars<T>
    DBuf histbuf;
    ConvState *cs = nullptr;
    KrylovScalars<T> *ks = nullptr;

    virtual ~Solver() {}
    virtual void setup();
    // x <- M^{-1} b (or the solve); xzero: caller guarantees x is zero (flag semantics of
    // the reference's set_const(x, 0)); halt: enclosing loop's early-exit flag.
    virtual void apply(T *x, bool xzero, const T *b, const int *halt) = 0;
    void stats(SolveStats &st, bool with_hist) override;
    void renew(const GlobalParams &gp) override;
    void read_conf();
    T *vec(int k) { return bufs[k].template as<T>(); }
    std::vector<DBuf> bufs;
    void alloc_bufs(int k);
    FinArgs<T> fin(int f, bool check_mask = false);
};

template <typename T>
std::unique_ptr<Solver<T>> make_solver(const std::string &method, DevMatrix *A,
                                       const ParamList &prm, const std::string &role, int nv,
                                       const GlobalParams *gp, int level,
                                       std::shared_ptr<MgHierarchy> hier = nullptr);

// Top-level construction from the parameter dictionary (solver.h:449-471).
std::unique_ptr<SolverBase> construct_solver_hierarchy(const GlobalParams &gp, DevMatrix *A, int nv,
                                                       Solver<double> **top64);

// Host loop with device early exit: runs body() up to max_iters times in batches,
// stopping once the device sets cs->halt.  Returns the number of batches issued.
int run_loop(int max_iters, ConvState *cs, const std::function<void()> &body);

// Host classical AMG setup (Ruge-Stueben coarsening, direct interpolation, Galerkin
// triple product) standing in for hypre's BoomerAMG setup.
std::shared_ptr<MgHierarchy> build_amg(DevMatrix *A, const ParamList &mgprm);
// Device copy of a host hierarchy whose level 0 is A0.
std::shared_ptr<MgHierarchy> upload_hierarchy(DevMatrix *A0, const amg::Hierarchy &h);

// Options of the execution engine (batch size of the early-exit polling).
struct EngineCfg {
    int batch = 8;
    int use_graphs = 0;
};
EngineCfg &engine_cfg();

} // namespace xamg
