// XAMG C API (include/xamg_c_api.h) and its MI355X extensions (include/xamg_ext.h).
// Handle tables follow api/c/xamg_c_api.cpp:64-85: indices into process-wide vectors,
// never reused; *_destroy empties a slot.
#include "../../include/xamg_c_api.h"
#include "../../include/xamg_ext.h"

#include "params.hpp"
#include "runtime.hpp"
#include "batch.hpp"
#include "solvers.hpp"

#include <sys/time.h>

#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

namespace xamg {
void fill_rhs(double *b, int64_t n, int nv, uint32_t seed);
int64_t poisson7_nnz(int nx, int ny, int nz);
void poisson7(int nx, int ny, int nz, int *row, int *col, double *val);
int64_t aniso27_nnz(int nx, int ny, int nz);
void aniso27(int nx, int ny, int nz, double kx, double ky, double kz, uint32_t seed, int *row,
             int *col, double *val);
int64_t var5_nnz(int s);
void var5(int s, uint32_t seed, int *row, int *col, double *val);
void batch_var5_layout(int nsys, uint32_t seed, int smin, int smax, int maxnv, int *side, int *nv);
int64_t batch_var5_nnz(int nsys, const int *side);
void batch_var5(int nsys, const int *side, uint32_t seed, int *sys_rows, int *row, int *col,
                double *val);
void read_system_info(const char *path, uint64_t *nrows, uint64_t *nnz, int *has_b, int *has_x);
void read_system(const char *path, int *row, int *col, double *val, double *b, double *x);
void write_system(const char *path, uint64_t nrows, const int *row, const int *col, const double *val,
                  const double *b, const double *x);
void write_solver_record(const char *path, int64_t nrows, int64_t nnz, int nv, const double *x0_norm,
                         const double *x_norm, const double *b_norm, int iters, const int *converged,
                         const double *abs_res, const double *rel_res, const double *init_res,
                         double seconds, const char *config);
} // namespace xamg

using namespace xamg;

namespace {

std::string g_err;

struct Part {
    int64_t size = 0;
};

struct Vec {
    int mode = XAMG_MEM_LOCAL;
    int64_t size = 0;
    int nv = 0;
    uint64_t ext_offset = 0;
    std::vector<double> host; // LOCAL vectors
    DBuf dev;                 // SHARED / DISTRIBUTED vectors (HBM)
    bool on_device() const { return mode != XAMG_MEM_LOCAL; }
};

struct HostCsr {
    int64_t nrows = 0, ncols = 0, nnz = 0;
    int64_t row_offset = 0, col_offset = 0;
    std::vector<uint32_t> row, col;
    std::vector<double> val;
};

struct Mat {
    std::unique_ptr<DevMatrix> m;
    int part = -1;
};

struct Sol {
    std::unique_ptr<SolverBase> s;
    Solver<double> *top = nullptr;
    GlobalParams params;
    int mid = -1;
    int x = -1, b = -1;
    int nv = 0;
    double last_ms = 0;
};

std::vector<GlobalParams *> g_params;
std::vector<std::map<std::string, std::string> *> g_ovl;
std::vector<Part *> g_parts;
std::vector<Vec *> g_vecs;
std::vector<HostCsr *> g_csr;
std::vector<Mat *> g_mats;
std::vector<Sol *> g_sols;
std::vector<amg::Hierarchy *> g_hiers;
std::vector<Batch *> g_batches;

// MultiGrid keys of a parameter dictionary: the first list whose method is MultiGrid,
// with the reference defaults applied (hypre_base.h:124-215 reads the same keys).
ParamList mg_list(const GlobalParams &p) {
    GlobalParams g = p;
    g.set_defaults();
    for (auto &r : g.roles)
        if (r.second.has("method") && r.second.get_string("method") == "MultiGrid")
            return r.second;
    ParamList d;
    set_family_defaults(d, "MultiGrid", "preconditioner");
    return d;
}

void copy_csr_out(const amg::Csr &c, int *row, int *col, double *val) {
    if (row)
        for (int64_t i = 0; i <= c.n; i++)
            row[i] = (int)c.row[i];
    if (col)
        for (int64_t j = 0; j < c.nnz(); j++)
            col[j] = (int)c.col[j];
    if (val)
        std::copy(c.val.begin(), c.val.end(), val);
}

amg::Csr csr_in(int64_t n, int64_t ncols, const int *row, const int *col, const double *val) {
    amg::Csr c;
    c.n = n;
    c.ncols = ncols;
    c.row.resize(n + 1);
    XAMG_REQUIRE(row[0] == 0, "row pointers must start at 0");
    for (int64_t i = 0; i <= n; i++) {
        XAMG_REQUIRE(i == 0 || row[i] >= row[i - 1], "row pointers must be non-decreasing");
        c.row[i] = (uint32_t)row[i];
    }
    const int64_t nnz = row[n];
    c.col.resize(nnz);
    c.val.assign(val, val + nnz);
    for (int64_t j = 0; j < nnz; j++) {
        XAMG_REQUIRE(col[j] >= 0 && col[j] < ncols, "column index out of range");
        c.col[j] = (uint32_t)col[j];
    }
    return c;
}

template <typename T>
T *slot(std::vector<T *> &v, size_t id, const char *what) {
    if (id >= v.size() || v[id] == nullptr)
        throw std::runtime_error(std::string("xamg: invalid ") + what + " handle");
    return v[id];
}

template <typename F>
int guard(F &&f) {
    try {
        f();
        return 0;
    } catch (const std::exception &e) {
        g_err = e.what();
        fprintf(stderr, "%s\n", e.what());
        return -1;
    }
}

int default_nv() {
    const char *e = getenv("XAMG_NV");
    return e ? atoi(e) : 1;
}

bool print_enabled() {
    const char *e = getenv("XAMG_LOG");
    if (!e)
        return true;
    std::string s(e);
    return !(s.find("^CONVERGENCE") != std::string::npos || s == "NONE" || s == "none" ||
             s == "quiet");
}

void check_nv(int nv) {
    XAMG_REQUIRE(nv == 1 || nv == 2 || nv == 4 || nv == 8 || nv == 16,
                 "nv must be one of 1, 2, 4, 8, 16");
}

// print.inl:119-142 residual table, from the device history after the solve.
void print_history(Sol *S, const SolveStats &st) {
    const std::string role = S->params.has("meta_solver") ? "meta_solver" : "solver";
    const std::string prefix = role == "meta_solver" ? "META" : "SOLV";
    const ParamList &p = S->params.get(role);
    if (!p.has("convergence_details") ||
        p.get_string("convergence_details").find("print") == std::string::npos ||
        p.get_string("convergence_details").find("noprint") != std::string::npos)
        return;
    if (!print_enabled())
        return;
    const int nv = S->nv;
    for (int k = 0; k < st.nhist; k++) {
        printf("%5s | %6d | ", prefix.c_str(), k);
        for (int l = 0; l < nv; l++) {
            double r = st.hist[(size_t)k * nv + l];
            if (r < 0)
                printf("%12s | %12s | ", "------------", "------------");
            else
                printf("%12.6e | %12.6e | ", std::sqrt(r), std::sqrt(r / st.res0[l]));
        }
        printf("\n");
    }
    fflush(stdout);
}

void do_solve(Sol *S, Vec *x, Vec *b, int nv) {
    check_nv(nv);
    XAMG_REQUIRE(nv == S->nv, "solve nv differs from the nv the solver was created with");
    XAMG_REQUIRE(x->on_device() && b->on_device(), "solver vectors must be distributed");
    XAMG_REQUIRE(x->nv == nv && b->nv == nv, "vector nv mismatch");
    Runtime &rt = Runtime::get();
    hipEvent_t e0, e1;
    XAMG_HIP_CHECK(hipEventCreate(&e0));
    XAMG_HIP_CHECK(hipEventCreate(&e1));
    XAMG_HIP_CHECK(hipEventRecord(e0, rt.stream));
    S->top->apply(x->dev.as<double>(), false, b->dev.as<double>(), nullptr);
    XAMG_HIP_CHECK(hipEventRecord(e1, rt.stream));
    XAMG_HIP_CHECK(hipEventSynchronize(e1));
    float ms = 0;
    XAMG_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
    S->last_ms = ms;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    SolveStats st;
    S->top->stats(st, true);
    print_history(S, st);
}

// Scratch of xamgx_dot_noalloc: the per-workgroup partial sums of one dot-product launch
// ([kMaxNV][kNoAllocBlocks] doubles) and the reduced result (kMaxNV doubles).  A
// file-scope device array, so the entry point allocates nothing; each GPU holds its own
// copy (256.1 KiB) until the process ends.  Every slot the finish pass reads is written
// by the same call, so nothing in it has to start at zero.  Calls on one GPU share it:
// calls that overlap in time (from several host threads) race.
constexpr int kNoAllocBlocks = 2048;
constexpr size_t kNoAllocWords = (size_t)kMaxNV * kNoAllocBlocks + kMaxNV;
static __device__ __align__(256) double scratch[kNoAllocWords];

// Device address of `scratch` per device ordinal.  In host code the array's name is a
// host-side stand-in; kernels get this address instead.
constexpr int kMaxDevices = 64;
double *g_scratch_dev[kMaxDevices] = {};

// Address of the current device's copy, resolved once per device (by XAMG_init, or by
// the first no-allocation call on a device).
double *noalloc_scratch() {
    int dev = 0;
    XAMG_HIP_CHECK(hipGetDevice(&dev));
    XAMG_REQUIRE(dev >= 0 && dev < kMaxDevices, "device ordinal out of range");
    if (g_scratch_dev[dev] == nullptr) {
        void *p = nullptr;
        XAMG_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        XAMG_REQUIRE(p != nullptr, "no device address for the scratch array");
        g_scratch_dev[dev] = (double *)p;
    }
    return g_scratch_dev[dev];
}

} // namespace

extern "C" {

int XAMG_init(int argc, char **argv, char *conf) {
    (void)argc;
    (void)argv;
    (void)conf;
    return guard([&] {
        Runtime::get();
        noalloc_scratch();
    });
}

int XAMG_finalize() {
    return guard([&] {
        if (Runtime::get().ready)
            Runtime::get().sync();
        for (auto &s : g_sols) {
            delete s;
            s = nullptr;
        }
        for (auto &p : g_params) {
            delete p;
            p = nullptr;
        }
        for (auto &o : g_ovl) {
            delete o;
            o = nullptr;
        }
        for (auto &p : g_parts) {
            delete p;
            p = nullptr;
        }
        for (auto &v : g_vecs) {
            delete v;
            v = nullptr;
        }
        for (auto &c : g_csr) {
            delete c;
            c = nullptr;
        }
        for (auto &m : g_mats) {
            delete m;
            m = nullptr;
        }
        for (auto &h : g_hiers) {
            delete h;
            h = nullptr;
        }
        for (auto &b : g_batches) {
            delete b;
            b = nullptr;
        }
    });
}

int XAMG_get_nv() { return default_nv(); }
int XAMG_id_get_proc() { return 0; }
int XAMG_id_get_nprocs() { return 1; }
void *XAMG_id_get_comm_ptr() { return nullptr; }

// ---- parameters (xamg_c_api.cpp:150-300) ----
int XAMG_param_create(XAMG_param_id *pid) {
    return guard([&] {
        g_params.push_back(new GlobalParams());
        *pid = g_params.size() - 1;
    });
}

static std::string num_str(double v) {
    char b[64];
    snprintf(b, sizeof(b), "%.17g", v);
    return b;
}

static int add_value(char *role, char *key, const std::string &v, XAMG_param_id pid) {
    return guard([&] { slot(g_params, pid, "param")->get(role).set(key, v); });
}
static int change_value(char *role, char *key, const std::string &v, XAMG_param_id pid, bool forced) {
    return guard([&] {
        GlobalParams *p = slot(g_params, pid, "param");
        XAMG_REQUIRE(p->has(role), std::string("no parameter block for role ") + role);
        ParamList &l = p->get(role);
        // change_value requires an existing key; forced_change_value creates it
        XAMG_REQUIRE(forced || l.has(key),
                     std::string("parameter '") + key + "' is not set for role " + role);
        l.set(key, v);
    });
}

int XAMG_param_add_value_s(char *r, char *k, char *v, XAMG_param_id p) { return add_value(r, k, v, p); }
int XAMG_param_add_value_i(char *r, char *k, uint32_t v, XAMG_param_id p) {
    return add_value(r, k, std::to_string(v), p);
}
int XAMG_param_add_value_d(char *r, char *k, double v, XAMG_param_id p) { return add_value(r, k, num_str(v), p); }
int XAMG_param_add_value_b(char *r, char *k, uint8_t v, XAMG_param_id p) {
    return add_value(r, k, v ? "1" : "0", p);
}
int XAMG_param_change_value_s(char *r, char *k, char *v, XAMG_param_id p) { return change_value(r, k, v, p, false); }
int XAMG_param_change_value_i(char *r, char *k, uint32_t v, XAMG_param_id p) {
    return change_value(r, k, std::to_string(v), p, false);
}
int XAMG_param_change_value_d(char *r, char *k, double v, XAMG_param_id p) {
    return change_value(r, k, num_str(v), p, false);
}
int XAMG_param_change_value_b(char *r, char *k, uint8_t v, XAMG_param_id p) {
    return change_value(r, k, v ? "1" : "0", p, false);
}
int XAMG_param_forced_change_value_s(char *r, char *k, char *v, XAMG_param_id p) {
    return change_value(r, k, v, p, true);
}
int XAMG_param_forced_change_value_i(char *r, char *k, uint32_t v, XAMG_param_id p) {
    return change_value(r, k, std::to_string(v), p, true);
}
int XAMG_param_forced_change_value_d(char *r, char *k, double v, XAMG_param_id p) {
    return change_value(r, k, num_str(v), p, true);
}
int XAMG_param_forced_change_value_b(char *r, char *k, uint8_t v, XAMG_param_id p) {
    return change_value(r, k, v ? "1" : "0", p, true);
}

int XAMG_param_override_create(XAMG_param_override_id *ov) {
    return guard([&] {
        g_ovl.push_back(new std::map<std::string, std::string>());
        *ov = g_ovl.size() - 1;
    });
}
int XAMG_param_override_add_value(char *key, char *val, XAMG_param_override_id ov) {
    return guard([&] { slot(g_ovl, ov, "override")->insert({key, val}); });
}
int XAMG_param_override_apply(char *role, XAMG_param_id pid, XAMG_param_override_id ov) {
    return guard([&] {
        GlobalParams *p = slot(g_params, pid, "param");
        for (auto &kv : *slot(g_ovl, ov, "override"))
            p->overrides[role].push_back(parse_override(kv.first, kv.second));
    });
}
int XAMG_param_override_destroy(XAMG_param_override_id ov) {
    return guard([&] {
        if (ov < g_ovl.size()) {
            delete g_ovl[ov];
            g_ovl[ov] = nullptr;
        }
    });
}
void XAMG_param_set_default(XAMG_param_id pid) {
    guard([&] { slot(g_params, pid, "param")->set_defaults(); });
}
int XAMG_param_print(XAMG_param_id pid) {
    return guard([&] {
        if (print_enabled())
            slot(g_params, pid, "param")->print(stdout);
    });
}
int XAMG_param_destroy(XAMG_param_id pid) {
    return guard([&] {
        if (pid < g_params.size()) {
            delete g_params[pid];
            g_params[pid] = nullptr;
        }
    });
}

// ---- partition ----
int XAMG_part_create(XAMG_part_id *pid) {
    return guard([&] {
        g_parts.push_back(new Part());
        *pid = g_parts.size() - 1;
    });
}
int XAMG_part_construct(int size, XAMG_part_id pid) {
    return guard([&] { slot(g_parts, pid, "part")->size = size; });
}
int XAMG_part_get_numa_size(XAMG_part_id pid) {
    int r = 0;
    if (guard([&] { r = (int)slot(g_parts, pid, "part")->size; }) != 0)
        return -1;
    return r;
}
int XAMG_part_destroy(XAMG_part_id pid) {
    return guard([&] {
        if (pid < g_parts.size()) {
            delete g_parts[pid];
            g_parts[pid] = nullptr;
        }
    });
}

// ---- vectors ----
int XAMG_vector_create(XAMG_vector_id *vid, vector_alloc_mode mode) {
    return guard([&] {
        Vec *v = new Vec();
        v->mode = mode;
        g_vecs.push_back(v);
        *vid = g_vecs.size() - 1;
    });
}
int XAMG_vector_offset(uint64_t off, XAMG_vector_id vid) {
    return guard([&] { slot(g_vecs, vid, "vector")->ext_offset = off; });
}
int XAMG_vector_alloc_d(int size, uint16_t nv, XAMG_vector_id vid) {
    return guard([&] {
        check_nv(nv);
        Vec *v = slot(g_vecs, vid, "vector");
        XAMG_REQUIRE(size >= 0, "negative vector size");
        v->size = size;
        v->nv = nv;
        if (v->on_device()) {
            v->dev.alloc(sizeof(double) * std::max<int64_t>(size, 1) * nv);
            XAMG_HIP_CHECK(hipMemset(v->dev.p, 0, sizeof(double) * std::max<int64_t>(size, 1) * nv));
        } else {
            v->host.assign((size_t)size * nv, 0.0);
        }
    });
}
int XAMG_vector_set_part(XAMG_part_id pid, XAMG_vector_id vid) {
    return guard([&] {
        slot(g_parts, pid, "part");
        slot(g_vecs, vid, "vector");
    });
}
int XAMG_vector_set_val_d(double val, XAMG_vector_id vid) {
    return guard([&] {
        Vec *v = slot(g_vecs, vid, "vector");
        if (v->on_device())
            fill<double>(v->dev.as<double>(), v->size, v->nv, val);
        else
            std::fill(v->host.begin(), v->host.end(), val);
        if (v->on_device())
            Runtime::get().sync();
    });
}
int XAMG_vector_download_d(double *ptr, uint64_t bs, uint64_t bo, uint16_t nv, XAMG_vector_id vid) {
    return guard([&] {
        Vec *v = slot(g_vecs, vid, "vector");
        XAMG_REQUIRE(nv == v->nv, "download nv mismatch");
        XAMG_REQUIRE(bo + bs <= (uint64_t)v->size, "download range outside the vector");
        if (v->on_device()) {
            Runtime::get().sync();
            XAMG_HIP_CHECK(hipMemcpy(ptr, v->dev.as<double>() + bo * nv, sizeof(double) * bs * nv,
                                     hipMemcpyDeviceToHost));
        } else {
            std::memcpy(ptr, v->host.data() + bo * nv, sizeof(double) * bs * nv);
        }
    });
}
int XAMG_vector_upload_v(XAMG_vector_id in, XAMG_vector_id out) {
    return guard([&] {
        Vec *a = slot(g_vecs, in, "vector"), *b = slot(g_vecs, out, "vector");
        XAMG_REQUIRE(a->nv == b->nv, "upload nv mismatch");
        XAMG_REQUIRE(a->ext_offset + a->size <= (uint64_t)b->size, "upload range outside the vector");
        size_t bytes = sizeof(double) * a->size * a->nv;
        double *dst = b->on_device() ? b->dev.as<double>() + a->ext_offset * a->nv
                                     : b->host.data() + a->ext_offset * a->nv;
        // the copies run on the null stream, which does not wait for the library's
        // non-blocking stream: finish the kernels that may still read or write either vector
        if (a->on_device() || b->on_device())
            Runtime::get().sync();
        if (a->on_device() && b->on_device())
            XAMG_HIP_CHECK(hipMemcpy(dst, a->dev.p, bytes, hipMemcpyDeviceToDevice));
        else if (!a->on_device() && b->on_device())
            XAMG_HIP_CHECK(hipMemcpy(dst, a->host.data(), bytes, hipMemcpyHostToDevice));
        else if (a->on_device())
            XAMG_HIP_CHECK(hipMemcpy(dst, a->dev.p, bytes, hipMemcpyDeviceToHost));
        else
            std::memcpy(dst, a->host.data(), bytes);
    });
}
int XAMG_vector_upload_d(double *ptr, uint64_t bs, uint64_t bo, uint16_t nv, XAMG_vector_id vid) {
    return guard([&] {
        Vec *v = slot(g_vecs, vid, "vector");
        XAMG_REQUIRE(nv == v->nv, "upload nv mismatch");
        XAMG_REQUIRE(bo + bs <= (uint64_t)v->size, "upload range outside the vector");
        if (v->on_device()) {
            Runtime::get().sync();
            XAMG_HIP_CHECK(hipMemcpy(v->dev.as<double>() + bo * nv, ptr, sizeof(double) * bs * nv,
                                     hipMemcpyHostToDevice));
        } else {
            std::memcpy(v->host.data() + bo * nv, ptr, sizeof(double) * bs * nv);
        }
    });
}
int XAMG_vector_destroy(XAMG_vector_id vid) {
    return guard([&] {
        if (vid < g_vecs.size()) {
            delete g_vecs[vid];
            g_vecs[vid] = nullptr;
        }
    });
}

// ---- host CSR ----
int XAMG_csr_matrix_create(XAMG_csr_matrix_id *id) {
    return guard([&] {
        g_csr.push_back(new HostCsr());
        *id = g_csr.size() - 1;
    });
}
int XAMG_csr_matrix_alloc(uint64_t nrows, uint64_t ncols, uint64_t nnz, XAMG_matrix_id id) {
    return guard([&] {
        HostCsr *c = slot(g_csr, id, "csr matrix");
        XAMG_REQUIRE(nnz < (1ull << 32), "nnz exceeds 32-bit row pointers");
        c->nrows = nrows;
        c->ncols = ncols;
        c->nnz = nnz;
        c->row.assign(nrows + 1, 0);
        c->col.assign(nnz, 0);
        c->val.assign(nnz, 0.0);
    });
}
int XAMG_csr_matrix_fill(int *row, int *col, double *val, XAMG_matrix_id id) {
    return guard([&] {
        HostCsr *c = slot(g_csr, id, "csr matrix");
        for (int64_t i = 0; i <= c->nrows; i++)
            c->row[i] = (uint32_t)row[i];
        for (int64_t j = 0; j < c->nnz; j++) {
            c->col[j] = (uint32_t)col[j];
            c->val[j] = val[j];
        }
    });
}
int XAMG_csr_matrix_offset(int ro, int co, XAMG_matrix_id id) {
    return guard([&] {
        HostCsr *c = slot(g_csr, id, "csr matrix");
        c->row_offset = ro;
        c->col_offset = co;
    });
}
uint32_t XAMG_csr_matrix_get_nrows(XAMG_csr_matrix_id id) {
    uint32_t r = 0;
    guard([&] { r = (uint32_t)slot(g_csr, id, "csr matrix")->nrows; });
    return r;
}
int XAMG_csr_matrix_destroy(XAMG_csr_matrix_id id) {
    return guard([&] {
        if (id < g_csr.size()) {
            delete g_csr[id];
            g_csr[id] = nullptr;
        }
    });
}

// ---- device matrix ----
int XAMG_matrix_create(XAMG_matrix_id *mid) {
    return guard([&] {
        g_mats.push_back(new Mat());
        *mid = g_mats.size() - 1;
    });
}
int XAMG_matrix_set_part(XAMG_part_id pid, XAMG_matrix_id mid) {
    return guard([&] {
        slot(g_parts, pid, "part");
        slot(g_mats, mid, "matrix")->part = (int)pid;
    });
}
int XAMG_matrix_construct(XAMG_csr_matrix_id cid, XAMG_matrix_id mid) {
    return guard([&] {
        HostCsr *c = slot(g_csr, cid, "csr matrix");
        Mat *m = slot(g_mats, mid, "matrix");
        XAMG_REQUIRE(c->row_offset == 0 && c->nrows == c->ncols,
                     "one GPU holds the whole square matrix: expected row offset 0 and nrows == ncols");
        XAMG_REQUIRE(c->nrows == 0 || c->row[c->nrows] == (uint32_t)c->nnz,
                     "row pointer end differs from nnz");
        for (int64_t j = 0; j < c->nnz; j++)
            XAMG_REQUIRE((int64_t)c->col[j] < c->ncols, "column index out of range");
        for (int64_t i = 0; i < c->nrows; i++)
            XAMG_REQUIRE(c->row[i] <= c->row[i + 1], "row pointers must be non-decreasing");
        auto dm = std::make_unique<DevMatrix>();
        dm->n = c->nrows;
        dm->ncols = c->ncols;
        dm->nnz = c->nnz;
        dm->hrow = c->row;
        dm->hcol = c->col;
        dm->hval = c->val;
        dm->upload();
        m->m = std::move(dm);
    });
}
int XAMG_matrix_destroy(XAMG_matrix_id mid) {
    return guard([&] {
        if (mid < g_mats.size()) {
            delete g_mats[mid];
            g_mats[mid] = nullptr;
        }
    });
}

// ---- solvers ----
static int create_solver(XAMG_matrix_id mid, XAMG_param_id pid, int x, int b, XAMG_solver_id *sid) {
    return guard([&] {
        Mat *m = slot(g_mats, mid, "matrix");
        XAMG_REQUIRE(m->m != nullptr, "matrix not constructed");
        GlobalParams *p = slot(g_params, pid, "param");
        Sol *S = new Sol();
        S->params = *p;
        S->mid = (int)mid;
        S->x = x;
        S->b = b;
        S->nv = x >= 0 ? slot(g_vecs, x, "vector")->nv : default_nv();
        check_nv(S->nv);
        try {
            S->s = construct_solver_hierarchy(S->params, m->m.get(), S->nv, &S->top);
        } catch (...) {
            delete S;
            throw;
        }
        g_sols.push_back(S);
        *sid = g_sols.size() - 1;
    });
}

int XAMG_solver_create1(XAMG_matrix_id mid, XAMG_param_id pid, XAMG_solver_id *sid) {
    return create_solver(mid, pid, -1, -1, sid);
}
int XAMG_solver_create2(XAMG_matrix_id mid, XAMG_param_id pid, XAMG_vector_id x, XAMG_vector_id b,
                        XAMG_solver_id *sid) {
    return create_solver(mid, pid, (int)x, (int)b, sid);
}
int XAMG_solver_renew_params(XAMG_param_id pid, XAMG_solver_id sid) {
    return guard([&] {
        Sol *S = slot(g_sols, sid, "solver");
        GlobalParams *p = slot(g_params, pid, "param");
        for (auto &r : p->roles)
            for (auto &kv : r.second.kv)
                S->params.get(r.first).set(kv.first, kv.second);
        S->s->renew(S->params);
    });
}
int XAMG_solver_solve1(uint16_t nv, XAMG_solver_id sid) {
    return guard([&] {
        Sol *S = slot(g_sols, sid, "solver");
        XAMG_REQUIRE(S->x >= 0 && S->b >= 0, "solve1 needs a solver created with vectors");
        do_solve(S, slot(g_vecs, S->x, "vector"), slot(g_vecs, S->b, "vector"), nv);
    });
}
int XAMG_solver_solve2(uint16_t nv, XAMG_vector_id x, XAMG_vector_id b, XAMG_solver_id sid) {
    return guard([&] {
        Sol *S = slot(g_sols, sid, "solver");
        do_solve(S, slot(g_vecs, x, "vector"), slot(g_vecs, b, "vector"), nv);
    });
}
int XAMG_solver_print_info(XAMG_solver_id sid) {
    return guard([&] {
        Sol *S = slot(g_sols, sid, "solver");
        DevMatrix *A = slot(g_mats, S->mid, "matrix")->m.get();
        if (print_enabled())
            printf("A: nrows %lld, nonzeros %lld, avg nnz/row %.3f\n", (long long)A->n,
                   (long long)A->nnz, A->n ? (double)A->nnz / A->n : 0.0);
    });
}
int XAMG_solver_get_convergence_info(int *conv, uint16_t nv, XAMG_solver_id sid) {
    return guard([&] {
        Sol *S = slot(g_sols, sid, "solver");
        XAMG_REQUIRE(nv <= S->nv, "nv larger than the solver's");
        SolveStats st;
        S->top->stats(st, false);
        for (int l = 0; l < nv; l++)
            conv[l] = st.converged[l];
    });
}
int XAMG_solver_get_iter_info(int *info, XAMG_solver_id sid) {
    return guard([&] {
        Sol *S = slot(g_sols, sid, "solver");
        SolveStats st;
        S->top->stats(st, false);
        info[0] = st.iters;
        info[1] = st.sub_iters;
    });
}
int XAMG_solver_destroy(XAMG_solver_id sid) {
    return guard([&] {
        if (sid < g_sols.size()) {
            delete g_sols[sid];
            g_sols[sid] = nullptr;
        }
    });
}

double XAMG_timer() {
    struct timeval tv;
    gettimeofday(&tv, nullptr);
    return (double)tv.tv_sec + 1e-6 * (double)tv.tv_usec;
}

int XAMG_barrier() {
    return guard([&] {
        if (Runtime::get().ready)
            Runtime::get().sync();
    });
}

// ---------------------------------------------------------------------------
// extensions
// ---------------------------------------------------------------------------
const char *xamgx_last_error(void) { return g_err.c_str(); }

int xamgx_device_info(char *buf, int len) {
    return guard([&] {
        Runtime &rt = Runtime::get();
        hipDeviceProp_t p;
        XAMG_HIP_CHECK(hipGetDeviceProperties(&p, rt.device));
        snprintf(buf, len, "%s | %s | CUs %d | HBM %.1f GiB", p.name, p.gcnArchName,
                 p.multiProcessorCount, p.totalGlobalMem / 1073741824.0);
    });
}

int64_t xamgx_gen_poisson7_nnz(int nx, int ny, int nz) { return poisson7_nnz(nx, ny, nz); }
int xamgx_gen_poisson7(int nx, int ny, int nz, int *row, int *col, double *val) {
    return guard([&] { poisson7(nx, ny, nz, row, col, val); });
}
int64_t xamgx_gen_aniso27_nnz(int nx, int ny, int nz) { return aniso27_nnz(nx, ny, nz); }
int xamgx_gen_aniso27(int nx, int ny, int nz, double kx, double ky, double kz, uint32_t seed,
                      int *row, int *col, double *val) {
    return guard([&] { aniso27(nx, ny, nz, kx, ky, kz, seed, row, col, val); });
}
int64_t xamgx_gen_var5_nnz(int s) { return var5_nnz(s); }
int xamgx_gen_var5(int s, uint32_t seed, int *row, int *col, double *val) {
    return guard([&] { var5(s, seed, row, col, val); });
}
int xamgx_fill_rhs(double *b, int64_t n, int nv, uint32_t seed) {
    return guard([&] { fill_rhs(b, n, nv, seed); });
}

int xamgx_spmv_host(XAMG_matrix_id mid, int prec, int mode, int nv, const void *x, void *y) {
    return guard([&] {
        check_nv(nv);
        DevMatrix *A = slot(g_mats, mid, "matrix")->m.get();
        XAMG_REQUIRE(A != nullptr, "matrix not constructed");
        XAMG_REQUIRE(prec >= 0 && prec <= 2, "prec must be 0, 1 or 2");
        XAMG_REQUIRE(mode == 0 || mode == 1, "mode must be 0 (Ax_y) or 1 (Axpy)");
        size_t es = prec == 0 ? 8 : 4;
        size_t bx = es * std::max<int64_t>(A->ncols, 1) * nv, by = es * std::max<int64_t>(A->n, 1) * nv;
        DBuf dx(bx), dy(by);
        XAMG_HIP_CHECK(hipMemcpy(dx.p, x, es * A->ncols * nv, hipMemcpyHostToDevice));
        XAMG_HIP_CHECK(hipMemcpy(dy.p, y, es * A->n * nv, hipMemcpyHostToDevice));
        const int epi = mode == 0 ? EPI_AXY : EPI_AXPY;
        if (prec == 0) {
            SpmvArgs<double> a{};
            a.x = dx.as<double>();
            a.y = dy.as<double>();
            spmv<double>(*A, Prec::F64, epi, 0, nv, a, red_args(), FinArgs<double>{});
        } else {
            SpmvArgs<float> a{};
            a.x = dx.as<float>();
            a.y = dy.as<float>();
            spmv<float>(*A, prec == 1 ? Prec::F32 : Prec::BF16, epi, 0, nv, a, red_args(),
                        FinArgs<float>{});
        }
        Runtime::get().sync();
        XAMG_HIP_CHECK(hipMemcpy(y, dy.p, es * A->n * nv, hipMemcpyDeviceToHost));
    });
}

int xamgx_dot(XAMG_vector_id x, XAMG_vector_id y, int nv, double *res) {
    return guard([&] {
        Vec *a = slot(g_vecs, x, "vector"), *b = slot(g_vecs, y, "vector");
        XAMG_REQUIRE(a->on_device() && b->on_device(), "dot needs device vectors");
        XAMG_REQUIRE(a->nv == nv && b->nv == nv && a->size == b->size, "dot shape mismatch");
        // z = 1*x (scratch copy) with the fused dot(z, y): the reduction path of the solvers
        DBuf z(sizeof(double) * std::max<int64_t>(a->size, 1) * nv);
        EwArgs<double> e;
        e.n = a->size;
        e.z = z.as<double>();
        e.x = a->dev.as<double>();
        e.y2 = z.as<double>();
        e.w = b->dev.as<double>();
        ew<double>(EW_COPY3, 1, nv, e, red_args(), FinArgs<double>{});
        Runtime::get().sync();
        XAMG_HIP_CHECK(hipMemcpy(res, Runtime::get().ws.result, sizeof(double) * nv,
                                 hipMemcpyDeviceToHost));
    });
}

int xamgx_dot_noalloc(XAMG_vector_id x, XAMG_vector_id y, int nv, double *res) {
    return guard([&] {
        // Runtime::get() would initialise (and allocate) on a first call
        XAMG_REQUIRE(Runtime::is_ready(), "xamgx_dot_noalloc: call XAMG_init first");
        Runtime &rt = Runtime::get();
        int dev = 0;
        XAMG_HIP_CHECK(hipGetDevice(&dev));
        XAMG_REQUIRE(dev == rt.device, "xamgx_dot_noalloc: the current device is not the library's");
        check_nv(nv);
        XAMG_REQUIRE(res != nullptr, "xamgx_dot_noalloc: null result pointer");
        Vec *a = slot(g_vecs, x, "vector"), *b = slot(g_vecs, y, "vector");
        XAMG_REQUIRE(a->on_device() && b->on_device(), "dot needs device vectors");
        XAMG_REQUIRE(a->nv == nv && b->nv == nv && a->size == b->size, "dot shape mismatch");
        // the partial sums of one workgroup per slot: the launch may not use more
        // workgroups than the scratch holds slots (the reduction order, and so the bits of
        // the result, follow the workgroup count, as in xamgx_dot)
        const int grid = launch_cfg().grid_red;
        if (grid > kNoAllocBlocks) {
            char m[192];
            snprintf(m, sizeof(m),
                     "xamgx_dot_noalloc: reduction grid of %d workgroups exceeds the %d partial "
                     "slots of its scratch (set grid_red <= %d)",
                     grid, kNoAllocBlocks, kNoAllocBlocks);
            throw std::runtime_error(m);
        }
        double *ws = noalloc_scratch();
        RedArgs ra;
        ra.ws.partials = ws;
        ra.ws.result = ws + (size_t)kMaxNV * kNoAllocBlocks;
        ra.ws.counter = nullptr;
        ra.ws.max_blocks = kNoAllocBlocks;
        EwArgs<double> e;
        e.n = a->size;
        e.z = nullptr; // EW_DOTS stores nothing
        e.x = a->dev.as<double>();
        e.w = b->dev.as<double>();
        // the kernels of xamgx_dot without its copy of x: the same products in the same
        // per-lane, per-workgroup and finish order
        launch_ew<double>(EW_DOTS, 1, nv, e, ra, FinArgs<double>{}, rt.stream);
        XAMG_HIP_CHECK(hipMemcpyAsync(res, ra.ws.result, sizeof(double) * nv, hipMemcpyDeviceToHost,
                                      rt.stream));
        XAMG_HIP_CHECK(hipStreamSynchronize(rt.stream));
    });
}

int xamgx_bench_spmv(XAMG_matrix_id mid, int prec, int nv, int iters, double *avg_ms) {
    return guard([&] {
        check_nv(nv);
        DevMatrix *A = slot(g_mats, mid, "matrix")->m.get();
        XAMG_REQUIRE(A != nullptr, "matrix not constructed");
        Runtime &rt = Runtime::get();
        size_t es = prec == 0 ? 8 : 4;
        DBuf dx(es * std::max<int64_t>(A->n, 1) * nv), dy(es * std::max<int64_t>(A->n, 1) * nv);
        auto launch = [&]() {
            if (prec == 0) {
                SpmvArgs<double> a{};
                a.x = dx.as<double>();
                a.y = dy.as<double>();
                spmv<double>(*A, Prec::F64, EPI_AXY, 0, nv, a, red_args(), FinArgs<double>{});
            } else {
                SpmvArgs<float> a{};
                a.x = dx.as<float>();
                a.y = dy.as<float>();
                spmv<float>(*A, prec == 1 ? Prec::F32 : Prec::BF16, EPI_AXY, 0, nv, a, red_args(),
                            FinArgs<float>{});
            }
        };
        if (prec == 0)
            fill<double>(dx.as<double>(), A->n, nv, 1.0);
        else
            fill<float>(dx.as<float>(), A->n, nv, 1.0f);
        // warm-up, until the in-solve choice of the launch configuration has settled (its
        // trial launches would otherwise fall into the timed loop)
        const Prec vp = prec == 0 ? Prec::F64 : prec == 1 ? Prec::F32 : Prec::BF16;
        for (int i = 0; i < 10 || (!A->choice_settled(vp, (int)es, nv, EPI_AXY, 0) && i < 1000);
             i++) {
            launch();
            if (i % 16 == 15)
                rt.sync();
        }
        hipEvent_t e0, e1;
        XAMG_HIP_CHECK(hipEventCreate(&e0));
        XAMG_HIP_CHECK(hipEventCreate(&e1));
        XAMG_HIP_CHECK(hipEventRecord(e0, rt.stream));
        for (int i = 0; i < iters; i++)
            launch();
        XAMG_HIP_CHECK(hipEventRecord(e1, rt.stream));
        XAMG_HIP_CHECK(hipEventSynchronize(e1));
        float ms = 0;
        XAMG_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
        *avg_ms = ms / iters;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    });
}

int xamgx_solver_last_ms(XAMG_solver_id sid, double *ms) {
    return guard([&] { *ms = slot(g_sols, sid, "solver")->last_ms; });
}

int xamgx_solver_get_history(XAMG_solver_id sid, double *out, int cap_rows, int *nrows) {
    return guard([&] {
        Sol *S = slot(g_sols, sid, "solver");
        SolveStats st;
        S->top->stats(st, true);
        int n = std::min(st.nhist, cap_rows);
        if (n > 0)
            std::memcpy(out, st.hist.data(), sizeof(double) * n * S->nv);
        *nrows = st.nhist;
    });
}

int xamgx_solver_get_residuals(XAMG_solver_id sid, double *res0, double *res) {
    return guard([&] {
        Sol *S = slot(g_sols, sid, "solver");
        SolveStats st;
        S->top->stats(st, false);
        for (int l = 0; l < S->nv; l++) {
            res0[l] = st.res0[l];
            res[l] = st.res[l];
        }
    });
}

int xamgx_solver_get_rr_info(XAMG_solver_id sid, int *checks, int *stopped) {
    return guard([&] {
        Sol *S = slot(g_sols, sid, "solver");
        SolveStats st;
        S->top->stats(st, false);
        if (checks)
            *checks = st.rr_checks;
        if (stopped)
            *stopped = st.rr_stopped;
    });
}

int xamgx_set_engine(int batch, int use_graphs, int grid_spmv, int grid_ew, int grid_red,
                     int xcd_remap) {
    return guard([&] {
        if (batch >= 0)
            engine_cfg().batch = std::max(1, batch);
        if (use_graphs >= 0)
            engine_cfg().use_graphs = use_graphs;
        LaunchCfg &c = launch_cfg();
        if (grid_spmv > 0)
            c.grid_spmv = grid_spmv;
        if (grid_ew > 0)
            c.grid_ew = grid_ew;
        if (grid_red > 0)
            c.grid_red = std::min(grid_red, 8192);
        if (xcd_remap >= 0)
            c.xcd_remap = xcd_remap;
    });
}

int xamgx_bench_spmv_epi(XAMG_matrix_id mid, int prec, int nv, int epi, int nd, int cold, int iters,
                         const int *cfg, double *avg_ms) {
    return guard([&] {
        check_nv(nv);
        DevMatrix *A = slot(g_mats, mid, "matrix")->m.get();
        XAMG_REQUIRE(A != nullptr, "matrix not constructed");
        XAMG_REQUIRE(prec >= 0 && prec <= 3, "prec must be 0 (FP64), 1 (FP32), 2 (BF16) or 3 (scaled FP32)");
        XAMG_REQUIRE(epi >= EPI_AXY && epi <= EPI_CHEB_LAST_U, "unknown epilogue");
        XAMG_REQUIRE(epi != EPI_RESID_MX || prec == 0, "the FP32-output residual needs FP64");
        XAMG_REQUIRE(iters > 0, "iters must be positive");
        const Prec vp = prec == 0 ? Prec::F64 : prec == 1 ? Prec::F32 : prec == 2 ? Prec::BF16 : Prec::F32S;
        *avg_ms = bench_spmv_epi(*A, vp, nv, epi, nd, cold, iters, cfg);
    });
}

int xamgx_kernel_timing(int on) {
    return guard([&] {
        if (on)
            KernelTimer::get().start();
        else
            KernelTimer::get().stop();
    });
}

int xamgx_kernel_stats_count(int *count) {
    return guard([&] { *count = (int)KernelTimer::get().stats().size(); });
}

int xamgx_kernel_stats_get(int i, char *name, int len, int64_t *calls, double *ms, double *bytes) {
    return guard([&] {
        const auto &st = KernelTimer::get().stats();
        XAMG_REQUIRE(i >= 0 && i < (int)st.size(), "kernel class index out of range");
        if (name != nullptr && len > 0) {
            std::strncpy(name, st[i].name.c_str(), (size_t)len - 1);
            name[len - 1] = 0;
        }
        if (calls)
            *calls = st[i].calls;
        if (ms)
            *ms = st[i].ms;
        if (bytes)
            *bytes = st[i].bytes;
    });
}

int xamgx_set_spmv_stage(int mode) {
    return guard([&] {
        XAMG_REQUIRE(mode >= -1 && mode <= 8, "mode must be -1 (tuned) or 0..8");
        launch_cfg().spmv_stage = mode;
    });
}

int xamgx_set_spmv_vpt(int nv, int es, int vpt) {
    return guard([&] {
        check_nv(nv);
        XAMG_REQUIRE(es == 4 || es == 8, "element size must be 4 or 8");
        XAMG_REQUIRE(vpt == 0 || (vpt <= nv && nv % vpt == 0), "vpt must divide nv");
        spmv_vpt_choice(nv, es) = vpt;
    });
}

int xamgx_matrix_set_eigs(XAMG_matrix_id mid, double lo, double hi) {
    return guard([&] {
        DevMatrix *A = slot(g_mats, mid, "matrix")->m.get();
        XAMG_REQUIRE(A != nullptr, "matrix not constructed");
        A->min_eig = lo;
        A->max_eig = hi;
    });
}
int xamgx_matrix_get_eigs(XAMG_matrix_id mid, double *lo, double *hi) {
    return guard([&] {
        DevMatrix *A = slot(g_mats, mid, "matrix")->m.get();
        XAMG_REQUIRE(A != nullptr, "matrix not constructed");
        if (std::isnan(A->min_eig) || std::isnan(A->max_eig))
            A->estimate_eigs();
        *lo = A->min_eig;
        *hi = A->max_eig;
    });
}

} // extern "C"

// Best rate of `iters` launches of a streaming kernel, each timed by HIP events its
// dispatch stamps at the kernel's start and end.
template <typename L>
static double best_rate(int iters, double bytes_moved, L &&launch) {
    Runtime &rt = Runtime::get();
    hipEvent_t e0, e1;
    XAMG_HIP_CHECK(hipEventCreate(&e0));
    XAMG_HIP_CHECK(hipEventCreate(&e1));
    double best = 0;
    launch(); // warm-up
    for (int i = 0; i < iters; i++) {
        launch_events() = LaunchEvents{e0, e1};
        launch();
        launch_events() = LaunchEvents{};
        XAMG_HIP_CHECK(hipEventSynchronize(e1));
        float ms = 0;
        XAMG_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
        best = std::max(best, bytes_moved / (ms * 1e-3) / 1e9);
    }
    (void)rt;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return best;
}

extern "C" int xamgx_bench_copy(uint64_t bytes, int iters, double *gbs) {
    return guard([&] {
        Runtime &rt = Runtime::get();
        bytes = (bytes / 16) * 16;
        DBuf a(bytes), b(bytes);
        XAMG_HIP_CHECK(hipMemset(a.p, 1, bytes));
        *gbs = best_rate(iters, 2.0 * (double)bytes,
                         [&] { launch_copy(a.p, b.p, (int64_t)bytes, rt.stream); });
    });
}

extern "C" int xamgx_bench_triad(uint64_t bytes, int iters, double *gbs) {
    return guard([&] {
        Runtime &rt = Runtime::get();
        bytes = (bytes / 16) * 16;
        DBuf a(bytes), b(bytes), c(bytes);
        fill<double>(b.as<double>(), (int64_t)(bytes / 8), 1, 1.0);
        fill<double>(c.as<double>(), (int64_t)(bytes / 8), 1, 2.0);
        *gbs = best_rate(iters, 3.0 * (double)bytes,
                         [&] { launch_triad(a.p, b.p, c.p, 3.0, (int64_t)bytes, rt.stream); });
    });
}

// ---- AMG hierarchies ----
// Level l >= 1 carries A_l; P_l (l < levels-1) maps level l+1 to level l.  The fine
// level's rows are recorded in the first interpolation operator.
extern "C" {

int xamgx_hierarchy_build(XAMG_matrix_id mid, XAMG_param_id pid, XAMG_hierarchy_id *hid) {
    return guard([&] {
        DevMatrix *A = slot(g_mats, mid, "matrix")->m.get();
        XAMG_REQUIRE(A != nullptr, "matrix not constructed");
        GlobalParams *p = slot(g_params, pid, "param");
        amg::Csr A0;
        A0.n = A->n;
        A0.ncols = A->ncols;
        A0.row = A->hrow;
        A0.col = A->hcol;
        A0.val = A->hval;
        g_hiers.push_back(new amg::Hierarchy(amg::setup(A0, amg::options_from(mg_list(*p)))));
        *hid = g_hiers.size() - 1;
    });
}
int xamgx_hierarchy_build_csr(XAMG_param_id pid, uint64_t nrows, const int *row, const int *col,
                              const double *val, XAMG_hierarchy_id *hid) {
    return guard([&] {
        GlobalParams *p = slot(g_params, pid, "param");
        amg::Csr A0 = csr_in((int64_t)nrows, (int64_t)nrows, row, col, val);
        g_hiers.push_back(new amg::Hierarchy(amg::setup(A0, amg::options_from(mg_list(*p)))));
        *hid = g_hiers.size() - 1;
    });
}
int xamgx_hierarchy_create(XAMG_hierarchy_id *hid) {
    return guard([&] {
        auto *h = new amg::Hierarchy();
        h->A.emplace_back();
        h->min_eig.push_back(std::numeric_limits<double>::quiet_NaN());
        h->max_eig.push_back(std::numeric_limits<double>::quiet_NaN());
        g_hiers.push_back(h);
        *hid = g_hiers.size() - 1;
    });
}
int xamgx_hierarchy_add_level(XAMG_hierarchy_id hid, uint64_t nfine, uint64_t ncoarse,
                              const int *prow, const int *pcol, const double *pval,
                              const int *arow, const int *acol, const double *aval) {
    return guard([&] {
        amg::Hierarchy *h = slot(g_hiers, hid, "hierarchy");
        const int l = h->levels() - 1;
        XAMG_REQUIRE(l == 0 || (int64_t)nfine == h->A[l].n,
                     "fine rows differ from the previous level's coarse rows");
        h->P.push_back(csr_in((int64_t)nfine, (int64_t)ncoarse, prow, pcol, pval));
        h->A.push_back(csr_in((int64_t)ncoarse, (int64_t)ncoarse, arow, acol, aval));
        h->min_eig.push_back(std::numeric_limits<double>::quiet_NaN());
        h->max_eig.push_back(std::numeric_limits<double>::quiet_NaN());
    });
}
int xamgx_hierarchy_num_levels(XAMG_hierarchy_id hid, int *nlevels) {
    return guard([&] { *nlevels = slot(g_hiers, hid, "hierarchy")->levels(); });
}
int xamgx_hierarchy_level_size(XAMG_hierarchy_id hid, int level, uint64_t sizes[4]) {
    return guard([&] {
        amg::Hierarchy *h = slot(g_hiers, hid, "hierarchy");
        XAMG_REQUIRE(level >= 0 && level < h->levels(), "level out of range");
        const bool hasP = level + 1 < h->levels();
        sizes[0] = hasP ? (uint64_t)h->P[level].n : (uint64_t)h->A[level].n;
        sizes[1] = level > 0 ? (uint64_t)h->A[level].nnz() : 0;
        sizes[2] = hasP ? (uint64_t)h->P[level].ncols : 0;
        sizes[3] = hasP ? (uint64_t)h->P[level].nnz() : 0;
    });
}
int xamgx_hierarchy_level_get(XAMG_hierarchy_id hid, int level, int *arow, int *acol, double *aval,
                              int *prow, int *pcol, double *pval) {
    return guard([&] {
        amg::Hierarchy *h = slot(g_hiers, hid, "hierarchy");
        XAMG_REQUIRE(level >= 0 && level < h->levels(), "level out of range");
        if (level > 0)
            copy_csr_out(h->A[level], arow, acol, aval);
        if (level + 1 < h->levels())
            copy_csr_out(h->P[level], prow, pcol, pval);
    });
}
int xamgx_hierarchy_set_eigs(XAMG_hierarchy_id hid, int level, double lo, double hi) {
    return guard([&] {
        amg::Hierarchy *h = slot(g_hiers, hid, "hierarchy");
        XAMG_REQUIRE(level >= 1 && level < h->levels(),
                     "level out of range (level 0 bounds belong to the matrix)");
        h->min_eig[level] = lo;
        h->max_eig[level] = hi;
    });
}
int xamgx_hierarchy_get_eigs(XAMG_hierarchy_id hid, int level, double *lo, double *hi) {
    return guard([&] {
        amg::Hierarchy *h = slot(g_hiers, hid, "hierarchy");
        XAMG_REQUIRE(level >= 1 && level < h->levels(),
                     "level out of range (level 0 bounds belong to the matrix)");
        if (std::isnan(h->min_eig[level]) || std::isnan(h->max_eig[level])) {
            const amg::Csr &A = h->A[level];
            amg::lanczos_bounds(A.n, A.row.data(), A.col.data(), A.val.data(), 60,
                                &h->min_eig[level], &h->max_eig[level]);
        }
        *lo = h->min_eig[level];
        *hi = h->max_eig[level];
    });
}
int xamgx_matrix_set_hierarchy(XAMG_matrix_id mid, XAMG_hierarchy_id hid) {
    return guard([&] {
        DevMatrix *A = slot(g_mats, mid, "matrix")->m.get();
        XAMG_REQUIRE(A != nullptr, "matrix not constructed");
        amg::Hierarchy *h = slot(g_hiers, hid, "hierarchy");
        XAMG_REQUIRE(h->levels() >= 2 && h->P[0].n == A->n,
                     "hierarchy does not match the matrix (fine rows differ)");
        A->hier = upload_hierarchy(A, *h);
    });
}
int xamgx_hierarchy_destroy(XAMG_hierarchy_id hid) {
    return guard([&] {
        if (hid < g_hiers.size()) {
            delete g_hiers[hid];
            g_hiers[hid] = nullptr;
        }
    });
}

} // extern "C"

// ---- batches of independent systems ----
extern "C" {

int xamgx_gen_batch_var5_layout(int nsys, uint32_t seed, int smin, int smax, int maxnv, int *side,
                                int *nv) {
    return guard([&] {
        XAMG_REQUIRE(nsys > 0 && smin >= 2 && smax >= smin && maxnv >= 1 && maxnv <= 4,
                     "invalid batch layout request");
        batch_var5_layout(nsys, seed, smin, smax, maxnv, side, nv);
    });
}
int64_t xamgx_gen_batch_var5_nnz(int nsys, const int *side) { return batch_var5_nnz(nsys, side); }
int xamgx_gen_batch_var5(int nsys, const int *side, uint32_t seed, int *sys_rows, int *row, int *col,
                         double *val) {
    return guard([&] { batch_var5(nsys, side, seed, sys_rows, row, col, val); });
}

int xamgx_batch_create(int nsys, const int *sys_rows, const int *nv, const int *row, const int *col,
                       const double *val, XAMG_batch_id *bid) {
    return guard([&] {
        auto *B = new Batch();
        try {
            B->build(nsys, sys_rows, nv, row, col, val);
        } catch (...) {
            delete B;
            throw;
        }
        g_batches.push_back(B);
        *bid = g_batches.size() - 1;
    });
}
int xamgx_batch_setup(XAMG_batch_id bid, XAMG_param_id pid) {
    return guard([&] {
        Batch *B = slot(g_batches, bid, "batch");
        GlobalParams *p = slot(g_params, pid, "param");
        // the previous solver (its graph, pinned slots and vectors) goes first, so two
        // solvers of one batch never coexist; a setup that fails leaves none
        B->solver.reset();
        B->solver = make_batch_solver(*B, *p);
    });
}
static double *batch_vec(Batch *B, XAMG_vector_id vid) {
    Vec *v = slot(g_vecs, vid, "vector");
    XAMG_REQUIRE(v->on_device(), "batch vectors must be device vectors");
    XAMG_REQUIRE(v->size * v->nv == B->nelem,
                 "batch vector size differs from the sum of rows x nv over the systems");
    return v->dev.as<double>();
}
int xamgx_batch_solve(XAMG_batch_id bid, XAMG_vector_id x, XAMG_vector_id b) {
    return guard([&] {
        Batch *B = slot(g_batches, bid, "batch");
        XAMG_REQUIRE(B->solver != nullptr, "batch: call xamgx_batch_setup first");
        B->solver->solve(batch_vec(B, x), batch_vec(B, b));
    });
}
int xamgx_batch_get_info(XAMG_batch_id bid, int *iters, int *checks, int *converged) {
    return guard([&] {
        Batch *B = slot(g_batches, bid, "batch");
        XAMG_REQUIRE(B->solver != nullptr, "batch: call xamgx_batch_setup first");
        B->solver->info(iters, checks, converged);
    });
}
int xamgx_batch_residuals(XAMG_batch_id bid, XAMG_vector_id x, XAMG_vector_id b, double *relres) {
    return guard([&] {
        Batch *B = slot(g_batches, bid, "batch");
        batch_residuals(*B, batch_vec(B, x), batch_vec(B, b), relres);
    });
}
int xamgx_batch_last_ms(XAMG_batch_id bid, double *ms) {
    return guard([&] { *ms = slot(g_batches, bid, "batch")->last_ms; });
}
int xamgx_batch_set_launch(int grid, int tile_rows, int batch, int use_graph) {
    return guard([&] {
        BatchCfg &c = batch_cfg();
        if (grid > 0) {
            c.grid = grid;
            c.tune = 0; // a fixed grid: one element per lane and one tile per step
            c.tps = 1;
            c.ept = 1;
        }
        if (tile_rows > 0)
            c.tile_rows = tile_rows;
        if (batch > 0)
            c.batch = batch;
        if (use_graph >= 0)
            c.use_graph = use_graph;
    });
}
int xamgx_batch_set_tuning(int on, int tps, int ept) {
    return guard([&] {
        BatchCfg &c = batch_cfg();
        XAMG_REQUIRE(tps >= 0 && (ept == 0 || ept == 1 || ept == 2),
                     "batch: tiles per step must be >= 1 and elements per lane 1 or 2");
        c.tune = on ? 1 : 0;
        if (tps > 0)
            c.tps = tps;
        if (ept > 0)
            c.ept = ept;
    });
}
int xamgx_batch_get_launch(XAMG_batch_id bid, int *grid, int *tps, int *ept, int *tuned) {
    return guard([&] {
        Batch *B = slot(g_batches, bid, "batch");
        XAMG_REQUIRE(B->solver != nullptr, "batch: call xamgx_batch_setup first");
        B->solver->launch(grid, tps, ept, tuned);
    });
}
int xamgx_batch_destroy(XAMG_batch_id bid) {
    return guard([&] {
        if (bid < g_batches.size()) {
            delete g_batches[bid];
            g_batches[bid] = nullptr;
        }
    });
}

} // extern "C"

// ---- system files and result records ----
extern "C" {

int xamgx_read_system_info(const char *path, uint64_t *nrows, uint64_t *nnz, int *has_b, int *has_x) {
    return guard([&] { read_system_info(path, nrows, nnz, has_b, has_x); });
}
int xamgx_read_system(const char *path, int *row, int *col, double *val, double *b, double *x) {
    return guard([&] { read_system(path, row, col, val, b, x); });
}
int xamgx_write_system(const char *path, uint64_t nrows, const int *row, const int *col,
                       const double *val, const double *b, const double *x) {
    return guard([&] { write_system(path, nrows, row, col, val, b, x); });
}
int xamgx_solver_write_record(XAMG_solver_id sid, XAMG_vector_id x, XAMG_vector_id b,
                              const double *x0_norm, double seconds, const char *config,
                              const char *path) {
    return guard([&] {
        Sol *S = slot(g_sols, sid, "solver");
        DevMatrix *A = slot(g_mats, S->mid, "matrix")->m.get();
        const int nv = S->nv;
        std::vector<double> xn(nv), bn(nv);
        XAMG_REQUIRE(xamgx_dot(x, x, nv, xn.data()) == 0 && xamgx_dot(b, b, nv, bn.data()) == 0,
                     g_err);
        SolveStats st;
        S->top->stats(st, false);
        std::vector<double> ar(nv), rr(nv), ir(nv);
        std::vector<int> cv(nv);
        for (int l = 0; l < nv; l++) {
            xn[l] = std::sqrt(xn[l]);
            bn[l] = std::sqrt(bn[l]);
            // stats.h:134-148: abs = sqrt(res), rel = sqrt(res / res0), initial = sqrt(res0)
            ar[l] = std::sqrt(st.res[l]);
            rr[l] = std::sqrt(st.res[l] / st.res0[l]);
            ir[l] = std::sqrt(st.res0[l]);
            cv[l] = st.converged[l];
        }
        write_solver_record(path, A->n, A->nnz, nv, x0_norm, xn.data(), bn.data(), st.iters,
                            cv.data(), ar.data(), rr.data(), ir.data(), seconds, config);
    });
}

} // extern "C"
