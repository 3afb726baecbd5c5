// Binary CSR system files and YAML result records (the file formats either side of the
// C API in the reference's test drivers).
//
// System file (written by examples/test/mtx_gen.cpp:116-147, read by
// src/io/detail/reader.inl:47-177): u32 nrows, u32 nnz, u32 row[nrows+1], u32 col[nnz],
// f64 val[nnz], then optionally f64 b[nrows] and f64 x[nrows].  A file without b reads
// as b = 1 and one without x as x = 0, as the reference's reader falls back.
//
// Result record (examples/test/tst_output.h:32-91, tst_solver.h:155-179): YAML blocks
// info / solver / timing / topo with flow-style lists per RHS.
#include "../../include/xamg_ext.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace xamg {

namespace {

struct File {
    FILE *f = nullptr;
    File(const char *path, const char *mode) : f(fopen(path, mode)) {
        if (!f)
            throw std::runtime_error(std::string("xamg: cannot open ") + path);
    }
    ~File() {
        if (f)
            fclose(f);
    }
    long size() {
        const long cur = ftell(f);
        fseek(f, 0, SEEK_END);
        const long s = ftell(f);
        fseek(f, cur, SEEK_SET);
        return s;
    }
    void read(void *p, size_t sz, size_t n) {
        if (n && fread(p, sz, n, f) != n)
            throw std::runtime_error("xamg: system file is truncated");
    }
    void write(const void *p, size_t sz, size_t n) {
        if (n && fwrite(p, sz, n, f) != n)
            throw std::runtime_error("xamg: write to system file failed");
    }
};

struct Header {
    uint32_t nrows = 0, nnz = 0;
    bool has_b = false, has_x = false;
};

Header read_header(File &F) {
    Header h;
    F.read(&h.nrows, 4, 1);
    F.read(&h.nnz, 4, 1);
    const long base = 8 + 4L * (h.nrows + 1) + 4L * h.nnz + 8L * h.nnz;
    const long sz = F.size();
    if (sz < base)
        throw std::runtime_error("xamg: system file is shorter than its header says");
    h.has_b = sz >= base + 8L * h.nrows;
    h.has_x = sz >= base + 16L * h.nrows;
    return h;
}

} // namespace

void read_system_info(const char *path, uint64_t *nrows, uint64_t *nnz, int *has_b, int *has_x) {
    File F(path, "rb");
    Header h = read_header(F);
    *nrows = h.nrows;
    *nnz = h.nnz;
    if (has_b)
        *has_b = h.has_b;
    if (has_x)
        *has_x = h.has_x;
}

void read_system(const char *path, int *row, int *col, double *val, double *b, double *x) {
    File F(path, "rb");
    Header h = read_header(F);
    std::vector<uint32_t> u(std::max<uint64_t>(h.nrows + 1, h.nnz));
    F.read(u.data(), 4, h.nrows + 1);
    for (uint64_t i = 0; i <= h.nrows; i++) {
        if (u[i] > h.nnz || (i > 0 && u[i] < u[i - 1]) || (i == 0 && u[0] != 0))
            throw std::runtime_error("xamg: system file has invalid row pointers");
        if (row)
            row[i] = (int)u[i];
    }
    if (u[h.nrows] != h.nnz)
        throw std::runtime_error("xamg: system file row pointers do not end at nnz");
    F.read(u.data(), 4, h.nnz);
    for (uint64_t j = 0; j < h.nnz; j++) {
        if (u[j] >= h.nrows)
            throw std::runtime_error("xamg: system file has a column index out of range");
        if (col)
            col[j] = (int)u[j];
    }
    std::vector<double> v(h.nnz);
    F.read(v.data(), 8, h.nnz);
    if (val)
        std::memcpy(val, v.data(), 8 * h.nnz);
    std::vector<double> bb(h.nrows, 1.0), xx(h.nrows, 0.0);
    if (h.has_b)
        F.read(bb.data(), 8, h.nrows);
    if (h.has_x)
        F.read(xx.data(), 8, h.nrows);
    if (b)
        std::memcpy(b, bb.data(), 8 * h.nrows);
    if (x)
        std::memcpy(x, xx.data(), 8 * h.nrows);
}

void write_system(const char *path, uint64_t nrows, const int *row, const int *col, const double *val,
                  const double *b, const double *x) {
    if (nrows >= (1ull << 32) || (uint64_t)row[nrows] >= (1ull << 32))
        throw std::runtime_error("xamg: the system file format holds 32-bit sizes");
    if (x != nullptr && b == nullptr)
        throw std::runtime_error("xamg: the system file stores x only after b");
    File F(path, "wb");
    const uint32_t n = (uint32_t)nrows, nnz = (uint32_t)row[nrows];
    F.write(&n, 4, 1);
    F.write(&nnz, 4, 1);
    std::vector<uint32_t> u(std::max<uint64_t>(nrows + 1, nnz));
    for (uint64_t i = 0; i <= nrows; i++)
        u[i] = (uint32_t)row[i];
    F.write(u.data(), 4, nrows + 1);
    for (uint64_t j = 0; j < nnz; j++)
        u[j] = (uint32_t)col[j];
    F.write(u.data(), 4, nnz);
    F.write(val, 8, nnz);
    if (b)
        F.write(b, 8, nrows);
    if (x)
        F.write(x, 8, nrows);
}

namespace {

std::string num(double v) {
    if (std::isnan(v))
        return ".nan";
    if (std::isinf(v))
        return v > 0 ? ".inf" : "-.inf";
    char buf[40];
    snprintf(buf, sizeof(buf), "%.17g", v);
    return buf;
}

template <typename T, typename F>
std::string flow(const T *v, int n, F fmt) {
    std::string s = "[";
    for (int i = 0; i < n; i++) {
        if (i)
            s += ", ";
        s += fmt(v[i]);
    }
    return s + "]";
}

} // namespace

// tst_solver.h:155-179 record of one solve.
void write_solver_record(const char *path, int64_t nrows, int64_t nnz, int nv, const double *x0_norm,
                         const double *x_norm, const double *b_norm, int iters, const int *converged,
                         const double *abs_res, const double *rel_res, const double *init_res,
                         double seconds, const char *config) {
    File F(path, "w");
    auto d = [](double v) { return num(v); };
    auto bo = [](int v) { return std::string(v ? "true" : "false"); };
    std::string s = "---\ninfo:\n";
    s += "  nrows: " + std::to_string(nrows) + "\n";
    s += "  nonzeros: " + std::to_string(nnz) + "\n";
    s += "  NV: " + std::to_string(nv) + "\n";
    s += "  nprocs: 1\n";
    if (config && *config)
        s += std::string("  config: ") + config + "\n";
    s += "solver:\n";
    if (x0_norm)
        s += "  X_initial_norm: " + flow(x0_norm, nv, d) + "\n";
    s += "  X_norm: " + flow(x_norm, nv, d) + "\n";
    s += "  Y_norm: " + flow(b_norm, nv, d) + "\n";
    s += "  iters: " + std::to_string(iters) + "\n";
    s += "  converged: " + flow(converged, nv, bo) + "\n";
    s += "  abs_residual: " + flow(abs_res, nv, d) + "\n";
    s += "  rel_residual: " + flow(rel_res, nv, d) + "\n";
    s += "  initial_residual: " + flow(init_res, nv, d) + "\n";
    s += "timing:\n  solver: " + num(seconds) + "\n";
    s += "topo:\n  nnumas: 1\n  ncores: 1\n";
    F.write(s.data(), 1, s.size());
}

} // namespace xamg
