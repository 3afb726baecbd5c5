// Golden-vector driver around the reference's own SpMV/BLAS layer -- TEST
// INFRASTRUCTURE ONLY, built by oracle/ref.mk into oracle/_ref/ from the headers under
// the reference checkout (never copied).  It runs as one MPI rank with one core and
// writes the outputs of csr_Ax_y / csr_Axpy (csr_matvec.inl), set_rand (blas.inl:389-429),
// dot (blas.inl:521-589), axpby (blas.inl:89-151), axpby_z/axpbypcz (blas_ext.inl) and
// bxor_global (blas.inl:431-484) for a CSR system read from a binary input file.
//
// input : int64 n, nnz, nv; u32 row[n+1]; u32 col[nnz]; f64 val[nnz]; f64 x[n*nv]
// output: f64 ax[n*nv]; f64 y0[n*nv] (set_rand); f64 axpy[n*nv] (y0 + A x);
//         f64 dot_xy[nv]; f64 dot_xx[nv]; f64 axpby[n*nv] (0.5 x + (-2) y0);
//         f64 axpbypcz[n*nv] (0.5 x + (-2) y0 + 3 ax); u64 bxor_ax[nv]
#include <xamg/xamg_headers.h>
#include <xamg/xamg_types.h>
#include <xamg/init.h>
#include <xamg/primitives/vector/vector.h>
#include <xamg/primitives/matrix/matrix.h>
#include <xamg/blas/blas.h>
#include <xamg/blas2/blas2.h>

#include <cstdio>
#include <vector>

extern ID id;

using matrix_t = XAMG::matrix::csr_matrix<float64_t, uint32_t, uint32_t, uint32_t, uint32_t>;

static void set_flags(XAMG::vector::vector &v) {
    v.if_initialized = true;
    v.if_zero = false;
}

template <uint16_t NV>
static int run(FILE *in, FILE *out, int64_t n, int64_t nnz) {
    std::vector<uint32_t> row(n + 1), col(nnz);
    std::vector<double> val(nnz), x(n * NV);
    if (fread(row.data(), 4, n + 1, in) != (size_t)(n + 1) || fread(col.data(), 4, nnz, in) != (size_t)nnz ||
        fread(val.data(), 8, nnz, in) != (size_t)nnz || fread(x.data(), 8, n * NV, in) != (size_t)(n * NV))
        return 2;

    auto csr = std::make_shared<matrix_t>();
    csr->nrows = n;
    csr->block_nrows = n;
    csr->ncols = n;
    csr->block_ncols = n;
    csr->block_row_offset = 0;
    csr->block_col_offset = 0;
    csr->nonzeros = nnz;
    csr->alloc();
    std::copy(row.begin(), row.end(), csr->row.get_aligned_ptr<uint32_t>());
    std::copy(col.begin(), col.end(), csr->col.get_aligned_ptr<uint32_t>());
    std::copy(val.begin(), val.end(), csr->val.get_aligned_ptr<double>());
    set_flags(csr->row);
    set_flags(csr->col);
    set_flags(csr->val);

    XAMG::vector::vector x0, z0;
    x0.alloc<double>(n, NV);
    z0.alloc<double>(n, NV);
    std::copy(x.begin(), x.end(), x0.get_aligned_ptr<double>());
    std::fill(z0.get_aligned_ptr<double>(), z0.get_aligned_ptr<double>() + n * NV, 0.0);
    x0.ext_offset = z0.ext_offset = 0;
    set_flags(x0);
    set_flags(z0);

    auto part = XAMG::part::make_partitioner(csr->nrows);
    XAMG::vector::vector xv(XAMG::mem::DISTRIBUTED), ax(XAMG::mem::DISTRIBUTED),
        y0(XAMG::mem::DISTRIBUTED), t(XAMG::mem::DISTRIBUTED);
    XAMG::vector::construct_distributed<double, NV>(part, x0, xv);
    XAMG::vector::construct_distributed<double, NV>(part, z0, ax);
    XAMG::vector::construct_distributed<double, NV>(part, z0, y0);
    XAMG::vector::construct_distributed<double, NV>(part, z0, t);
    XAMG::matrix::matrix m(XAMG::mem::DISTRIBUTED);
    XAMG::matrix::construct_distributed<matrix_t>(part, *csr, m);

    auto dump = [&](XAMG::vector::vector &v, size_t cnt, size_t es) {
        fwrite(v.get_aligned_ptr<double>(), es, cnt, out);
    };
    XAMG::blas2::Ax_y<double, NV>(m, xv, ax, NV);
    dump(ax, n * NV, 8);
    XAMG::blas::set_rand<double, NV>(y0, false);
    dump(y0, n * NV, 8);
    XAMG::blas::copy<double, NV>(y0, t);
    XAMG::blas2::Axpy<double, NV>(m, xv, t, NV);
    dump(t, n * NV, 8);

    XAMG::vector::vector d1, d2;
    d1.alloc<double>(1, NV);
    d2.alloc<double>(1, NV);
    XAMG::blas::dot<double, NV>(xv, ax, d1);
    XAMG::blas::dot<double, NV>(xv, xv, d2);
    dump(d1, NV, 8);
    dump(d2, NV, 8);

    XAMG::blas::copy<double, NV>(y0, t);
    XAMG::blas::axpby<double, NV>(0.5, xv, -2.0, t);
    dump(t, n * NV, 8);

    XAMG::vector::vector ca, cb, cc;
    ca.alloc<double>(1, NV);
    cb.alloc<double>(1, NV);
    cc.alloc<double>(1, NV);
    XAMG::blas::set_const<double, NV>(ca, 0.5, true);
    XAMG::blas::set_const<double, NV>(cb, -2.0, true);
    XAMG::blas::set_const<double, NV>(cc, 3.0, true);
    XAMG::blas::copy<double, NV>(ax, t);
    XAMG::blas::axpbypcz<double, NV>(ca, xv, cb, y0, cc, t);
    dump(t, n * NV, 8);

    XAMG::vector::vector cs;
    cs.alloc<uint64_t>(1, NV);
    XAMG::blas::bxor_global<double, NV>(ax, cs);
    fwrite(cs.get_aligned_ptr<uint64_t>(), 8, NV, out);
    return 0;
}

int main(int argc, char **argv) {
    if (argc < 3) {
        fprintf(stderr, "usage: %s input.bin output.bin\n", argv[0]);
        return 1;
    }
    XAMG::init(argc, argv, "");
    FILE *in = fopen(argv[1], "rb");
    FILE *out = fopen(argv[2], "wb");
    if (!in || !out)
        return 1;
    int64_t hdr[3];
    if (fread(hdr, 8, 3, in) != 3)
        return 2;
    int rc = 3;
    switch (hdr[2]) {
    case 1:
        rc = run<1>(in, out, hdr[0], hdr[1]);
        break;
    case 2:
        rc = run<2>(in, out, hdr[0], hdr[1]);
        break;
    case 4:
        rc = run<4>(in, out, hdr[0], hdr[1]);
        break;
    case 8:
        rc = run<8>(in, out, hdr[0], hdr[1]);
        break;
    case 16:
        rc = run<16>(in, out, hdr[0], hdr[1]);
        break;
    }
    fclose(in);
    fclose(out);
    XAMG::finalize();
    return rc;
}
