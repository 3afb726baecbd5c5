// SpMV staging mode 6 for long rows (stencils of 27 points, Galerkin coarse operators):
// the nonzeros of a tile of rows are spread over the lanes instead of the rows.
//
// Each wavefront takes tiles of RT = 64 / NV rows (tile k * 4 + w of the workgroup's row
// range goes to wave w).  Gather phase: consecutive lanes take consecutive nonzeros of
// the tile (NV / GV lanes per nonzero, GV right-hand sides each), so the column indices
// and values are read as whole cache lines and every lane has several independent x
// gathers in flight; each lane stores its products x[c][l] * a_ij (after the Chebyshev
// column scaling, if any) into the wave's slice of LDS.  Sum phase: one lane per (row,
// right-hand side) adds its row's products from LDS in ascending column order, the order
// of csr_matvec.inl:213-231, so every result is bit-identical to the other staging modes;
// the epilogue, stores and fused dots are those of the other modes with one right-hand
// side per lane (the lane -> (row, RHS) map is theirs at one RHS per lane, so the dots are
// summed in the same order too).  A tile with more nonzeros than its LDS slice holds is
// summed by each (row, RHS) lane straight from memory.
#include "spmv_kernel.hpp"

namespace xamg {

// Keeps the compiler from moving memory accesses across this point.  A wavefront's LDS
// instructions execute in the order they are issued, so a store before the point is seen
// by any lane's load after it.
__device__ __forceinline__ void lds_order() { asm volatile("" ::: "memory"); }

// LDS products per wavefront: 28 per row on average at 64 (row, RHS) pairs per tile.
template <typename T>
constexpr int nzp_cap() {
    return 64 * 28;
}

template <typename V, typename T, int NV, int GV, int EPI, int ND>
__global__ void __launch_bounds__(kBlock) k_spmv_nzp(SpmvArgs<T> a, RedArgs r, FinArgs<T> f,
                                                     int remap) {
    const StartFlags fl = start_flags(r);
    if (fl.halt || !fl.gate)
        return;
    constexpr int RT = 64 / NV;        // rows of a wave's tile
    constexpr int RPB = kBlock / NV;   // rows of the workgroup's four tiles
    constexpr int LRG = NV / GV;       // gather lanes per nonzero
    constexpr int NS = 64 / LRG;       // nonzeros per gather round
    constexpr int CAPP = nzp_cap<T>(); // products per wave
    constexpr int CAP = CAPP / NV;     // nonzeros per wave tile
    constexpr int UG = 4;              // gather rounds issued together
    constexpr int US = 8;              // sum-phase LDS reads issued together
    __shared__ __align__(16) T prod[kBlock / 64][CAPP];
    const V *__restrict__ val = (const V *)a.val;
    const uint32_t *__restrict__ row = a.row;
    const uint32_t *__restrict__ col = a.col;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int rr = lane / NV, l = lane % NV;       // sum phase
    const int slot = lane / LRG, sub = lane % LRG; // gather phase
    const int G = gridDim.x;
    const int L = logical_block(blockIdx.x, G, remap);
    const int64_t r0 = a.n * L / G, r1 = a.n * (L + 1) / G;
    constexpr int NDX = ND > 0 ? ND : 1;
    double dacc[NDX][1];
#pragma unroll
    for (int k = 0; k < NDX; k++)
        dacc[k][0] = 0.0;
    T *pw = prod[wave];
    // the tile's first and last row pointers, loaded one tile ahead
    int64_t base = r0 + (int64_t)wave * RT;
    uint32_t jA = 0, jB = 0;
    if (base < r1) {
        jA = row[base];
        jB = row[base + RT < r1 ? base + RT : r1];
    }
    for (; base < r1; base += RPB) {
        const int64_t tend = base + RT < r1 ? base + RT : r1;
        const int64_t nb = base + RPB;
        uint32_t nA = 0, nB = 0;
        if (nb < r1) {
            nA = row[nb];
            nB = row[nb + RT < r1 ? nb + RT : r1];
        }
        const int64_t i = base + rr;
        const bool active = i < tend;
        const int64_t e = i * NV + l;
        uint32_t j0 = jA, j1 = jA;
        RowOps<T, 1> o;
        T acc[1] = {(T)0};
        if (active) {
            j0 = row[i];
            j1 = row[i + 1];
            row_ops_load<T, NV, 1, EPI, ND>(a, i, e, o);
            if constexpr (EPI == EPI_AXPY)
                acc[0] = a.y[e];
        }
        const uint32_t cnt = jB - jA;
        if (cnt <= (uint32_t)CAP) {
            for (uint32_t k0 = slot; k0 < cnt; k0 += UG * NS) {
                uint32_t c[UG];
                T v[UG];
                T xv[UG][GV];
#pragma unroll
                for (int u = 0; u < UG; u++) {
                    const uint32_t k = k0 + u * NS;
                    const uint32_t j = jA + (k < cnt ? k : 0);
                    c[u] = ld_nt(col + j);
                    v[u] = to_t<T>(ld_nt(val + j));
                }
#pragma unroll
                for (int u = 0; u < UG; u++)
                    ld_vec<T, GV>(a.x + (int64_t)c[u] * NV + sub * GV, xv[u]);
#pragma unroll
                for (int u = 0; u < UG; u++) {
                    const uint32_t k = k0 + u * NS;
                    x_transform<T, GV, EPI>(a, c[u], xv[u]);
                    if (k < cnt) {
                        T p[GV];
#pragma unroll
                        for (int g = 0; g < GV; g++)
                            p[g] = xv[u][g] * v[u];
                        st_vec<T, GV>(pw + (size_t)k * NV + sub * GV, p);
                    }
                }
            }
            // one wave's LDS accesses execute in issue order, so the products are read
            // after they are stored once the compiler keeps the order (no wait on the
            // wave's global loads and stores)
            lds_order();
            if (active) {
                const T *pr = pw + (size_t)(j0 - jA) * NV + l;
                const uint32_t len = j1 - j0;
                uint32_t q = 0;
                for (; q + US <= len; q += US) {
                    T t[US];
#pragma unroll
                    for (int u = 0; u < US; u++)
                        t[u] = pr[(size_t)(q + u) * NV];
#pragma unroll
                    for (int u = 0; u < US; u++)
                        acc[0] = acc[0] + t[u];
                }
                for (; q < len; q++)
                    acc[0] = acc[0] + pr[(size_t)q * NV];
            }
            // this tile's products are read before the next tile's are stored
            lds_order();
        } else if (active) {
            for (uint32_t j = j0; j < j1; j++) {
                const uint32_t c = col[j];
                T xv[1] = {a.x[(int64_t)c * NV + l]};
                x_transform<T, 1, EPI>(a, c, xv);
                acc[0] = acc[0] + xv[0] * to_t<T>(val[j]);
            }
        }
        if (active)
            row_finish<T, NV, 1, EPI, ND>(a, e, l, acc, o, dacc);
        jA = nA;
        jB = nB;
    }
    if constexpr (ND > 0)
        reduce_and_finish<T, NV, ND, 1>(dacc, r, f, L);
}

// Right-hand sides per gather lane: one 16-byte x load.
template <typename T, int NV>
constexpr int nzp_gv() {
    return (int)(16 / sizeof(T)) < NV ? (int)(16 / sizeof(T)) : NV;
}

template <typename V, typename T, int NV>
static void spmv_nzp_nv(int epi, int nd, const SpmvArgs<T> &a, const RedArgs &r, const FinArgs<T> &f,
                        hipStream_t s) {
    constexpr int GV = nzp_gv<T, NV>();
    LaunchCfg c = launch_cfg();
    if (a.cfg_grid > 0) {
        c.grid_spmv = a.cfg_grid;
        c.grid_red = a.cfg_grid;
    }
    if (a.cfg_remap >= 0)
        c.xcd_remap = a.cfg_remap;
    int G = nd > 0 ? c.grid_red : c.grid_spmv;
    if (nd > 0 && G > r.ws.max_blocks)
        G = r.ws.max_blocks;
    const int64_t need = (a.n * NV + kBlock - 1) / kBlock;
    if (need < G)
        G = (int)(need > 0 ? need : 1);
    dim3 g(G), b(kBlock);
#define XAMG_NZP_CASE(E, D)                                                                        \
    if (epi == E && nd == D) {                                                                     \
        XAMG_LAUNCH((k_spmv_nzp<V, T, NV, GV, E, D>), g, b, 0, s, a, r, f, c.xcd_remap);           \
        launch_finish<T, NV>(D, r, f, G, s);                                                       \
        return;                                                                                    \
    }
    XAMG_NZP_CASE(EPI_AXY, 0)
    XAMG_NZP_CASE(EPI_AXY, 1)
    XAMG_NZP_CASE(EPI_AXY, 3)
    XAMG_NZP_CASE(EPI_AXPY, 0)
    XAMG_NZP_CASE(EPI_RESID, 0)
    XAMG_NZP_CASE(EPI_RESID, 1)
    XAMG_NZP_CASE(EPI_JACOBI, 0)
    XAMG_NZP_CASE(EPI_CHEB, 0)
    XAMG_NZP_CASE(EPI_CHEB_LAST, 0)
    XAMG_NZP_CASE(EPI_CHEB_S, 0)
    XAMG_NZP_CASE(EPI_CHEB_LAST_S, 0)
    XAMG_NZP_CASE(EPI_CHEB_SF, 0)
    XAMG_NZP_CASE(EPI_CHEB_FU, 0)
    XAMG_NZP_CASE(EPI_CHEB_U, 0)
    XAMG_NZP_CASE(EPI_CHEB_LAST_U, 0)
    if constexpr (std::is_same<T, double>::value && std::is_same<V, double>::value) {
        XAMG_NZP_CASE(EPI_RESID_MX, 0)
        XAMG_NZP_CASE(EPI_RESID_MX, 1)
    }
#undef XAMG_NZP_CASE
    throw std::runtime_error("xamg: unsupported spmv epilogue/dot combination");
}

template <typename V, typename T>
void launch_spmv_nzp(int epi, int ndots, int nv, const SpmvArgs<T> &a, const RedArgs &r,
                     const FinArgs<T> &f, hipStream_t s) {
    switch (nv) {
    case 1:
        spmv_nzp_nv<V, T, 1>(epi, ndots, a, r, f, s);
        break;
    case 2:
        spmv_nzp_nv<V, T, 2>(epi, ndots, a, r, f, s);
        break;
    case 4:
        spmv_nzp_nv<V, T, 4>(epi, ndots, a, r, f, s);
        break;
    case 8:
        spmv_nzp_nv<V, T, 8>(epi, ndots, a, r, f, s);
        break;
    case 16:
        spmv_nzp_nv<V, T, 16>(epi, ndots, a, r, f, s);
        break;
    default:
        throw std::runtime_error("xamg: nv must be one of 1, 2, 4, 8, 16");
    }
    XAMG_HIP_CHECK(hipGetLastError());
}

template void launch_spmv_nzp<double, double>(int, int, int, const SpmvArgs<double> &, const RedArgs &,
                                              const FinArgs<double> &, hipStream_t);
template void launch_spmv_nzp<float, float>(int, int, int, const SpmvArgs<float> &, const RedArgs &,
                                            const FinArgs<float> &, hipStream_t);
template void launch_spmv_nzp<__hip_bfloat16, float>(int, int, int, const SpmvArgs<float> &,
                                                     const RedArgs &, const FinArgs<float> &,
                                                     hipStream_t);

} // namespace xamg
