// Device code of the CSR SpMV family (kernel template and its launchers), shared by
// the translation units that instantiate its staging modes.
#pragma once

#include "kernels.hpp"

#include <type_traits>

namespace xamg {

// ---------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------

template <typename T, typename V>
__device__ __forceinline__ T to_t(V v) {
    return (T)v;
}
template <>
__device__ __forceinline__ float to_t<float, __hip_bfloat16>(__hip_bfloat16 v) {
    return __bfloat162float(v);
}
template <>
__device__ __forceinline__ double to_t<double, __hip_bfloat16>(__hip_bfloat16 v) {
    return (double)__bfloat162float(v);
}

// XCD-aware workgroup -> contiguous-range map (see kernels.hpp).
__device__ __forceinline__ int logical_block(int b, int G, int remap) {
    if (remap && (G % kNumXcd) == 0)
        return (b % kNumXcd) * (G / kNumXcd) + b / kNumXcd;
    return b;
}

// Workgroup partial sums of the fused dots: lanes, then waves, in a fixed order; the
// partial of logical block L (the L-th row range) goes to slot L, so the finish kernel
// combines the row ranges in row order whatever the XCD mapping.
template <typename T, int NV, int ND, int VPT>
__device__ void reduce_and_finish(double (&acc)[ND][VPT], const RedArgs &r, const FinArgs<T> &f,
                                  int L) {
    constexpr int LR = NV / VPT;
    __shared__ double lds[kBlock / 64][ND][NV];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int G = gridDim.x;
#pragma unroll
    for (int k = 0; k < ND; k++) {
#pragma unroll
        for (int v = 0; v < VPT; v++) {
            double s = acc[k][v];
#pragma unroll
            for (int off = 32; off >= LR; off >>= 1)
                s += __shfl_xor(s, off);
            if (lane < LR)
                lds[wave][k][lane * VPT + v] = s;
        }
    }
    __syncthreads();
    if (tid < ND * NV) {
        const int k = tid / NV, l = tid % NV;
        double s = lds[0][k][l];
#pragma unroll
        for (int w = 1; w < kBlock / 64; w++)
            s += lds[w][k][l];
        r.ws.partials[(int64_t)tid * r.ws.max_blocks + L] = s;
    }
    (void)G;
    (void)f;
}

// Non-temporal (streaming) load of an operand read once per launch, so that it does not
// displace the gathered x from L2.
template <typename T>
__device__ __forceinline__ T ld_nt(const T *p) {
    return __builtin_nontemporal_load(p);
}
template <>
__device__ __forceinline__ __hip_bfloat16 ld_nt<__hip_bfloat16>(const __hip_bfloat16 *p) {
    const unsigned short u = __builtin_nontemporal_load((const unsigned short *)p);
    __hip_bfloat16 r;
    __builtin_memcpy(&r, &u, 2);
    return r;
}

typedef unsigned int u32x4_t __attribute__((ext_vector_type(4)));
typedef unsigned int u32x2_t __attribute__((ext_vector_type(2)));

// Loads / stores of VPT consecutive vector elements (16-byte accesses where possible);
// NT: non-temporal loads.
template <typename T, int VPT, bool NT = false>
__device__ __forceinline__ void ld_vec(const T *__restrict__ p, T (&v)[VPT]) {
    if constexpr (sizeof(T) * VPT % 16 == 0) {
#pragma unroll
        for (int q = 0; q < (int)(sizeof(T) * VPT / 16); q++) {
            u32x4_t t;
            if constexpr (NT)
                t = __builtin_nontemporal_load((const u32x4_t *)p + q);
            else
                t = ((const u32x4_t *)p)[q];
            __builtin_memcpy(&v[q * 16 / sizeof(T)], &t, 16);
        }
    } else if constexpr (sizeof(T) * VPT == 8) {
        u32x2_t t;
        if constexpr (NT)
            t = __builtin_nontemporal_load((const u32x2_t *)p);
        else
            t = *(const u32x2_t *)p;
        __builtin_memcpy(&v[0], &t, 8);
    } else {
#pragma unroll
        for (int q = 0; q < VPT; q++)
            v[q] = NT ? ld_nt(p + q) : p[q];
    }
}

template <typename T, int VPT>
__device__ __forceinline__ void st_vec(T *__restrict__ p, const T (&v)[VPT]) {
    if constexpr (sizeof(T) * VPT % 16 == 0) {
#pragma unroll
        for (int q = 0; q < (int)(sizeof(T) * VPT / 16); q++) {
            uint4 t;
            __builtin_memcpy(&t, &v[q * 16 / sizeof(T)], 16);
            ((uint4 *)p)[q] = t;
        }
    } else if constexpr (sizeof(T) * VPT == 8) {
        uint2 t;
        __builtin_memcpy(&t, &v[0], 8);
        *(uint2 *)p = t;
    } else {
#pragma unroll
        for (int q = 0; q < VPT; q++)
            p[q] = v[q];
    }
}

// ---------------------------------------------------------------------------
// CSR SpMV with fused epilogue and optional fused dot products.
// LR = NV/VPT lanes per row; each lane keeps VPT right-hand sides in registers and
// reads them with 16-byte loads.  The nonzeros of a row are walked in chunks of U:
// all column indices, values and x segments of a chunk are loaded first, then added
// in ascending column order (the reference's order, csr_matvec.inl:213-231).
//
// STAGE = 1: the workgroup first copies the column indices and values of its current
// tile of RPB rows into LDS with consecutive lanes reading consecutive entries (whole
// cache lines per wavefront load), and the rows' lanes read them from LDS.  Per row the
// lanes of one row would otherwise read 4-byte entries a row length apart, so one
// wavefront load touches one cache line per few lanes.  A tile whose nonzeros exceed
// the LDS capacity is read directly, as with STAGE = 0.
//
// STAGE = 2: as STAGE = 1 with the operands that are read once per launch (column
// indices, values, row operands of the epilogue and the dots) loaded non-temporally.
// On a 7-point 128^3 grid the gathered x is reused 16384 rows later (the next plane),
// which the L2 of an XCD only holds while the streamed operands do not evict it.
//
// STAGE = 4: as STAGE = 2 reading the column indices as 16-bit offsets from the row
// (col16, kept for matrices whose entries all lie within 32767 of the diagonal): half
// the index bytes.
// ---------------------------------------------------------------------------
// Row-wise operands of an epilogue (b, the previous iterate, the second dot operand, the
// row scaling), loaded before the row's gathers so their latency overlaps them.  They are
// read once per launch and not again before other kernels have streamed more than the
// last-level cache holds, so they are loaded non-temporally: they do not displace the
// matrix and the vector the next kernel gathers from the 256 MB last-level cache (a first
// Chebyshev step's b is its gathered x and keeps ordinary loads).
template <typename T, int VPT>
struct RowOps {
    T pb[VPT], pw[VPT], px[VPT];
    T pdi;
};

template <typename T, int NV, int VPT, int EPI, int ND>
__device__ __forceinline__ void row_ops_load(const SpmvArgs<T> &a, int64_t i, int64_t e,
                                             RowOps<T, VPT> &o) {
    constexpr bool kNeedB = EPI != EPI_AXY && EPI != EPI_AXPY;
    constexpr bool NTB = EPI != EPI_CHEB_SF;
    constexpr bool kNeedD = EPI == EPI_JACOBI || epi_cheb(EPI);
    o.pdi = (T)0;
    if constexpr (kNeedB)
        ld_vec<T, VPT, NTB>(a.b + e, o.pb);
    if constexpr (kNeedD)
        o.pdi = a.d[i];
    if constexpr (EPI == EPI_JACOBI) {
        ld_vec<T, VPT, true>(a.xin + e, o.px);
    } else if constexpr (epi_cheb_last(EPI)) {
        if (a.xin != nullptr)
            ld_vec<T, VPT, true>(a.xin + e, o.px);
    }
    if constexpr (ND == 3) {
        ld_vec<T, VPT, true>(a.w + e, o.pw);
    } else if constexpr (ND == 1) {
        if (a.w != nullptr)
            ld_vec<T, VPT, true>(a.w + e, o.pw);
    }
}

// The gathered x segment of one nonzero as the row sum uses it: the Chebyshev epilogues
// scale it by the column's diagonal factor first (chebyshev.inl:169-175), in the
// reference's operation order.
template <typename T, int VPT, int EPI>
__device__ __forceinline__ void x_transform(const SpmvArgs<T> &a, uint32_t c, T (&xv)[VPT]) {
    if constexpr (EPI == EPI_CHEB_SF) {
        // first Horner step on the scaled operator: t1 = c_{k-1} (b d)
        const T dc = a.dcol[c];
#pragma unroll
        for (int v = 0; v < VPT; v++)
            xv[v] = a.xcoef * (xv[v] * dc);
    } else if constexpr (EPI == EPI_CHEB_FU) {
        // first Horner step read straight from the right-hand side:
        // t1 = c_{k-1} * (b * d), then t1 * d (chebyshev.inl:169-173)
        const T dc = a.dcol[c];
#pragma unroll
        for (int v = 0; v < VPT; v++)
            xv[v] = (a.xcoef * (xv[v] * dc)) * dc;
    } else if constexpr (EPI == EPI_CHEB || EPI == EPI_CHEB_LAST) {
        if (a.x_is_rhs && a.prescaled) {
            // first Horner step on the scaled operator: t1 = c_{k-1} (b d)
            const T dc = a.dcol[c];
#pragma unroll
            for (int v = 0; v < VPT; v++)
                xv[v] = a.xcoef * (xv[v] * dc);
        } else if (a.x_is_rhs) {
            // first Horner step read straight from the right-hand side:
            // t1 = c_{k-1} * (b * d), then t1 * d (chebyshev.inl:169-173)
            const T dc = a.dcol[c];
#pragma unroll
            for (int v = 0; v < VPT; v++)
                xv[v] = (a.xcoef * (xv[v] * dc)) * dc;
        } else if (!a.prescaled) {
            const T dc = a.dcol[c];
#pragma unroll
            for (int v = 0; v < VPT; v++)
                xv[v] = xv[v] * dc;
        }
    }
}

// Epilogue of one row's VPT right-hand sides from e = i * NV + l0 on: the row sums acc,
// the output stores and the row's terms of the fused dots.
template <typename T, int NV, int VPT, int EPI, int ND>
__device__ __forceinline__ void row_finish(const SpmvArgs<T> &a, int64_t e, int l0,
                                           const T (&acc)[VPT], const RowOps<T, VPT> &o,
                                           double (&dacc)[ND > 0 ? ND : 1][VPT]) {
    T out[VPT];
    if constexpr (EPI == EPI_AXY || EPI == EPI_AXPY) {
#pragma unroll
        for (int v = 0; v < VPT; v++)
            out[v] = acc[v];
    } else if constexpr (EPI == EPI_RESID || EPI == EPI_RESID_MX) {
#pragma unroll
        for (int v = 0; v < VPT; v++)
            out[v] = (T)1.0 * o.pb[v] + (T)(-1.0) * acc[v];
    } else if constexpr (EPI == EPI_JACOBI) {
        const T di = o.pdi;
#pragma unroll
        for (int v = 0; v < VPT; v++)
            out[v] = o.px[v] + (a.wcoef[l0 + v] * di) * (o.pb[v] - acc[v]);
    } else {
        const T di = o.pdi;
#pragma unroll
        for (int v = 0; v < VPT; v++) {
            const T rr = o.pb[v] * di;
            T t2 = acc[v];
            if constexpr (epi_cheb_unscaled(EPI))
                t2 = acc[v] * di;
            else if constexpr (!epi_cheb_scaled(EPI))
                t2 = a.prescaled ? acc[v] : acc[v] * di;
            out[v] = a.coef * rr + (T)1.0 * t2;
        }
        if constexpr (epi_cheb_last(EPI)) {
#pragma unroll
            for (int v = 0; v < VPT; v++)
                out[v] = out[v] * di;
            if (a.xin != nullptr) {
#pragma unroll
                for (int v = 0; v < VPT; v++)
                    out[v] = (T)1.0 * out[v] + (T)1.0 * o.px[v];
            }
        }
    }
    if constexpr (EPI == EPI_RESID_MX) {
        float o32[VPT];
#pragma unroll
        for (int v = 0; v < VPT; v++)
            o32[v] = (float)out[v];
        st_vec<float, VPT>(a.yo + e, o32);
        if (a.yo2 != nullptr)
            st_vec<float, VPT>(a.yo2 + e, o32);
        if (a.yo3 != nullptr)
            st_vec<float, VPT>(a.yo3 + e, o32);
    } else {
        st_vec<T, VPT>(a.y + e, out);
        if (a.y2 != nullptr)
            st_vec<T, VPT>(a.y2 + e, out);
        if (a.y3 != nullptr)
            st_vec<T, VPT>(a.y3 + e, out);
    }
    if constexpr (ND == 1) {
        if (a.w != nullptr) {
#pragma unroll
            for (int v = 0; v < VPT; v++)
                dacc[0][v] += (double)out[v] * (double)o.pw[v];
        } else {
#pragma unroll
            for (int v = 0; v < VPT; v++)
                dacc[0][v] += (double)out[v] * (double)out[v];
        }
    } else if constexpr (ND == 3) {
#pragma unroll
        for (int v = 0; v < VPT; v++) {
            const double ov = (double)out[v], wv = (double)o.pw[v];
            dacc[0][v] += ov * wv;
            dacc[1][v] += ov * ov;
            dacc[2][v] += wv * wv;
        }
    }
}

// One row of the SpMV family: the lanes of row i (VPT right-hand sides each, from l0)
// walk nonzeros j0..j1 in chunks of U (all column indices, values and x segments of a
// chunk loaded first, then added in ascending column order, csr_matvec.inl:213-231),
// apply the epilogue, store, and add the row's terms to the fused dots.  nz(q, c, v)
// fetches column index and value of entry q (from global memory or an LDS copy).
template <typename V, typename T, int NV, int VPT, int EPI, int ND, bool NT, typename NZ,
          int UF = 1>
__device__ __forceinline__ void spmv_row(const SpmvArgs<T> &a, int64_t i, int l0, uint32_t j0,
                                         uint32_t j1, NZ &&nz,
                                         double (&dacc)[ND > 0 ? ND : 1][VPT]) {
    constexpr int U = ((16 / VPT) > 2 ? (16 / VPT) : 2) * UF;
    const int64_t e = i * NV + l0;
    T acc[VPT];
    if constexpr (EPI == EPI_AXPY) {
        ld_vec<T, VPT, NT>(a.y + e, acc);
    } else {
#pragma unroll
        for (int v = 0; v < VPT; v++)
            acc[v] = (T)0;
    }
    RowOps<T, VPT> o;
    row_ops_load<T, NV, VPT, EPI, ND>(a, i, e, o);
    for (uint32_t j = j0; j < j1; j += U) {
        uint32_t c[U];
        T av[U];
        T xv[U][VPT];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const bool ok = j + u < j1;
            T v;
            nz(ok ? j + u : j0, c[u], v);
            av[u] = ok ? v : (T)0;
        }
#pragma unroll
        for (int u = 0; u < U; u++)
            ld_vec<T, VPT>(a.x + (int64_t)c[u] * NV + l0, xv[u]);
#pragma unroll
        for (int u = 0; u < U; u++)
            x_transform<T, VPT, EPI>(a, c[u], xv[u]);
#pragma unroll
        for (int u = 0; u < U; u++)
            if (j + u < j1)
#pragma unroll
                for (int v = 0; v < VPT; v++)
                    acc[v] = acc[v] + xv[u][v] * av[u];
    }
    row_finish<T, NV, VPT, EPI, ND>(a, e, l0, acc, o, dacc);
}

constexpr int kStageCap = kSpmvStageCap;
constexpr int kStageCapVec = kSpmvStageCapVec;

template <typename V, typename T, int NV, int VPT, int EPI, int ND, int STAGE>
__global__ void __launch_bounds__(kBlock)
    k_spmv(SpmvArgs<T> a, RedArgs r, FinArgs<T> f, int remap) {
    const StartFlags fl = start_flags(r);
    if (fl.halt || !fl.gate)
        return;
    constexpr int LR = NV / VPT;
    constexpr int RPB = kBlock / LR;
    const V *__restrict__ val = (const V *)a.val;
    const uint32_t *__restrict__ row = a.row;
    const uint32_t *__restrict__ col = a.col;
    const int tid = threadIdx.x, lr = tid / LR, sub = tid % LR;
    const int l0 = sub * VPT;
    const int G = gridDim.x;
    const int L = logical_block(blockIdx.x, G, remap);
    const int64_t r0 = a.n * L / G, r1 = a.n * (L + 1) / G;
    constexpr int NDX = ND > 0 ? ND : 1;
    double dacc[NDX][VPT];
#pragma unroll
    for (int k = 0; k < NDX; k++)
#pragma unroll
        for (int v = 0; v < VPT; v++)
            dacc[k][v] = 0.0;
    // STAGE 7 / 8: as 2 / 4 with the tile copied by 16-byte loads per lane (a
    // wavefront load moves 1 KB instead of 128-256 bytes: the texture-address unit's cost
    // goes by instructions), twice the staging capacity and twice the nonzero chunk
    // (rows of 27 nonzeros and more)
    constexpr bool VEC = STAGE >= 7;
    constexpr int CAP = STAGE ? (VEC ? kStageCapVec : kStageCap) : 1;
    constexpr bool NT = STAGE == 2 || STAGE == 4 || VEC;
    constexpr bool D16 = STAGE == 4 || STAGE == 8; // 16-bit column deltas
    // entries per 16-byte word of the column and value arrays; a vector copy starts at
    // the word holding the tile's first entry and ends with the word holding its last
    constexpr int EC = D16 ? 8 : 4, EV = (int)(16 / sizeof(V));
    constexpr int SLACK = VEC ? 16 : 0;
    __shared__ __align__(16) uint32_t s_col[D16 ? 4 : CAP + SLACK];
    __shared__ __align__(16) int16_t s_c16[D16 ? CAP + SLACK : 8];
    __shared__ __align__(16) V s_val[CAP + SLACK];
    const int16_t *__restrict__ col16 = a.col16;
    constexpr int UF = VEC ? 2 : 1;
    for (int64_t base = r0; base < r1; base += RPB) {
        const int64_t i = base + lr;
        uint32_t jc = 0, jv = 0; // entries at LDS index 0 of the column and value copies
        bool staged = false;
        if constexpr (STAGE) {
            const int64_t bend = base + RPB < r1 ? base + RPB : r1;
            const uint32_t jA = row[base], jB = row[bend];
            staged = jB - jA <= (uint32_t)CAP;
            jc = jv = jA;
            __syncthreads(); // the previous tile's rows are done with the LDS copy
            if constexpr (VEC) {
                jc = jA & ~(uint32_t)(EC - 1);
                jv = jA & ~(uint32_t)(EV - 1);
                if (staged) {
                    const uint32_t nwc = (jB - jc + EC - 1) / EC, nwv = (jB - jv + EV - 1) / EV;
                    const u32x4_t *gc = D16 ? (const u32x4_t *)(col16 + jc) : (const u32x4_t *)(col + jc);
                    u32x4_t *sc = D16 ? (u32x4_t *)s_c16 : (u32x4_t *)s_col;
                    for (uint32_t k = tid; k < nwc; k += kBlock)
                        sc[k] = __builtin_nontemporal_load(gc + k);
                    const u32x4_t *gv = (const u32x4_t *)(val + jv);
                    for (uint32_t k = tid; k < nwv; k += kBlock)
                        ((u32x4_t *)s_val)[k] = __builtin_nontemporal_load(gv + k);
                }
            } else if (staged) {
                for (uint32_t k = tid; k < jB - jA; k += kBlock) {
                    if constexpr (D16)
                        s_c16[k] = ld_nt(col16 + jA + k);
                    else
                        s_col[k] = NT ? ld_nt(col + jA + k) : col[jA + k];
                    s_val[k] = NT ? ld_nt(val + jA + k) : val[jA + k];
                }
            }
            __syncthreads();
        }
        if (i < r1) {
            const uint32_t j0 = row[i], j1 = row[i + 1];
            auto nzf = [&](uint32_t q, uint32_t &c, T &v) {
                                                     if (STAGE && staged) {
                                                         if constexpr (D16)
                                                             c = (uint32_t)(i + s_c16[q - jc]);
                                                         else
                                                             c = s_col[q - jc];
                                                         v = to_t<T>(s_val[q - jv]);
                                                     } else {
                                                         // a tile too large for LDS: plain loads
                                                         // (non-temporal loads of entries a row
                                                         // apart would fetch each line per lane)
                                                         if constexpr (D16)
                                                             c = (uint32_t)(i + col16[q]);
                                                         else
                                                             c = col[q];
                                                         v = to_t<T>(val[q]);
                                                     }
                                                 };
            spmv_row<V, T, NV, VPT, EPI, ND, NT, decltype(nzf) &, UF>(a, i, l0, j0, j1, nzf, dacc);
        }
    }
    if constexpr (ND > 0)
        reduce_and_finish<T, NV, ND, VPT>(dacc, r, f, L);
}

// STAGE = 3 (values of 4 or 8 bytes): tiles of RPB rows staged through LDS with a
// double buffer filled by the direct global-to-LDS loads of gfx950 (global_load_lds:
// each lane's dword goes to LDS without passing through registers).  While the rows of
// tile t gather x, the row pointers, column indices and values of tile t+1 are already
// on their way into the other buffer, so a tile costs one dependent round trip to
// memory (the x gathers) instead of three (row pointers, column indices, x).  Tile t+1
// starts at tile t's last row pointer, so one scalar row-pointer load, issued a tile
// ahead, bounds each copy.  A tile with more than CAP nonzeros reads them directly.
// D16 (STAGE = 5): the column indices are copied as 16-bit offsets from the row.
constexpr int kPipeNnzPerRow = 8;

template <typename V, typename T, int NV, int VPT, int EPI, int ND, bool D16>
__global__ void __launch_bounds__(kBlock)
    k_spmv_pipe(SpmvArgs<T> a, RedArgs r, FinArgs<T> f, int remap) {
    static_assert(sizeof(V) % 4 == 0, "values are copied as dwords");
    const StartFlags fl = start_flags(r);
    if (fl.halt || !fl.gate)
        return;
    constexpr int LR = NV / VPT;
    constexpr int RPB = kBlock / LR;
    constexpr int CAP = RPB * kPipeNnzPerRow;
    constexpr int VW = (int)(sizeof(V) / 4);
    const V *__restrict__ val = (const V *)a.val;
    const uint32_t *__restrict__ row = a.row;
    const uint32_t *__restrict__ col = a.col;
    const int tid = threadIdx.x, lr = tid / LR, sub = tid % LR;
    const int wave = tid >> 6, lane = tid & 63;
    const int l0 = sub * VPT;
    const int G = gridDim.x;
    const int L = logical_block(blockIdx.x, G, remap);
    const int64_t r0 = a.n * L / G, r1 = a.n * (L + 1) / G;
    constexpr int NDX = ND > 0 ? ND : 1;
    double dacc[NDX][VPT];
#pragma unroll
    for (int k = 0; k < NDX; k++)
#pragma unroll
        for (int v = 0; v < VPT; v++)
            dacc[k][v] = 0.0;
    // 16-bit offsets: a staged tile's [jA/2, (jB+1)/2) holds at most CAP/2 + 1 dwords
    constexpr int CAPC = D16 ? CAP / 2 + 2 : CAP;
    __shared__ uint32_t s_row[2][RPB + 1];
    __shared__ uint32_t s_col[2][CAPC];
    __shared__ __align__(16) uint32_t s_vw[2][CAP * VW];
    // copy of tile rows [base, bend] (row pointers) and its nonzeros [jA, jB) into buffer b
    auto issue = [&](int b, int64_t base, int64_t bend, uint32_t jA, uint32_t jB) {
        const uint32_t nr = (uint32_t)(bend - base) + 1;
        for (uint32_t k0 = wave * 64; k0 < nr; k0 += kBlock)
            if (k0 + lane < nr)
                __builtin_amdgcn_global_load_lds(row + base + k0 + lane, &s_row[b][k0], 4, 0, 0);
        const uint32_t cnt = jB - jA;
        if (cnt <= (uint32_t)CAP) {
            if constexpr (D16) {
                // whole dwords of 16-bit offsets: [jA/2, (jB+1)/2) (the array is padded)
                const uint32_t d0 = jA >> 1, nw = ((jB + 1) >> 1) - d0;
                const uint32_t *cw = (const uint32_t *)a.col16 + d0;
                for (uint32_t k0 = wave * 64; k0 < nw; k0 += kBlock)
                    if (k0 + lane < nw)
                        __builtin_amdgcn_global_load_lds(cw + k0 + lane, &s_col[b][k0], 4, 0, 0);
            } else {
                for (uint32_t k0 = wave * 64; k0 < cnt; k0 += kBlock)
                    if (k0 + lane < cnt)
                        __builtin_amdgcn_global_load_lds(col + jA + k0 + lane, &s_col[b][k0], 4, 0, 0);
            }
            const uint32_t *vw = (const uint32_t *)val + (size_t)jA * VW;
            for (uint32_t k0 = wave * 64; k0 < cnt * VW; k0 += kBlock)
                if (k0 + lane < cnt * VW)
                    __builtin_amdgcn_global_load_lds(vw + k0 + lane, &s_vw[b][k0], 4, 0, 0);
        }
    };
    const int ntiles = (int)((r1 - r0 + RPB - 1) / RPB);
    int64_t base = r0, bend = r0 + RPB < r1 ? r0 + RPB : r1;
    uint32_t jA = 0, jB = 0, jN = 0;
    if (ntiles > 0) {
        jA = row[base];
        jB = row[bend];
        issue(0, base, bend, jA, jB);
        if (bend < r1)
            jN = row[bend + RPB < r1 ? bend + RPB : r1];
    }
    for (int t = 0; t < ntiles; t++) {
        const int b = t & 1;
        // every lane's copies of tile t have landed and every row of tile t-1 is done
        // with buffer b^1
        __syncthreads();
        const int64_t nbase = bend, nbend = nbase + RPB < r1 ? nbase + RPB : r1;
        uint32_t jN2 = 0;
        if (t + 1 < ntiles) {
            issue(b ^ 1, nbase, nbend, jB, jN);
            if (nbend < r1)
                jN2 = row[nbend + RPB < r1 ? nbend + RPB : r1];
        }
        const int64_t i = base + lr;
        if (i < bend) {
            const uint32_t j0 = s_row[b][lr], j1 = s_row[b][lr + 1];
            const bool staged = jB - jA <= (uint32_t)CAP;
            spmv_row<V, T, NV, VPT, EPI, ND, false>(a, i, l0, j0, j1,
                                                    [&](uint32_t q, uint32_t &c, T &v) {
                                                        if (staged) {
                                                            if constexpr (D16)
                                                                c = (uint32_t)(
                                                                    i + ((const int16_t *)s_col[b])
                                                                            [q - 2 * (jA >> 1)]);
                                                            else
                                                                c = s_col[b][q - jA];
                                                            V w;
                                                            __builtin_memcpy(&w, &s_vw[b][(q - jA) * VW],
                                                                             sizeof(V));
                                                            v = to_t<T>(w);
                                                        } else {
                                                            if constexpr (D16)
                                                                c = (uint32_t)(i + a.col16[q]);
                                                            else
                                                                c = col[q];
                                                            v = to_t<T>(val[q]);
                                                        }
                                                    },
                                                    dacc);
        }
        base = nbase;
        bend = nbend;
        jA = jB;
        jB = jN;
        jN = jN2;
    }
    if constexpr (ND > 0)
        reduce_and_finish<T, NV, ND, VPT>(dacc, r, f, L);
}

template <typename V, typename T, int NV, int VPT, int STAGE>
void spmv_nvv(int epi, int nd, const SpmvArgs<T> &a, const RedArgs &r, const FinArgs<T> &f,
                     hipStream_t s) {
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
    constexpr int LR = NV / VPT;
    int64_t need = (a.n * LR + kBlock - 1) / kBlock;
    if (need < G)
        G = (int)(need > 0 ? need : 1);
    dim3 g(G), b(kBlock);
#define XAMG_SPMV_CASE(E, D)                                                                       \
    if (epi == E && nd == D) {                                                                     \
        if constexpr ((STAGE == 3 || STAGE == 5) && sizeof(V) % 4 == 0)                            \
            XAMG_LAUNCH((k_spmv_pipe<V, T, NV, VPT, E, D, STAGE == 5>), g, b, 0, s, a, r, f,           \
                               c.xcd_remap);                                                        \
        else                                                                                       \
            XAMG_LAUNCH((k_spmv<V, T, NV, VPT, E, D, (STAGE == 3 ? 1 : STAGE == 5 ? 4 : STAGE)>),    \
                               g, b, 0, s, a, r, f, c.xcd_remap);                                   \
        launch_finish<T, NV>(D, r, f, G, s);                                                       \
        return;                                                                                    \
    }
    XAMG_SPMV_CASE(EPI_AXY, 0)
    XAMG_SPMV_CASE(EPI_AXY, 1)
    XAMG_SPMV_CASE(EPI_AXY, 3)
    XAMG_SPMV_CASE(EPI_AXPY, 0)
    XAMG_SPMV_CASE(EPI_RESID, 0)
    XAMG_SPMV_CASE(EPI_RESID, 1)
    XAMG_SPMV_CASE(EPI_JACOBI, 0)
    XAMG_SPMV_CASE(EPI_CHEB, 0)
    XAMG_SPMV_CASE(EPI_CHEB_LAST, 0)
    XAMG_SPMV_CASE(EPI_CHEB_S, 0)
    XAMG_SPMV_CASE(EPI_CHEB_LAST_S, 0)
    XAMG_SPMV_CASE(EPI_CHEB_SF, 0)
    XAMG_SPMV_CASE(EPI_CHEB_FU, 0)
    XAMG_SPMV_CASE(EPI_CHEB_U, 0)
    XAMG_SPMV_CASE(EPI_CHEB_LAST_U, 0)
    if constexpr (std::is_same<T, double>::value && std::is_same<V, double>::value) {
        XAMG_SPMV_CASE(EPI_RESID_MX, 0)
        XAMG_SPMV_CASE(EPI_RESID_MX, 1)
    }
#undef XAMG_SPMV_CASE
    throw std::runtime_error("xamg: unsupported spmv epilogue/dot combination");
}

template <typename V, typename T, int NV, int STAGE>
void spmv_nv(int epi, int nd, const SpmvArgs<T> &a, const RedArgs &r, const FinArgs<T> &f,
                    hipStream_t s) {
    int vpt = a.cfg_vpt > 0 ? a.cfg_vpt : spmv_vpt_choice(NV, (int)sizeof(T));
    if constexpr (NV == 1) {
        spmv_nvv<V, T, 1, 1, STAGE>(epi, nd, a, r, f, s);
    } else if constexpr (NV == 2) {
        if (vpt == 1)
            spmv_nvv<V, T, 2, 1, STAGE>(epi, nd, a, r, f, s);
        else
            spmv_nvv<V, T, 2, 2, STAGE>(epi, nd, a, r, f, s);
    } else if constexpr (NV == 4) {
        if (vpt == 2)
            spmv_nvv<V, T, 4, 2, STAGE>(epi, nd, a, r, f, s);
        else
            spmv_nvv<V, T, 4, 4, STAGE>(epi, nd, a, r, f, s);
    } else if constexpr (NV == 8) {
        if (vpt == 2)
            spmv_nvv<V, T, 8, 2, STAGE>(epi, nd, a, r, f, s);
        else if (vpt == 8)
            spmv_nvv<V, T, 8, 8, STAGE>(epi, nd, a, r, f, s);
        else
            spmv_nvv<V, T, 8, 4, STAGE>(epi, nd, a, r, f, s);
    } else {
        if (vpt == 8)
            spmv_nvv<V, T, 16, 8, STAGE>(epi, nd, a, r, f, s);
        else if (vpt == 16)
            spmv_nvv<V, T, 16, 16, STAGE>(epi, nd, a, r, f, s);
        else
            spmv_nvv<V, T, 16, 4, STAGE>(epi, nd, a, r, f, s);
    }
}

// Launch of the SpMV family for one staging mode (0: column indices and values read by
// each row's lanes, 1: staged through LDS per tile of rows, 2: as 1 with non-temporal
// streamed operands, 3: double-buffered LDS staging by direct global-to-LDS loads; BF16
// values take mode 1 there; 4 and 5: modes 2 and 3 with 16-bit column offsets, BF16
// values in mode 5 take mode 4).
template <typename V, typename T, int STAGE>
void launch_spmv_mode(int epi, int ndots, int nv, const SpmvArgs<T> &a, const RedArgs &r,
                      const FinArgs<T> &f, hipStream_t s) {
    switch (nv) {
    case 1:
        spmv_nv<V, T, 1, STAGE>(epi, ndots, a, r, f, s);
        break;
    case 2:
        spmv_nv<V, T, 2, STAGE>(epi, ndots, a, r, f, s);
        break;
    case 4:
        spmv_nv<V, T, 4, STAGE>(epi, ndots, a, r, f, s);
        break;
    case 8:
        spmv_nv<V, T, 8, STAGE>(epi, ndots, a, r, f, s);
        break;
    case 16:
        spmv_nv<V, T, 16, STAGE>(epi, ndots, a, r, f, s);
        break;
    default:
        throw std::runtime_error("xamg: nv must be one of 1, 2, 4, 8, 16");
    }
    XAMG_HIP_CHECK(hipGetLastError());
}

} // namespace xamg
