// Batched Krylov solves over a concatenated block-diagonal CSR (see batch.hpp).
#include "batch.hpp"

#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <type_traits>

namespace xamg {

BatchCfg &batch_cfg() {
    static BatchCfg c;
    return c;
}

namespace {

constexpr int BT = 256;      // threads per workgroup
// Second Jacobi sweep from a zero start, with the first folded in: the gathered and the
// row's own first-sweep values z1_j = (w d_j) q_j are formed from q as the first sweep
// computes them (EW_JACOBI0), so no z1 vector is written or read.
constexpr int BEPI_JACOBI2 = 100;
constexpr int PSTRIDE = kBatchDots * kBatchMaxNV; // partial sums per tile

template <typename T>
struct BSpmvArgs {
    const void *val = nullptr; // V[nnz]
    const T *x = nullptr;
    T *y = nullptr, *y2 = nullptr, *y3 = nullptr;
    const T *b = nullptr;
    const T *xin = nullptr;
    const T *d = nullptr;     // inverse diagonal per global row
    const double *d64 = nullptr; // FP64 inverse diagonal (Jacobi sweeps with FP64 values)
    const T *w = nullptr;     // dot partner
    const T *wcoef = nullptr; // JACOBI relax factors, per (system, RHS)
    float *yo = nullptr, *yo2 = nullptr, *yo3 = nullptr; // EPI_RESID_MX outputs
    const int *gate = nullptr;    // per system: tiles run only while set
    const int *copy_if = nullptr; // per system: yo2 / yo3 written only while set
};

template <typename T>
struct BEwArgs {
    T *z = nullptr;
    const T *x = nullptr, *y = nullptr;
    const T *ca = nullptr, *cb = nullptr, *cc = nullptr; // per (system, RHS); null = 1
    const T *d = nullptr;
    T c0 = 0;
    const void *xo = nullptr, *yo = nullptr; // FP32 inputs of EW_MX_XUPD
    const float *ao = nullptr, *bo = nullptr; // FP32 coefficients of EW_MX_XUPD
    T *y2 = nullptr, *y3 = nullptr;
    const T *w = nullptr;
    const int *keep = nullptr; // per system: z unchanged while set
    // per system, while set: AXPBY_Z also stores z into y2 and dots z with itself;
    // AXPBYPCZ sets z = x (restart of the recurrence from its residual)
    const int *copy = nullptr;
    int freeze_zero = 0;       // AXPBYPCZ / MX_XUPD: z unchanged where a and b are both 0
};

struct BFinArgs {
    int fin = FIN_NONE;
    int nd = 0;
    double abs_tol = 0, rel_tol = 0, delta2 = 0;
    double theta = 0; // FIN_BETA with (r, r0), (r, r), (r0, r0): restart threshold
    double gap2 = 0;  // FIN_MX_REPL: restart when true > gap2 * recurrence residual^2
    int max_iters = 1;
    const int *gate = nullptr; // per system
};

// Sum of the per-thread accumulators of one tile over the workgroup, in a fixed order,
// stored as the tile's partial sums by thread 0 (slots of RHS the system does not have
// are stored as 0).
template <int ND>
__device__ void tile_partial(double (&acc)[ND][kBatchMaxNV], int nv, double *out) {
    __shared__ double lds[BT / 64][ND][kBatchMaxNV];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < ND; k++)
#pragma unroll
        for (int l = 0; l < kBatchMaxNV; l++) {
            if (l >= nv) // nv is the same for the whole workgroup
                break;
            double v = acc[k][l];
#pragma unroll
            for (int off = 32; off > 0; off >>= 1)
                v += __shfl_xor(v, off, 64);
            if (lane == 0)
                lds[wave][k][l] = v;
        }
    __syncthreads();
    if (threadIdx.x == 0)
        for (int k = 0; k < ND; k++)
            for (int l = 0; l < kBatchMaxNV; l++) {
                double v = 0.0;
                if (l < nv) {
                    v = lds[0][k][l];
                    for (int q = 1; q < BT / 64; q++)
                        v += lds[q][k][l];
                }
                out[k * kBatchMaxNV + l] = v;
            }
    __syncthreads();
}

// Element (row, RHS) of a tile: q = local row * nv + l.  Division by nv in {1..4}
// without a divide instruction.
__device__ __forceinline__ uint32_t div_nv(uint32_t q, int nv) {
    if (nv == 1)
        return q;
    if (nv == 2)
        return q >> 1;
    if (nv == 4)
        return q >> 2;
    return (uint32_t)(((uint64_t)q * 0xAAAAAAABull) >> 33); // exact q / 3 for 32-bit q
}

template <typename T>
__device__ __forceinline__ T sel4(const T (&v)[kBatchMaxNV], int l) {
    T r = v[0];
    r = l == 1 ? v[1] : r;
    r = l == 2 ? v[2] : r;
    r = l == 3 ? v[3] : r;
    return r;
}

// Adds v to slot l of a per-RHS accumulator (predicated, no dynamic register index).
__device__ __forceinline__ void acc_slot(double (&acc)[kBatchMaxNV], int l, double v) {
#pragma unroll
    for (int q = 0; q < kBatchMaxNV; q++)
        if (q == l)
            acc[q] += v;
}

// y = epilogue(A x) over the tiles of the running systems.  One lane per vector element
// (row, RHS): lanes of one row read the same column / value entries; the row's nonzeros
// are loaded U at a time and summed in ascending column order (csr_matvec.inl:213-231,
// x * val).  U is the smallest compiled chunk that holds the batch's longest row (5 for
// 5-point stencils), so a row costs one round of loads and no padding loads.  A lane
// keeps EPT elements (BT apart) in flight: their row operands, nonzeros and gathers are
// all loaded before the first sum, and the lane adds them to its dot accumulators in
// element order, so the dots do not depend on EPT.  A workgroup takes L.tps consecutive
// tiles per step of the grid stride (tiles are the dot blocks: neither changes a sum).
// With FP64 values and FP32 vectors (the recurrence products of the FP32 solver with the
// FP64 operator) the row sums run in FP64 and are rounded once, at the store.
template <typename V, typename T, int EPI, int ND, int U, int EPT>
__global__ void __launch_bounds__(BT)
    kb_spmv(BLayout L, const int *halt, BSpmvArgs<T> a, double *part) {
    using Acc = typename std::conditional<(sizeof(V) > sizeof(T)), V, T>::type;
    constexpr bool kWide = !std::is_same<Acc, T>::value;
    constexpr bool kJac = EPI == EPI_JACOBI || EPI == BEPI_JACOBI2;
    const V *val = (const V *)a.val;
    const int tps = L.tps;
    for (int t0 = blockIdx.x * tps; t0 < L.ntiles; t0 += gridDim.x * tps) {
        const int t1 = min(L.ntiles, t0 + tps);
        for (int t = t0; t < t1; t++) {
            const int s = L.tsys[t];
            if ((halt != nullptr && halt[s]) || (a.gate != nullptr && !a.gate[s]))
                continue;
            const int nv = L.nv[s];
            const uint32_t rs = L.srow[s];
            const int64_t eb = L.voff[s];
            const int64_t vb = eb - (int64_t)rs * nv;
            const uint32_t q0 = (L.trow[t] - rs) * nv, q1 = (L.trow[t + 1] - rs) * nv;
            const bool cp = a.copy_if == nullptr || a.copy_if[s];
            constexpr int NDX = ND > 0 ? ND : 1;
            double dacc[NDX][kBatchMaxNV];
#pragma unroll
            for (int k = 0; k < NDX; k++)
#pragma unroll
                for (int l = 0; l < kBatchMaxNV; l++)
                    dacc[k][l] = 0.0;
            for (uint32_t qb = q0 + threadIdx.x; qb < q1; qb += EPT * BT) {
                bool on[EPT];
                int l[EPT];
                uint32_t i[EPT], j0[EPT], n[EPT];
                int64_t e[EPT];
                // the epilogue's and the dots' row operands are loaded before the gathers,
                // so their latency overlaps the nonzero loads
                T pb[EPT], px[EPT], pw[EPT];
                Acc pd[EPT], wl[EPT]; // Jacobi: inverse diagonal and relax factor
#pragma unroll
                for (int k = 0; k < EPT; k++) {
                    on[k] = qb + k * BT < q1;
                    // an element past the tile stands in for the step's first element
                    const uint32_t q = on[k] ? qb + k * BT : qb;
                    const uint32_t li = div_nv(q, nv);
                    l[k] = (int)(q - li * nv);
                    i[k] = rs + li;
                    e[k] = eb + q;
                    pb[k] = px[k] = pw[k] = (T)0;
                    pd[k] = wl[k] = (Acc)0;
                    if constexpr (EPI == EPI_RESID || EPI == EPI_RESID_MX || kJac)
                        pb[k] = a.b[e[k]];
                    if constexpr (EPI == EPI_JACOBI)
                        px[k] = a.xin[e[k]];
                    if constexpr (kJac) {
                        if constexpr (kWide)
                            pd[k] = a.d64[i[k]];
                        else
                            pd[k] = a.d[i[k]];
                        wl[k] = (Acc)a.wcoef[s * kBatchMaxNV + l[k]];
                    }
                    if constexpr (ND == 3)
                        pw[k] = a.w[e[k]];
                    else if constexpr (ND == 1)
                        pw[k] = a.w != nullptr ? a.w[e[k]] : (T)0;
                    j0[k] = L.row[i[k]];
                    n[k] = L.row[i[k] + 1] - j0[k];
                }
                uint32_t nmax = n[0];
#pragma unroll
                for (int k = 1; k < EPT; k++)
                    nmax = max(nmax, n[k]);
                Acc acc[EPT];
#pragma unroll
                for (int k = 0; k < EPT; k++)
                    acc[k] = (Acc)0;
                for (uint32_t jo = 0; jo < nmax; jo += U) {
                    uint32_t c[EPT][U];
                    Acc v[EPT][U], xa[EPT][U];
                    T xv[EPT][U];
#pragma unroll
                    for (int k = 0; k < EPT; k++)
#pragma unroll
                        for (int u = 0; u < U; u++) {
                            const bool ok = jo + u < n[k];
                            // a padding slot gathers the row's own element (in its system)
                            c[k][u] = ok ? L.col[j0[k] + jo + u] : i[k];
                            v[k][u] = ok ? (Acc)val[j0[k] + jo + u] : (Acc)0;
                        }
#pragma unroll
                    for (int k = 0; k < EPT; k++)
#pragma unroll
                        for (int u = 0; u < U; u++)
                            xv[k][u] = a.x[vb + (int64_t)c[k][u] * nv + l[k]];
                    if constexpr (EPI == BEPI_JACOBI2) {
                        Acc dv[EPT][U];
#pragma unroll
                        for (int k = 0; k < EPT; k++)
#pragma unroll
                            for (int u = 0; u < U; u++) {
                                if constexpr (kWide)
                                    dv[k][u] = a.d64[c[k][u]];
                                else
                                    dv[k][u] = a.d[c[k][u]];
                            }
#pragma unroll
                        for (int k = 0; k < EPT; k++)
#pragma unroll
                            for (int u = 0; u < U; u++)
                                xa[k][u] = (wl[k] * dv[k][u]) * (Acc)xv[k][u];
                    } else {
#pragma unroll
                        for (int k = 0; k < EPT; k++)
#pragma unroll
                            for (int u = 0; u < U; u++)
                                xa[k][u] = (Acc)xv[k][u];
                    }
#pragma unroll
                    for (int k = 0; k < EPT; k++)
#pragma unroll
                        for (int u = 0; u < U; u++)
                            if (jo + u < n[k])
                                acc[k] = acc[k] + xa[k][u] * v[k][u];
                }
#pragma unroll
                for (int k = 0; k < EPT; k++) {
                    if (!on[k])
                        break;
                    const T ax = (T)acc[k];
                    T out;
                    if constexpr (EPI == EPI_AXY) {
                        out = ax;
                    } else if constexpr (EPI == EPI_RESID || EPI == EPI_RESID_MX) {
                        out = (T)1.0 * pb[k] + (T)(-1.0) * ax;
                    } else if constexpr (EPI == EPI_JACOBI) { // x_new = x + (w d)(b - A x)
                        out = (T)((Acc)px[k] + (wl[k] * pd[k]) * ((Acc)pb[k] - acc[k]));
                    } else { // BEPI_JACOBI2: x1 = (w d) b, then x_new = x1 + (w d)(b - A x1)
                        const Acc x1 = (wl[k] * pd[k]) * (Acc)pb[k];
                        out = (T)(x1 + (wl[k] * pd[k]) * ((Acc)pb[k] - acc[k]));
                    }
                    if constexpr (EPI == EPI_RESID_MX) {
                        const float o = (float)out;
                        a.yo[e[k]] = o;
                        if (cp) {
                            if (a.yo2 != nullptr)
                                a.yo2[e[k]] = o;
                            if (a.yo3 != nullptr)
                                a.yo3[e[k]] = o;
                        }
                    } else {
                        a.y[e[k]] = out;
                        if (a.y2 != nullptr)
                            a.y2[e[k]] = out;
                        if (a.y3 != nullptr)
                            a.y3[e[k]] = out;
                    }
                    if constexpr (ND == 1) {
                        const double wv = a.w != nullptr ? (double)pw[k] : (double)out;
                        acc_slot(dacc[0], l[k], (double)out * wv);
                    } else if constexpr (ND == 3) {
                        const double o = (double)out, wv = (double)pw[k];
                        acc_slot(dacc[0], l[k], o * wv);
                        acc_slot(dacc[1], l[k], o * o);
                        acc_slot(dacc[2], l[k], wv * wv);
                    }
                }
            }
            if constexpr (ND > 0)
                tile_partial<ND>(dacc, nv, part + (int64_t)t * PSTRIDE);
        }
    }
}

// Element-wise updates over the tiles of the running systems (blas.inl / blas_ext.inl
// operation order, per-(system, RHS) coefficients), one lane per vector element.
template <typename T, int OP, int ND>
__global__ void __launch_bounds__(BT)
    kb_ew(BLayout L, const int *halt, BEwArgs<T> a, double *part) {
    const int tps = L.tps;
    for (int t0 = blockIdx.x * tps; t0 < L.ntiles; t0 += gridDim.x * tps)
    for (int t = t0; t < min(L.ntiles, t0 + tps); t++) {
        const int s = L.tsys[t];
        if (halt != nullptr && halt[s])
            continue;
        const bool keep = a.keep != nullptr && a.keep[s];
        const bool copy = a.copy != nullptr && a.copy[s];
        const int nv = L.nv[s];
        const uint32_t rs = L.srow[s];
        const int64_t eb = L.voff[s];
        const uint32_t q0 = (L.trow[t] - rs) * nv, q1 = (L.trow[t + 1] - rs) * nv;
        T ca[kBatchMaxNV], cb[kBatchMaxNV], cc[kBatchMaxNV];
        bool frozen[kBatchMaxNV];
#pragma unroll
        for (int l = 0; l < kBatchMaxNV; l++) {
            const int qq = s * kBatchMaxNV + l;
            ca[l] = a.ca != nullptr ? a.ca[qq] : (T)1.0;
            cb[l] = a.cb != nullptr ? a.cb[qq] : (T)1.0;
            cc[l] = a.cc != nullptr ? a.cc[qq] : (T)1.0;
            frozen[l] = false;
            if constexpr (OP == EW_AXPBYPCZ)
                frozen[l] = a.freeze_zero && ca[l] == (T)0 && cb[l] == (T)0;
            else if constexpr (OP == EW_MX_XUPD)
                frozen[l] = a.freeze_zero && a.ao[qq] == 0.0f && a.bo[qq] == 0.0f;
        }
        float ao[kBatchMaxNV], bo[kBatchMaxNV];
        if constexpr (OP == EW_MX_XUPD) {
#pragma unroll
            for (int l = 0; l < kBatchMaxNV; l++) {
                ao[l] = a.ao[s * kBatchMaxNV + l];
                bo[l] = a.bo[s * kBatchMaxNV + l];
            }
        }
        constexpr int NDX = ND > 0 ? ND : 1;
        double dacc[NDX][kBatchMaxNV];
#pragma unroll
        for (int k = 0; k < NDX; k++)
#pragma unroll
            for (int l = 0; l < kBatchMaxNV; l++)
                dacc[k][l] = 0.0;
        // Every operand of an element is loaded before the first test or use of any of
        // them (absent optional operands read x again), and the K elements of a thread in
        // a step (BT apart) are all loaded before the first store: the elements of a tile
        // cost one round trip to memory per step instead of one each.  The dots add the
        // elements in the thread's ascending order.
        constexpr int K = 4;
        const T *zsrc = a.keep != nullptr ? a.z : a.x;
        const T *wsrc = (ND > 0 && a.w != nullptr) ? a.w : a.x;
        for (uint32_t qb = q0 + threadIdx.x; qb < q1; qb += K * BT) {
            T xe[K], ye[K], ze[K], we[K], de[K];
            float xo[K], yo[K];
#pragma unroll
            for (int u = 0; u < K; u++) {
                // an element past the tile reads the step's first element again
                const uint32_t q = qb + u * BT < q1 ? qb + u * BT : qb;
                const uint32_t li = div_nv(q, nv);
                const int64_t e = eb + q;
                xe[u] = ye[u] = ze[u] = we[u] = de[u] = (T)0;
                xo[u] = yo[u] = 0.0f;
                if constexpr (OP == EW_AXPBY_Z) {
                    xe[u] = a.x[e];
                    ye[u] = a.y[e];
                    ze[u] = zsrc[e];
                } else if constexpr (OP == EW_AXPBYPCZ) {
                    xe[u] = a.x[e];
                    ye[u] = a.y[e];
                    ze[u] = a.z[e];
                } else if constexpr (OP == EW_JACOBI0) {
                    xe[u] = a.x[e];
                    de[u] = a.d[rs + li];
                } else if constexpr (OP == EW_COPY3) {
                    xe[u] = a.x[e];
                } else if constexpr (OP == EW_MX_XUPD) {
                    xo[u] = ((const float *)a.xo)[e];
                    yo[u] = ((const float *)a.yo)[e];
                    ze[u] = a.z[e];
                }
                if constexpr (ND > 0)
                    we[u] = wsrc[e];
            }
#pragma unroll
            for (int u = 0; u < K; u++) {
                const uint32_t q = qb + u * BT;
                if (q >= q1)
                    break;
                const uint32_t li = div_nv(q, nv);
                const int l = (int)(q - li * nv);
                const int64_t e = eb + q;
                T out;
                if constexpr (OP == EW_FILL) {
                    out = a.c0;
                } else if constexpr (OP == EW_AXPBY_Z) {
                    out = keep ? ze[u] : sel4(ca, l) * xe[u] + sel4(cb, l) * ye[u];
                    if (copy && a.y2 != nullptr)
                        a.y2[e] = out;
                } else if constexpr (OP == EW_AXPBYPCZ) {
                    if (copy)
                        out = xe[u];
                    else
                        out = (keep || sel4(frozen, l))
                                  ? ze[u]
                                  : sel4(ca, l) * xe[u] + sel4(cb, l) * ye[u] +
                                        sel4(cc, l) * ze[u];
                } else if constexpr (OP == EW_JACOBI0) {
                    out = (sel4(ca, l) * de[u]) * xe[u];
                } else if constexpr (OP == EW_COPY3) {
                    out = (T)1.0 * xe[u];
                    if (a.y2 != nullptr)
                        a.y2[e] = out;
                    if (a.y3 != nullptr)
                        a.y3[e] = out;
                } else { // EW_MX_XUPD
                    out = sel4(frozen, l) ? ze[u]
                                          : (T)sel4(ao, l) * (T)xo[u] +
                                                (T)sel4(bo, l) * (T)yo[u] + (T)1.0 * ze[u];
                }
                a.z[e] = out;
                if constexpr (ND == 1) {
                    const double wv = (a.w != nullptr && !copy) ? (double)we[u] : (double)out;
                    acc_slot(dacc[0], l, (double)out * wv);
                } else if constexpr (ND == 3) {
                    const double o = (double)out, wv = copy ? o : (double)we[u];
                    acc_slot(dacc[0], l, o * wv);
                    acc_slot(dacc[1], l, o * o);
                    acc_slot(dacc[2], l, wv * wv);
                }
            }
        }
        if constexpr (ND > 0)
            tile_partial<ND>(dacc, nv, part + (int64_t)t * PSTRIDE);
    }
}

// Per-system convergence test (convergence.h:209-262 semantics, NaN residuals end the
// system unconverged).  Returns true when the system is done.
template <typename F>
__device__ bool bconv_check(BState<F> S, int ns, int s, int nv, const F *rn, bool initial,
                            double abs_tol, double rel_tol, F *iconv) {
    double *res0 = S.cv + BC_RES0 * ns * kBatchMaxNV, *res = S.cv + BC_RES * ns * kBatchMaxNV,
           *tol = S.cv + BC_TOL * ns * kBatchMaxNV;
    const int q0 = s * kBatchMaxNV;
    if (initial)
        for (int l = 0; l < nv; l++) {
            res0[q0 + l] = (double)rn[l];
            tol[q0 + l] = fmax(abs_tol, rel_tol * sqrt(res0[q0 + l]));
        }
    for (int l = 0; l < nv; l++)
        if (!S.status[q0 + l])
            res[q0 + l] = (double)rn[l];
    int flag = 1, bad = 0;
    for (int l = 0; l < nv; l++) {
        if (S.status[q0 + l])
            continue;
        const F ar = (F)sqrt(res[q0 + l]);
        if (ar > (F)tol[q0 + l]) {
            flag = 0;
        } else if (ar == ar) {
            S.status[q0 + l] = 1;
        } else {
            flag = 0;
            bad = 1;
        }
    }
    if (flag || bad)
        return true;
    if (iconv != nullptr)
        for (int l = 0; l < nv; l++)
            if (S.status[q0 + l])
                iconv[q0 + l] = (F)0;
    return false;
}

template <typename F>
__device__ void bhalt(BState<F> S, int ns, int s) {
    if (!S.sy[BY_HALT * ns + s]) {
        S.sy[BY_HALT * ns + s] = 1;
        atomicAdd(S.ndone, 1);
    }
}

// Per-system sums of the tile partials (tile order) and the solver's scalar update,
// one thread per system (pbicgstab.inl:108-204 scalar steps; ir.inl / residual
// replacement for the FP32 recurrences).  F is the precision of the recurrences.
template <typename F>
__global__ void __launch_bounds__(256) kb_finish(BLayout L, BState<F> S, BFinArgs f) {
    // one wavefront per system: lanes sum interleaved tiles, then a fixed xor tree
    const int s = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int ns = L.nsys;
    if (s >= ns)
        return;
    int *sy = S.sy;
    auto SY = [&](int fld) -> int & { return sy[fld * ns + s]; };
    const bool init = f.fin == FIN_INIT || f.fin == FIN_MX_INIT;
    if (!init && SY(BY_HALT))
        return;
    if (f.gate != nullptr && !f.gate[s])
        return;
    const int nv = L.nv[s];
    double sums[kBatchDots][kBatchMaxNV];
#pragma unroll
    for (int k = 0; k < kBatchDots; k++)
#pragma unroll
        for (int l = 0; l < kBatchMaxNV; l++)
            sums[k][l] = 0.0;
    for (int t = L.stile[s] + lane; t < L.stile[s + 1]; t += 64) {
        const double *p = S.part + (int64_t)t * PSTRIDE;
#pragma unroll
        for (int k = 0; k < kBatchDots; k++)
#pragma unroll
            for (int l = 0; l < kBatchMaxNV; l++)
                if (k < f.nd && l < nv)
                    sums[k][l] += p[k * kBatchMaxNV + l];
    }
#pragma unroll
    for (int k = 0; k < kBatchDots; k++)
#pragma unroll
        for (int l = 0; l < kBatchMaxNV; l++)
#pragma unroll
            for (int off = 32; off > 0; off >>= 1)
                sums[k][l] += __shfl_xor(sums[k][l], off, 64);
    if (lane != 0)
        return;
    const int NS4 = ns * kBatchMaxNV, q0 = s * kBatchMaxNV;
    auto SC = [&](int fld, int l) -> F & { return S.sc[fld * NS4 + q0 + l]; };
    F v[kBatchMaxNV];
    switch (f.fin) {
    case FIN_INIT: { // rho = (r, r0) = (r, r); initial check
        for (int l = 0; l < nv; l++) {
            v[l] = (F)sums[0][l];
            SC(BS_RHO, l) = v[l];
        }
        if (bconv_check<F>(S, ns, s, nv, v, true, f.abs_tol, f.rel_tol, &S.sc[BS_ICONV * NS4]))
            bhalt(S, ns, s);
        break;
    }
    case FIN_ALPHA:
        SY(BY_ITER)++;
        for (int l = 0; l < nv; l++) {
            const F al = SC(BS_RHO, l) / (F)sums[0][l];
            SC(BS_ALPHA, l) = al;
            SC(BS_NALPHA, l) = al * (F)-1.0;
        }
        SY(BY_RESTART) = 0;
        SY(BY_RSTNOW) = 0;
        break;
    case FIN_OMEGA: {
        for (int l = 0; l < nv; l++) {
            const F s2 = (F)sums[0][l], s3 = (F)sums[1][l], s4 = (F)sums[2][l];
            const F om = s2 / s3;
            SC(BS_OMEGA, l) = om;
            SC(BS_NOMEGA, l) = om * (F)-1.0;
            const F m = SC(BS_ICONV, l);
            SC(BS_AC, l) = m != (F)0 ? m * SC(BS_ALPHA, l) : (F)0;
            SC(BS_OC, l) = m != (F)0 ? m * om : (F)0;
            F res = om * s2;
            res = (F)1.0 * s4 + (F)(-1.0) * res;
            v[l] = res;
        }
        if (bconv_check<F>(S, ns, s, nv, v, false, f.abs_tol, f.rel_tol, &S.sc[BS_ICONV * NS4]))
            SY(BY_DONE) = 1;
        break;
    }
    case FIN_BETA:
        for (int l = 0; l < nv; l++) {
            const F s1 = (F)sums[0][l];
            const F s2 = s1 / SC(BS_RHO, l);
            const F s3 = SC(BS_ALPHA, l) / SC(BS_OMEGA, l);
            const F be = s2 * s3;
            SC(BS_BETA, l) = be;
            SC(BS_RHO, l) = s1;
            const F t = SC(BS_OMEGA, l) * be;
            SC(BS_NBO, l) = t * (F)-1.0;
            if (f.theta > 0 && !S.status[q0 + l] && !SY(BY_RESTART) &&
                fabs(sums[0][l]) < f.theta * sqrt(sums[1][l] * sums[2][l]))
                SY(BY_RSTNOW) = 1;
        }
        SY(BY_REPLACED) = 0;
        if (SY(BY_DONE) || SY(BY_ITER) >= f.max_iters)
            bhalt(S, ns, s);
        break;
    case FIN_MX_INIT: { // FP64 residual norms; FP32 rho; initial check on FP64 values
        double rn[kBatchMaxNV];
        double *maxres = S.cv + BC_MAXRES * NS4;
        for (int l = 0; l < nv; l++) {
            rn[l] = sums[0][l];
            SC(BS_RHO, l) = (F)sums[0][l];
            maxres[q0 + l] = sums[0][l];
        }
        SY(BY_REPLACE) = SY(BY_REPLACED) = SY(BY_RESTART) = SY(BY_RSTNOW) = 0;
        BState<double> S64{nullptr, S.cv, S.status, S.sy, S.ndone, S.part, nullptr};
        if (bconv_check<double>(S64, ns, s, nv, rn, true, f.abs_tol, f.rel_tol, nullptr)) {
            bhalt(S, ns, s);
        } else {
            for (int l = 0; l < nv; l++)
                if (S.status[q0 + l])
                    SC(BS_ICONV, l) = (F)0;
        }
        break;
    }
    case FIN_MX_OMEGA: {
        double *res = S.cv + BC_RES * NS4, *tol = S.cv + BC_TOL * NS4,
               *maxres = S.cv + BC_MAXRES * NS4;
        int req = 0, brk = 0;
        for (int l = 0; l < nv; l++) {
            const F s2 = (F)sums[0][l], s3 = (F)sums[1][l], s4 = (F)sums[2][l];
            const F om = s2 / s3;
            SC(BS_OMEGA, l) = om;
            SC(BS_NOMEGA, l) = om * (F)-1.0;
            const F m = SC(BS_ICONV, l);
            SC(BS_AC, l) = m != (F)0 ? m * SC(BS_ALPHA, l) : (F)0;
            SC(BS_OC, l) = m != (F)0 ? m * om : (F)0;
            F r = om * s2;
            r = (F)1.0 * s4 + (F)(-1.0) * r;
            if (S.status[q0 + l])
                continue;
            if (!(isfinite(om) && om != (F)0 && isfinite(SC(BS_ALPHA, l)) && isfinite(r))) {
                SC(BS_AC, l) = SC(BS_OC, l) = (F)0;
                brk = 1;
                continue;
            }
            const double rr = (double)r;
            res[q0 + l] = rr;
            if (rr <= tol[q0 + l] * tol[q0 + l] || rr <= f.delta2 * maxres[q0 + l])
                req = 1;
        }
        if (brk) {
            SY(BY_RESTART) = 1;
            req = 1;
        }
        if (req)
            SY(BY_REPLACE) = 1;
        break;
    }
    case FIN_MX_REPL:   // FP64 true residual: convergence decided here
    case FIN_MX_RST: {  // the same after a restart (r0 = p = r)
        double *res = S.cv + BC_RES * NS4, *tol = S.cv + BC_TOL * NS4,
               *maxres = S.cv + BC_MAXRES * NS4;
        SY(BY_SUB)++;
        int flag = 1, bad = 0, rst = 0;
        for (int l = 0; l < nv; l++) {
            if (f.fin == FIN_MX_RST)
                SC(BS_RHO, l) = (F)sums[0][l];
            if (S.status[q0 + l])
                continue;
            // recurrence 10x off the true residual, or no progress since the last
            // check: restart the recurrence from the true residual
            if (f.fin == FIN_MX_REPL &&
                ((f.gap2 > 0 && sums[0][l] > f.gap2 * res[q0 + l]) || sums[0][l] > maxres[q0 + l]))
                rst = 1;
            res[q0 + l] = sums[0][l];
            maxres[q0 + l] = sums[0][l];
            if (sqrt(sums[0][l]) <= tol[q0 + l])
                S.status[q0 + l] = 1;
            else
                flag = 0;
            if (sums[0][l] != sums[0][l])
                bad = 1;
        }
        if (flag || bad) {
            SY(BY_DONE) = 1;
            if (f.fin == FIN_MX_RST)
                bhalt(S, ns, s);
        } else {
            for (int l = 0; l < nv; l++)
                if (S.status[q0 + l])
                    SC(BS_ICONV, l) = (F)0;
        }
        if (f.fin == FIN_MX_REPL) {
            if (rst)
                SY(BY_RESTART) = 1;
            SY(BY_REPLACE) = 0;
            SY(BY_REPLACED) = 1;
        }
        break;
    }
    default:
        break;
    }
}

// State reset: scalars 0, masks 1, relax factors, flags 0 (a kernel, so the reset can
// sit next to captured work without memset nodes).
template <typename F>
__global__ void kb_state_init(BState<F> S, int ns, F relax) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    const int NS4 = ns * kBatchMaxNV;
    if (q < NS4) {
        for (int f = 0; f < BS_NF; f++)
            S.sc[f * NS4 + q] = (F)0;
        S.sc[BS_ICONV * NS4 + q] = (F)1;
        S.sc[BS_RELAX * NS4 + q] = relax;
        for (int f = 0; f < BC_NF; f++)
            S.cv[f * NS4 + q] = 0.0;
        S.status[q] = 0;
    }
    if (q < ns)
        for (int f = 0; f < BY_NF; f++)
            S.sy[f * ns + q] = 0;
    if (q == 0)
        *S.ndone = 0;
}

// Per-system sums of the tile partials into out[(s*4 + l)*nd + k].
__global__ void kb_collect(BLayout L, const double *part, int nd, double *out) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= L.nsys)
        return;
    for (int k = 0; k < nd; k++)
        for (int l = 0; l < kBatchMaxNV; l++) {
            double v = 0;
            for (int t = L.stile[s]; t < L.stile[s + 1]; t++)
                v += part[(int64_t)t * PSTRIDE + k * kBatchMaxNV + l];
            out[((int64_t)s * kBatchMaxNV + l) * nd + k] = v;
        }
}

} // namespace

// ---------------------------------------------------------------------------
// Batch (layout and matrix in HBM)
// ---------------------------------------------------------------------------
void Batch::build(int ns, const int *sys_rows, const int *nvs, const int *r, const int *c,
                  const double *v) {
    XAMG_REQUIRE(ns > 0, "batch: no systems");
    nsys = ns;
    h_srow.resize(ns + 1);
    h_nv.assign(nvs, nvs + ns);
    h_voff.resize(ns + 1);
    XAMG_REQUIRE(sys_rows[0] == 0, "batch: system rows must start at 0");
    h_voff[0] = 0;
    for (int s = 0; s < ns; s++) {
        XAMG_REQUIRE(sys_rows[s + 1] > sys_rows[s], "batch: every system needs at least one row");
        XAMG_REQUIRE(nvs[s] >= 1 && nvs[s] <= kBatchMaxNV, "batch: nv per system must be 1..4");
        h_srow[s] = (uint32_t)sys_rows[s];
        h_voff[s + 1] = h_voff[s] + (int64_t)(sys_rows[s + 1] - sys_rows[s]) * nvs[s];
    }
    h_srow[ns] = (uint32_t)sys_rows[ns];
    nrows = sys_rows[ns];
    nelem = h_voff[ns];
    nnz = r[nrows];
    XAMG_REQUIRE(r[0] == 0, "batch: row pointers must start at 0");
    hrow.resize(nrows + 1);
    hcol.resize(nnz);
    hval.assign(v, v + nnz);
    max_row_nnz = 0;
    for (int64_t i = 0; i <= nrows; i++) {
        XAMG_REQUIRE(i == 0 || r[i] >= r[i - 1], "batch: row pointers must be non-decreasing");
        hrow[i] = (uint32_t)r[i];
        if (i > 0)
            max_row_nnz = std::max(max_row_nnz, r[i] - r[i - 1]);
    }
    for (int s = 0; s < ns; s++)
        for (int64_t i = sys_rows[s]; i < sys_rows[s + 1]; i++)
            for (int j = r[i]; j < r[i + 1]; j++) {
                XAMG_REQUIRE(c[j] >= sys_rows[s] && c[j] < sys_rows[s + 1],
                             "batch: a column index leaves its system's block");
                hcol[j] = (uint32_t)c[j];
            }
    auto up = [](DBuf &d, const void *p, size_t bytes) {
        d.alloc(bytes);
        if (bytes)
            XAMG_HIP_CHECK(hipMemcpy(d.p, p, bytes, hipMemcpyHostToDevice));
    };
    up(srow, h_srow.data(), sizeof(uint32_t) * h_srow.size());
    up(nv, h_nv.data(), sizeof(int) * h_nv.size());
    up(voff, h_voff.data(), sizeof(int64_t) * h_voff.size());
    up(row, hrow.data(), sizeof(uint32_t) * hrow.size());
    up(col, hcol.data(), sizeof(uint32_t) * hcol.size());
    up(val64, hval.data(), sizeof(double) * hval.size());
    val32.release();
    dinv64.release();
    dinv32.release();
    tile_rows = 0;
    set_tiles(batch_cfg().tile_rows);
}

void Batch::set_tiles(int rows) {
    XAMG_REQUIRE(rows >= 1, "batch: tile rows must be positive");
    if (rows == tile_rows)
        return;
    tile_rows = rows;
    std::vector<int> ts;
    std::vector<uint32_t> tr;
    h_stile.assign(nsys + 1, 0);
    for (int s = 0; s < nsys; s++) {
        h_stile[s] = (int)ts.size();
        for (uint32_t i = h_srow[s]; i < h_srow[s + 1]; i += (uint32_t)rows) {
            ts.push_back(s);
            tr.push_back(i);
        }
    }
    h_stile[nsys] = (int)ts.size();
    tr.push_back((uint32_t)nrows);
    ntiles = (int)ts.size();
    auto up = [](DBuf &d, const void *p, size_t bytes) {
        d.alloc(bytes);
        XAMG_HIP_CHECK(hipMemcpy(d.p, p, bytes, hipMemcpyHostToDevice));
    };
    up(tsys, ts.data(), sizeof(int) * ts.size());
    up(trow, tr.data(), sizeof(uint32_t) * tr.size());
    up(stile, h_stile.data(), sizeof(int) * h_stile.size());
    gen++; // a captured graph holds the old arrays' addresses
}

BLayout Batch::layout() const {
    BLayout L;
    L.nsys = nsys;
    L.ntiles = ntiles;
    L.srow = srow.as<uint32_t>();
    L.nv = nv.as<int>();
    L.voff = voff.as<int64_t>();
    L.tsys = tsys.as<int>();
    L.trow = trow.as<uint32_t>();
    L.stile = stile.as<int>();
    L.row = row.as<uint32_t>();
    L.col = col.as<uint32_t>();
    L.max_row_nnz = max_row_nnz;
    return L;
}

const void *Batch::values(bool fp32) {
    if (!fp32)
        return val64.p;
    if (!val32.p) {
        std::vector<float> v(std::max<int64_t>(nnz, 1));
        for (int64_t j = 0; j < nnz; j++)
            v[j] = (float)hval[j];
        val32.alloc(sizeof(float) * v.size());
        XAMG_HIP_CHECK(hipMemcpy(val32.p, v.data(), sizeof(float) * v.size(), hipMemcpyHostToDevice));
    }
    return val32.p;
}

// csr_matrix.h:323-355: 1.0 / a_ii computed in double from the values in use
template <typename T>
const T *Batch::inv_diag() {
    DBuf &d = std::is_same<T, double>::value ? dinv64 : dinv32;
    if (!d.p) {
        std::vector<T> h(std::max<int64_t>(nrows, 1), (T)0);
        for (int64_t i = 0; i < nrows; i++)
            for (uint32_t j = hrow[i]; j < hrow[i + 1]; j++)
                if (hcol[j] == (uint32_t)i)
                    h[i] = (T)(1.0 / (double)(T)hval[j]);
        d.alloc(sizeof(T) * h.size());
        XAMG_HIP_CHECK(hipMemcpy(d.p, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice));
    }
    return d.as<T>();
}
template const double *Batch::inv_diag<double>();
template const float *Batch::inv_diag<float>();

// ---------------------------------------------------------------------------
// Batched (P)BiCGStab, FP64, or FP32 recurrences with FP64 residual replacement
// ---------------------------------------------------------------------------
namespace {

template <typename V, typename T, int EPI, int ND, int U>
void launch_bspmv_u(const BLayout &L, int G, const int *halt, const BSpmvArgs<T> &a, double *part,
                    hipStream_t st) {
    if (L.ept == 2)
        hipLaunchKernelGGL((kb_spmv<V, T, EPI, ND, U, 2>), dim3(G), dim3(BT), 0, st, L, halt, a, part);
    else
        hipLaunchKernelGGL((kb_spmv<V, T, EPI, ND, U, 1>), dim3(G), dim3(BT), 0, st, L, halt, a, part);
}
int grid_of(const BLayout &L) {
    return std::max(1, std::min(L.grid, (L.ntiles + L.tps - 1) / L.tps));
}
template <typename V, typename T, int EPI, int ND>
void launch_bspmv(const BLayout &L, const int *halt, const BSpmvArgs<T> &a, double *part,
                  hipStream_t st) {
    const int G = grid_of(L);
    if (L.max_row_nnz <= 3)
        launch_bspmv_u<V, T, EPI, ND, 3>(L, G, halt, a, part, st);
    else if (L.max_row_nnz <= 5)
        launch_bspmv_u<V, T, EPI, ND, 5>(L, G, halt, a, part, st);
    else if (L.max_row_nnz <= 7)
        launch_bspmv_u<V, T, EPI, ND, 7>(L, G, halt, a, part, st);
    else
        launch_bspmv_u<V, T, EPI, ND, 9>(L, G, halt, a, part, st);
}
template <typename T, int OP, int ND>
void launch_bew(const BLayout &L, const int *halt, const BEwArgs<T> &a, double *part,
                hipStream_t st) {
    hipLaunchKernelGGL((kb_ew<T, OP, ND>), dim3(grid_of(L)), dim3(BT), 0, st, L, halt, a, part);
}
template <typename F>
void launch_bfin(const BLayout &L, const BState<F> &S, const BFinArgs &f, hipStream_t st) {
    hipLaunchKernelGGL((kb_finish<F>), dim3((L.nsys + 3) / 4), dim3(256), 0, st, L, S, f);
}

template <typename T>
class BatchKrylov : public BatchKrylovBase {
    static constexpr bool kMixed = std::is_same<T, float>::value;

  public:
    Batch &B;
    bool pre = false;
    int sweeps = 2;
    double relax = 0.99;
    int max_iters = 50;
    double abs_tol = 0, rel_tol = 0, delta2 = 1e-4, theta = 1e-7, gap2 = 100.0;
    // FP32 solver: the products v = A p^ and t = A s^ of the recurrences read the FP64
    // values and sum in FP64 (key recurrence_operator, default fp64): the recurrence then
    // tracks b - A x for the operator itself, not for its FP32 rounding, and the
    // ill-conditioned systems need 10% fewer row-iterations; the Jacobi sweeps keep the
    // FP32 values
    bool op64 = true;
    // FP32 solver: the Jacobi sweeps of two or more read the FP64 values and inverse
    // diagonal and compute in FP64, rounding each sweep's result once (key
    // preconditioner_operator = fp64)
    bool pre64 = false;
    std::vector<DBuf> vec;
    DBuf sc, cv, status, sy, ndone, part;
    int *pinned = nullptr;
    BState<T> S;
    hipGraphExec_t gexec = nullptr;
    const double *g_x = nullptr, *g_b = nullptr;
    int g_k = 0, g_tiles = 0, g_grid = 0, g_tps = 0, g_ept = 0;
    int g_gen = -1; // tile-array generation the graph's launches point into
    // launch shape: workgroups, tiles per workgroup step, elements per SpMV lane; tuned
    // for the tile-array generation t_gen
    int lc_grid = 8192, lc_tps = 1, lc_ept = 1, t_gen = -1;
    bool tuned = false;

    explicit BatchKrylov(Batch &b) : B(b) {}
    ~BatchKrylov() override {
        if (gexec)
            (void)hipGraphExecDestroy(gexec);
        if (pinned)
            (void)hipHostFree(pinned);
    }
    T *V(int k) { return vec[k].as<T>(); }

    void setup(const GlobalParams &gp) {
        const bool meta = gp.has("meta_solver");
        const ParamList &sp = gp.get("solver");
        const std::string m = sp.get_string("method");
        XAMG_REQUIRE(m == "BiCGStab" || m == "PBiCGStab" || m == "MergedBiCGStab" ||
                         m == "MergedPBiCGStab",
                     "batch: the solver must be BiCGStab or PBiCGStab");
        pre = (m == "PBiCGStab" || m == "MergedPBiCGStab");
        if (pre) {
            XAMG_REQUIRE(gp.has("preconditioner"), "batch: PBiCGStab needs a preconditioner block");
            const ParamList &pp = gp.get("preconditioner");
            XAMG_REQUIRE(pp.get_string("method") == "Jacobi",
                         "batch: the preconditioner must be Jacobi");
            sweeps = (int)pp.get_int("max_iters");
            relax = pp.get_float("relax_factor");
            XAMG_REQUIRE(sweeps >= 1, "batch: Jacobi needs at least one sweep");
        }
        const ParamList &top = meta ? gp.get("meta_solver") : sp;
        abs_tol = top.get_float("abs_tolerance");
        rel_tol = top.get_float("rel_tolerance");
        max_iters = (int)sp.get_int("max_iters");
        if (kMixed) {
            // replacement after each 1000x drop of the recurrence residual: with the FP64
            // recurrence operator, over three 2,000-system batches 2-3% faster than 10x or
            // 100x and 2-5% faster than 1e4x or 1e6x
            const double step = top.has("replace_step") ? top.get_float("replace_step") : 1e-3;
            delta2 = step * step;
            theta = top.has("restart_threshold") ? top.get_float("restart_threshold") : 1e-7;
            const double gap = top.has("restart_gap") ? top.get_float("restart_gap") : 10.0;
            gap2 = gap * gap;
            if (top.has("recurrence_operator")) {
                const std::string o = top.get_string("recurrence_operator");
                XAMG_REQUIRE(o == "fp64" || o == "fp32", "batch: recurrence_operator must be fp64 or fp32");
                op64 = o == "fp64";
            }
            if (top.has("preconditioner_operator")) {
                const std::string o = top.get_string("preconditioner_operator");
                XAMG_REQUIRE(o == "fp64" || o == "fp32", "batch: preconditioner_operator must be fp64 or fp32");
                pre64 = o == "fp64";
            }
        }
        const int nb = pre ? 9 : 6;
        vec.clear();
        for (int i = 0; i < nb; i++)
            vec.emplace_back(sizeof(T) * std::max<int64_t>(B.nelem, 1));
        const int NS4 = B.nsys * kBatchMaxNV;
        sc.alloc(sizeof(T) * BS_NF * NS4);
        cv.alloc(sizeof(double) * BC_NF * NS4);
        status.alloc(sizeof(int) * NS4);
        sy.alloc(sizeof(int) * BY_NF * B.nsys);
        ndone.alloc(256);
        XAMG_HIP_CHECK(hipHostMalloc((void **)&pinned, 2 * sizeof(int), hipHostMallocDefault));
        B.values(kMixed);
        B.inv_diag<T>();
        if (kMixed && pre64)
            B.inv_diag<double>();
        S.sc = sc.as<T>();
        S.cv = cv.as<double>();
        S.status = status.as<int>();
        S.sy = sy.as<int>();
        S.ndone = ndone.as<int>();
        prepare(Runtime::get().stream);
    }

    BLayout lay() const {
        BLayout L = B.layout();
        L.grid = lc_grid;
        L.tps = lc_tps;
        L.ept = lc_ept;
        return L;
    }

    // Tile arrays for the configured tile size, the partial-sum buffer, and the launch
    // shape: the configured one, or the autotuner's for this tile-array generation.
    void prepare(hipStream_t st) {
        const BatchCfg &cfg = batch_cfg();
        B.set_tiles(cfg.tile_rows);
        if (part.bytes < sizeof(double) * PSTRIDE * (size_t)B.ntiles)
            part.alloc(sizeof(double) * PSTRIDE * (size_t)B.ntiles);
        S.part = part.as<double>();
        if (!cfg.tune) {
            lc_grid = std::max(1, cfg.grid);
            lc_tps = std::max(1, cfg.tps);
            lc_ept = cfg.ept == 2 ? 2 : 1;
            tuned = false;
        } else if (!tuned || t_gen != B.gen) {
            tune(st);
        }
    }

    // Times the launch shapes on the kernels that dominate an iteration while every system
    // runs (the three-dot product t = A s^, the folded Jacobi sweeps, the r update with its
    // dots) and keeps the fastest.  None of the shapes changes a result: tiles fix the dot
    // order, and a lane's elements are summed in element order whatever it keeps in flight.
    void tune(hipStream_t st) {
        const int NS4 = B.nsys * kBatchMaxNV;
        hipLaunchKernelGGL((kb_state_init<T>), dim3((NS4 + 255) / 256), dim3(256), 0, st, S,
                           B.nsys, (T)relax);
        BLayout L = lay();
        for (size_t k = 0; k < vec.size(); k++) {
            BEwArgs<T> a;
            a.z = V((int)k);
            a.x = V((int)k);
            a.c0 = (T)1;
            launch_bew<T, EW_FILL, 0>(L, nullptr, a, nullptr, st);
        }
        double *pt = part.as<double>();
        auto round = [&](const BLayout &Lc) {
            BSpmvArgs<T> m;
            m.x = V(4);
            m.y = V(5);
            m.w = V(4);
            if (op64) {
                m.val = B.values(false);
                launch_bspmv<double, T, EPI_AXY, 3>(Lc, nullptr, m, pt, st);
            } else {
                m.val = B.values(kMixed);
                launch_bspmv<T, T, EPI_AXY, 3>(Lc, nullptr, m, pt, st);
            }
            if (pre) {
                BSpmvArgs<T> j;
                j.x = V(2);
                j.b = V(2);
                j.y = V(6);
                j.d = B.inv_diag<T>();
                j.wcoef = scf(BS_RELAX);
                if (kMixed && pre64) {
                    j.val = B.values(false);
                    j.d64 = B.inv_diag<double>();
                    launch_bspmv<double, T, BEPI_JACOBI2, 0>(Lc, nullptr, j, nullptr, st);
                } else {
                    j.val = B.values(kMixed);
                    launch_bspmv<T, T, BEPI_JACOBI2, 0>(Lc, nullptr, j, nullptr, st);
                }
            }
            BEwArgs<T> r;
            r.z = V(1);
            r.x = V(4);
            r.y = V(5);
            r.cb = scf(BS_NOMEGA);
            r.w = V(0);
            launch_bew<T, EW_AXPBY_Z, 3>(Lc, nullptr, r, pt, st);
        };
        hipEvent_t e0, e1;
        XAMG_HIP_CHECK(hipEventCreate(&e0));
        XAMG_HIP_CHECK(hipEventCreate(&e1));
        float best = 0;
        for (int g : {2048, 4096, 8192, 16384, 32768})
            for (int tp : {1, 2, 4})
                for (int ep : {1, 2}) {
                    BLayout Lc = L;
                    Lc.grid = g;
                    Lc.tps = tp;
                    Lc.ept = ep;
                    round(Lc);
                    XAMG_HIP_CHECK(hipEventRecord(e0, st));
                    for (int r = 0; r < 3; r++)
                        round(Lc);
                    XAMG_HIP_CHECK(hipEventRecord(e1, st));
                    XAMG_HIP_CHECK(hipEventSynchronize(e1));
                    float ms = 0;
                    XAMG_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
                    if (best == 0 || ms < best) {
                        best = ms;
                        lc_grid = g;
                        lc_tps = tp;
                        lc_ept = ep;
                    }
                }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        tuned = true;
        t_gen = B.gen;
    }

    void launch(int *grid, int *tps, int *ept, int *was_tuned) const override {
        if (grid)
            *grid = lc_grid;
        if (tps)
            *tps = lc_tps;
        if (ept)
            *ept = lc_ept;
        if (was_tuned)
            *was_tuned = tuned ? 1 : 0;
    }

    const int *halt() const { return S.sy + BY_HALT * B.nsys; }
    const int *syf(int f) const { return S.sy + f * B.nsys; }
    T *scf(int f) const { return S.sc + f * B.nsys * kBatchMaxNV; }

    BFinArgs fin(int op, int nd, const int *gate = nullptr) const {
        BFinArgs f;
        f.fin = op;
        f.nd = nd;
        f.abs_tol = abs_tol;
        f.rel_tol = rel_tol;
        f.delta2 = delta2;
        f.gap2 = gap2;
        f.max_iters = max_iters;
        f.gate = gate;
        return f;
    }

    // z = M^{-1} q: Jacobi sweeps from zero (jacobi.inl no-check mode, out of place)
    void jacobi(const T *q, T *out, T *tmp, hipStream_t st, const BLayout &L) {
        const T *d = B.inv_diag<T>();
        const T *src = nullptr;
        if (sweeps == 1) {
            BEwArgs<T> a;
            a.z = out;
            a.x = q;
            a.d = d;
            a.ca = scf(BS_RELAX);
            launch_bew<T, EW_JACOBI0, 0>(L, halt(), a, nullptr, st);
            return;
        }
        { // sweeps 1 and 2 in one launch
            T *dst = ((sweeps - 2) % 2 == 0) ? out : tmp;
            BSpmvArgs<T> a;
            a.x = q;
            a.y = dst;
            a.b = q;
            a.d = d;
            a.wcoef = scf(BS_RELAX);
            if (kMixed && pre64) {
                a.val = B.values(false);
                a.d64 = B.inv_diag<double>();
                launch_bspmv<double, T, BEPI_JACOBI2, 0>(L, halt(), a, nullptr, st);
            } else {
                a.val = B.values(kMixed);
                launch_bspmv<T, T, BEPI_JACOBI2, 0>(L, halt(), a, nullptr, st);
            }
            src = dst;
        }
        for (int k = 3; k <= sweeps; k++) {
            T *dst = ((sweeps - k) % 2 == 0) ? out : tmp;
            BSpmvArgs<T> a;
            a.val = B.values(kMixed);
            a.x = src;
            a.xin = src;
            a.y = dst;
            a.b = q;
            a.d = d;
            a.wcoef = scf(BS_RELAX);
            if (kMixed && pre64) {
                a.val = B.values(false);
                a.d64 = B.inv_diag<double>();
                launch_bspmv<double, T, EPI_JACOBI, 0>(L, halt(), a, nullptr, st);
            } else {
                launch_bspmv<T, T, EPI_JACOBI, 0>(L, halt(), a, nullptr, st);
            }
            src = dst;
        }
    }

    void init(double *x, const double *b, hipStream_t st, const BLayout &L) {
        const int NS4 = B.nsys * kBatchMaxNV;
        hipLaunchKernelGGL((kb_state_init<T>), dim3((NS4 + 255) / 256), dim3(256), 0, st, S,
                           B.nsys, (T)relax);
        T *r0 = V(0), *r = V(1), *p = V(2);
        if constexpr (kMixed) {
            BSpmvArgs<double> a;
            a.val = B.values(false);
            a.x = x;
            a.b = b;
            a.yo = r0;
            a.yo2 = r;
            a.yo3 = p;
            launch_bspmv<double, double, EPI_RESID_MX, 1>(L, nullptr, a, part.as<double>(), st);
            launch_bfin<T>(L, S, fin(FIN_MX_INIT, 1), st);
        } else {
            BSpmvArgs<double> a;
            a.val = B.values(false);
            a.x = x;
            a.b = b;
            a.y = r0;
            a.y2 = r;
            a.y3 = p;
            launch_bspmv<double, double, EPI_RESID, 1>(L, nullptr, a, part.as<double>(), st);
            launch_bfin<T>(L, S, fin(FIN_INIT, 1), st);
        }
    }

    void iteration(double *x, const double *b, hipStream_t st, const BLayout &L) {
        T *r0 = V(0), *r = V(1), *p = V(2), *v = V(3), *s = V(4), *t = V(5);
        T *p_ = pre ? V(6) : p, *s_ = pre ? V(7) : s, *jt = pre ? V(8) : nullptr;
        double *pt = part.as<double>();
        const void *val = B.values(kMixed);
        if (pre)
            jacobi(p, p_, jt, st, L);
        { // v = A p_, (v, r0) -> alpha
            BSpmvArgs<T> a;
            a.val = val;
            a.x = p_;
            a.y = v;
            a.w = r0;
            if (op64) {
                a.val = B.values(false);
                launch_bspmv<double, T, EPI_AXY, 1>(L, halt(), a, pt, st);
            } else {
                launch_bspmv<T, T, EPI_AXY, 1>(L, halt(), a, pt, st);
            }
            launch_bfin<T>(L, S, fin(FIN_ALPHA, 1), st);
        }
        { // s = r - alpha v
            BEwArgs<T> a;
            a.z = s;
            a.x = r;
            a.y = v;
            a.cb = scf(BS_NALPHA);
            launch_bew<T, EW_AXPBY_Z, 0>(L, halt(), a, nullptr, st);
        }
        if (pre)
            jacobi(s, s_, jt, st, L);
        { // t = A s_, (t, s), (t, t), (s, s) -> omega, residual, check
            BSpmvArgs<T> a;
            a.val = val;
            a.x = s_;
            a.y = t;
            a.w = s;
            if (op64) {
                a.val = B.values(false);
                launch_bspmv<double, T, EPI_AXY, 3>(L, halt(), a, pt, st);
            } else {
                launch_bspmv<T, T, EPI_AXY, 3>(L, halt(), a, pt, st);
            }
            launch_bfin<T>(L, S, fin(kMixed ? FIN_MX_OMEGA : FIN_OMEGA, 3), st);
        }
        if constexpr (kMixed) {
            { // x(FP64) += alpha_c p_ + omega_c s_
                BEwArgs<double> a;
                a.z = x;
                a.xo = p_;
                a.yo = s_;
                a.ao = scf(BS_AC);
                a.bo = scf(BS_OC);
                a.freeze_zero = 1;
                launch_bew<double, EW_MX_XUPD, 0>(L, halt(), a, nullptr, st);
            }
            { // FP64 residual replacement for the systems that asked for it
                BSpmvArgs<double> a;
                a.val = B.values(false);
                a.x = x;
                a.b = b;
                a.yo = r;
                a.gate = syf(BY_REPLACE);
                launch_bspmv<double, double, EPI_RESID_MX, 1>(L, halt(), a, pt, st);
                launch_bfin<T>(L, S, fin(FIN_MX_REPL, 1, syf(BY_REPLACE)), st);
            }
        } else {
            BEwArgs<T> a;
            a.z = (T *)x;
            a.x = p_;
            a.ca = scf(BS_AC);
            a.y = s_;
            a.cb = scf(BS_OC);
            a.freeze_zero = 1;
            launch_bew<T, EW_AXPBYPCZ, 0>(L, halt(), a, nullptr, st);
        }
        { // r = s - omega t, (r, r0) -> beta
            BEwArgs<T> a;
            a.z = r;
            a.x = s;
            a.y = t;
            a.cb = scf(BS_NOMEGA);
            a.w = r0;
            if constexpr (kMixed) {
                a.keep = syf(BY_REPLACED);
                a.y2 = r0; // restart: r0 = r, rho = (r, r)
                a.copy = syf(BY_RESTART);
                launch_bew<T, EW_AXPBY_Z, 3>(L, halt(), a, pt, st);
                BFinArgs fb = fin(FIN_BETA, 3);
                fb.theta = theta;
                launch_bfin<T>(L, S, fb, st);
            } else {
                launch_bew<T, EW_AXPBY_Z, 1>(L, halt(), a, pt, st);
                launch_bfin<T>(L, S, fin(FIN_BETA, 1), st);
            }
        }
        if constexpr (kMixed) { // (r, r0) lost: FP64 residual, r0 = p = r
            BSpmvArgs<double> a;
            a.val = B.values(false);
            a.x = x;
            a.b = b;
            a.yo = r;
            a.yo2 = r0;
            a.yo3 = p;
            a.gate = syf(BY_RSTNOW);
            launch_bspmv<double, double, EPI_RESID_MX, 1>(L, halt(), a, pt, st);
            launch_bfin<T>(L, S, fin(FIN_MX_RST, 1, syf(BY_RSTNOW)), st);
        }
        { // p = r + beta p - beta omega v
            BEwArgs<T> a;
            a.z = p;
            a.x = r;
            a.y = v;
            a.cb = scf(BS_NBO);
            a.cc = scf(BS_BETA);
            a.copy = kMixed ? syf(BY_RESTART) : nullptr; // p = r on a restart
            a.keep = kMixed ? syf(BY_RSTNOW) : nullptr;  // p = r already
            launch_bew<T, EW_AXPBYPCZ, 0>(L, halt(), a, nullptr, st);
        }
    }

    void solve(double *x, const double *b) override {
        Runtime &rt = Runtime::get();
        hipStream_t st = rt.stream;
        const BatchCfg &cfg = batch_cfg();
        prepare(st);
        const BLayout L = lay();
        const int K = std::max(1, cfg.batch);
        hipEvent_t e0, e1;
        XAMG_HIP_CHECK(hipEventCreate(&e0));
        XAMG_HIP_CHECK(hipEventCreate(&e1));
        XAMG_HIP_CHECK(hipEventRecord(e0, st));
        init(x, b, st, L);
        // one graph of K iterations, rebuilt when the vectors or the launch shape change
        if (cfg.use_graph &&
            (!gexec || g_x != x || g_b != b || g_k != K || g_tiles != B.ntiles ||
             g_grid != L.grid || g_tps != L.tps || g_ept != L.ept || g_gen != B.gen)) {
            if (gexec) {
                (void)hipGraphExecDestroy(gexec);
                gexec = nullptr;
            }
            hipGraph_t g;
            XAMG_HIP_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            for (int k = 0; k < K; k++)
                iteration(x, b, st, L);
            XAMG_HIP_CHECK(hipStreamEndCapture(st, &g));
            XAMG_HIP_CHECK(hipGraphInstantiate(&gexec, g, nullptr, nullptr, 0));
            (void)hipGraphDestroy(g);
            g_x = x;
            g_b = b;
            g_k = K;
            g_tiles = B.ntiles;
            g_grid = L.grid;
            g_tps = L.tps;
            g_ept = L.ept;
            g_gen = B.gen;
        }
        hipEvent_t ev[2];
        XAMG_HIP_CHECK(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
        XAMG_HIP_CHECK(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
        // batches of K iterations; the counter of finished systems is read back after
        // each batch, one batch behind (the next batch is queued before waiting)
        const int nbatch = (max_iters + K - 1) / K;
        int issued = 0, waited = 0;
        bool finished = false;
        auto enqueue = [&]() {
            if (cfg.use_graph)
                XAMG_HIP_CHECK(hipGraphLaunch(gexec, st));
            else
                for (int k = 0; k < K; k++)
                    iteration(x, b, st, L);
            XAMG_HIP_CHECK(hipMemcpyAsync(&pinned[issued % 2], S.ndone, sizeof(int),
                                          hipMemcpyDeviceToHost, st));
            XAMG_HIP_CHECK(hipEventRecord(ev[issued % 2], st));
            issued++;
        };
        while (!finished && waited < nbatch) {
            while (issued < nbatch && issued < waited + 2)
                enqueue();
            XAMG_HIP_CHECK(hipEventSynchronize(ev[waited % 2]));
            if (pinned[waited % 2] >= B.nsys)
                finished = true;
            waited++;
        }
        XAMG_HIP_CHECK(hipEventRecord(e1, st));
        XAMG_HIP_CHECK(hipEventSynchronize(e1));
        float ms = 0;
        XAMG_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
        B.last_ms = ms;
        (void)hipEventDestroy(ev[0]);
        (void)hipEventDestroy(ev[1]);
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    }

    void info(int *iters, int *sub_iters, int *converged) override {
        std::vector<int> h(BY_NF * B.nsys), stt(B.nsys * kBatchMaxNV);
        XAMG_HIP_CHECK(hipMemcpy(h.data(), S.sy, sizeof(int) * h.size(), hipMemcpyDeviceToHost));
        XAMG_HIP_CHECK(hipMemcpy(stt.data(), S.status, sizeof(int) * stt.size(), hipMemcpyDeviceToHost));
        int q = 0;
        for (int s = 0; s < B.nsys; s++) {
            if (iters)
                iters[s] = h[BY_ITER * B.nsys + s];
            if (sub_iters)
                sub_iters[s] = h[BY_SUB * B.nsys + s];
            if (converged)
                for (int l = 0; l < B.h_nv[s]; l++)
                    converged[q++] = stt[s * kBatchMaxNV + l];
        }
    }
};

} // namespace

std::unique_ptr<BatchKrylovBase> make_batch_solver(Batch &B, const GlobalParams &gp0) {
    GlobalParams gp = gp0;
    gp.set_defaults();
    XAMG_REQUIRE(gp.has("solver"), "batch: the parameters need a solver block");
    bool mixed = false;
    if (gp.has("meta_solver")) {
        const ParamList &m = gp.get("meta_solver");
        XAMG_REQUIRE(m.get_string("method") == "IterativeRefinement",
                     "batch: the meta solver must be IterativeRefinement");
        mixed = m.has("subsolver_reduced_precision") && m.get_bool("subsolver_reduced_precision");
        XAMG_REQUIRE(mixed, "batch: IterativeRefinement runs with subsolver_reduced_precision");
    }
    if (mixed) {
        auto s = std::make_unique<BatchKrylov<float>>(B);
        s->setup(gp);
        return s;
    }
    auto s = std::make_unique<BatchKrylov<double>>(B);
    s->setup(gp);
    return s;
}

void batch_residuals(Batch &B, const double *x, const double *b, double *relres) {
    hipStream_t st = Runtime::get().stream;
    B.set_tiles(batch_cfg().tile_rows);
    const BLayout L = B.layout();
    DBuf y(sizeof(double) * std::max<int64_t>(B.nelem, 1));
    DBuf part(sizeof(double) * PSTRIDE * (size_t)B.ntiles);
    DBuf sums(sizeof(double) * 2 * B.nsys * kBatchMaxNV);
    {
        BSpmvArgs<double> a;
        a.val = B.values(false);
        a.x = x;
        a.b = b;
        a.y = y.as<double>();
        launch_bspmv<double, double, EPI_RESID, 1>(L, nullptr, a, part.as<double>(), st);
        hipLaunchKernelGGL(kb_collect, dim3((B.nsys + 63) / 64), dim3(64), 0, st, L,
                           part.as<const double>(), 1, sums.as<double>());
    }
    {
        BEwArgs<double> a;
        a.z = y.as<double>();
        a.x = b;
        launch_bew<double, EW_COPY3, 1>(L, nullptr, a, part.as<double>(), st);
        hipLaunchKernelGGL(kb_collect, dim3((B.nsys + 63) / 64), dim3(64), 0, st, L,
                           part.as<const double>(), 1, sums.as<double>() + B.nsys * kBatchMaxNV);
    }
    std::vector<double> h(2 * B.nsys * kBatchMaxNV);
    XAMG_HIP_CHECK(hipMemcpyAsync(h.data(), sums.p, sizeof(double) * h.size(), hipMemcpyDeviceToHost, st));
    XAMG_HIP_CHECK(hipStreamSynchronize(st));
    int q = 0;
    for (int s = 0; s < B.nsys; s++)
        for (int l = 0; l < B.h_nv[s]; l++) {
            const double rr = h[s * kBatchMaxNV + l], bb = h[B.nsys * kBatchMaxNV + s * kBatchMaxNV + l];
            relres[q++] = bb > 0 ? std::sqrt(rr / bb) : std::sqrt(rr);
        }
}

} // namespace xamg
