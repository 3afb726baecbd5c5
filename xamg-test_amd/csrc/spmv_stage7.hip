// The SpMV instantiations of staging modes 7 and 8 (modes 2 and 4 with 16-byte staging
// loads, twice the staging capacity and twice the nonzero chunk), compiled apart from
// kernels.hip so the sets build in parallel.
#include "spmv_kernel.hpp"

namespace xamg {

template <typename V, typename T>
void launch_spmv_stage7(int epi, int ndots, int nv, const SpmvArgs<T> &a, const RedArgs &r,
                        const FinArgs<T> &f, hipStream_t s) {
    launch_spmv_mode<V, T, 7>(epi, ndots, nv, a, r, f, s);
}

template <typename V, typename T>
void launch_spmv_stage8(int epi, int ndots, int nv, const SpmvArgs<T> &a, const RedArgs &r,
                        const FinArgs<T> &f, hipStream_t s) {
    launch_spmv_mode<V, T, 8>(epi, ndots, nv, a, r, f, s);
}

#define XAMG_STAGE78_INST(V, T)                                                                    \
    template void launch_spmv_stage7<V, T>(int, int, int, const SpmvArgs<T> &, const RedArgs &,    \
                                           const FinArgs<T> &, hipStream_t);                      \
    template void launch_spmv_stage8<V, T>(int, int, int, const SpmvArgs<T> &, const RedArgs &,    \
                                           const FinArgs<T> &, hipStream_t);
XAMG_STAGE78_INST(double, double)
XAMG_STAGE78_INST(float, float)
XAMG_STAGE78_INST(__hip_bfloat16, float)
#undef XAMG_STAGE78_INST

} // namespace xamg
