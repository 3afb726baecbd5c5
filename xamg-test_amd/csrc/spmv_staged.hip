// The LDS-staged instantiations of the SpMV family (STAGE = 1), compiled apart from
// kernels.hip so the sets build in parallel.
#include "spmv_kernel.hpp"

namespace xamg {

template <typename V, typename T>
void launch_spmv_staged(int epi, int ndots, int nv, const SpmvArgs<T> &a, const RedArgs &r,
                        const FinArgs<T> &f, hipStream_t s) {
    if (a.cfg_mode == 2)
        launch_spmv_stage2<V, T>(epi, ndots, nv, a, r, f, s);
    else
        launch_spmv_mode<V, T, 1>(epi, ndots, nv, a, r, f, s);
}

template void launch_spmv_staged<double, double>(int, int, int, const SpmvArgs<double> &,
                                                 const RedArgs &, const FinArgs<double> &, hipStream_t);
template void launch_spmv_staged<float, float>(int, int, int, const SpmvArgs<float> &, const RedArgs &,
                                               const FinArgs<float> &, hipStream_t);
template void launch_spmv_staged<__hip_bfloat16, float>(int, int, int, const SpmvArgs<float> &,
                                                        const RedArgs &, const FinArgs<float> &,
                                                        hipStream_t);

} // namespace xamg
