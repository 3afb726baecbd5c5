// The SpMV instantiations of staging mode 5 (mode 3 reading 16-bit column offsets),
// compiled apart from kernels.hip so the sets build in parallel.
#include "spmv_kernel.hpp"

namespace xamg {

template <typename V, typename T>
void launch_spmv_stage5(int epi, int ndots, int nv, const SpmvArgs<T> &a, const RedArgs &r,
                        const FinArgs<T> &f, hipStream_t s) {
    launch_spmv_mode<V, T, 5>(epi, ndots, nv, a, r, f, s);
}

template void launch_spmv_stage5<double, double>(int, int, int, const SpmvArgs<double> &,
                                                 const RedArgs &, const FinArgs<double> &, hipStream_t);
template void launch_spmv_stage5<float, float>(int, int, int, const SpmvArgs<float> &, const RedArgs &,
                                               const FinArgs<float> &, hipStream_t);
template void launch_spmv_stage5<__hip_bfloat16, float>(int, int, int, const SpmvArgs<float> &,
                                                        const RedArgs &, const FinArgs<float> &,
                                                        hipStream_t);

} // namespace xamg
