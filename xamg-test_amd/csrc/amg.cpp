// Host classical AMG setup (placeholder until the Ruge-Stueben setup lands).
#include "solvers.hpp"

namespace xamg {

std::shared_ptr<MgHierarchy> build_amg(DevMatrix *A, const ParamList &mgprm) {
    (void)A;
    (void)mgprm;
    throw std::runtime_error("xamg: MultiGrid hierarchy setup is not available in this build");
}

} // namespace xamg
