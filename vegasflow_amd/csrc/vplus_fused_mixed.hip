// vplus_fused_mixed.hip - the mixed-precision instances of the fused VEGAS+
// producer (vplus_fused.hpp).
#include "vplus_fused.hpp"

namespace mv {

int launch_vplus_fused_mixed(int family, bool train, const StepK& k, const StrataK& st,
                         const mv_integrand& fn, const double* grid, double* ress, double* ress2,
                         const RowsK& R, hipStream_t s) {
  return launch_vplus_precision<true>(family, train, k, st, fn, grid, ress, ress2, R, s);
}

}  // namespace mv
