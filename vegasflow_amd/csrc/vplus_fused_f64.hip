// vplus_fused_f64.hip - the FP64 instances of the fused VEGAS+
// producer (vplus_fused.hpp).
#include "vplus_fused.hpp"

namespace mv {

int launch_vplus_fused_f64(int family, bool train, const StepK& k, const StrataK& st,
                         const mv_integrand& fn, const double* grid, double* ress, double* ress2,
                         const RowsK& R, hipStream_t s) {
  return launch_vplus_precision<false>(family, train, k, st, fn, grid, ress, ress2, R, s);
}

}  // namespace mv
