// batch_vplus_mixed.hip - the mixed-precision producers of batched VegasFlowPlus
// integrals (batch_vplus.hpp); the two precisions are separate translation
// units so that they compile in parallel.
#include "batch_vplus.hpp"

namespace mv {

int launch_batch_vplus_mixed(const mv_batch_vplus_item* items, const mv_batch_tile* tiles, int n_tiles, int family,
                             int min_dim, int max_dim, const int32_t* draws, const double* grids,
                             const int32_t* n_ev, const int32_t* offsets, double* ress, double* ress2, double* rows,
                             int blocks, hipStream_t stream) {
  return launch_batch_vplus_precision<true>(items, tiles, n_tiles, family, min_dim, max_dim, draws, grids, n_ev,
                                              offsets, ress, ress2, rows, blocks, stream);
}

}  // namespace mv
