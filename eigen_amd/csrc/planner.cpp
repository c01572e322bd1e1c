// planner.cpp -- chooses the tile / split-K configuration of a product.
//
// Replaces Eigen's blocking-size choice (computeProductBlockingSizes and
// evaluateProductBlockingSizesHeuristic,
// Eigen/src/Core/products/GeneralBlockPanelKernel.h:201-452), which derives
// kc/mc/nc from the host's cache sizes.  Here the choice is one of the 24
// candidate configurations and comes from the current tuning mode.
#include "internal.hpp"

namespace eh {

eh_status plan_config(int dtype, int opa, int opb, int64_t m, int64_t n, int64_t k, int* config) {
  (void)opa;
  (void)opb;
  *config = heuristic_config(dtype, m, n, k);
  return EH_SUCCESS;
}

}  // namespace eh
