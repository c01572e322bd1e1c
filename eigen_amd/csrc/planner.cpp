// planner.cpp -- chooses the tile / split-K configuration of a product.
//
// Replaces Eigen's blocking-size choice (computeProductBlockingSizes and
// evaluateProductBlockingSizesHeuristic,
// Eigen/src/Core/products/GeneralBlockPanelKernel.h:201-452), which derives
// kc/mc/nc from the host's cache sizes.  Here the choice is one of the 24
// candidate configurations and comes from the current tuning mode:
//   HEURISTIC  closed-form rule (heuristic_config)
//   MLP        the trained selector, evaluated on the host from its
//              parameters (falls back to the heuristic when none is set)
//   TABLE      the exhaustive timing table; an unseen shape is tuned on
//              first use (on scratch buffers) and remembered
#include "internal.hpp"

namespace eh {

int mlp_plan(int dtype, int opa, int opb, int64_t m, int64_t n, int64_t k);
int table_plan(int dtype, int opa, int opb, int64_t m, int64_t n, int64_t k);

eh_status plan_config(int dtype, int opa, int opb, int64_t m, int64_t n, int64_t k, int* config) {
  int c = -1;
  switch (tune_mode()) {
    case EH_TUNE_MLP:
      c = mlp_plan(dtype, opa, opb, m, n, k);
      break;
    case EH_TUNE_TABLE:
      c = table_plan(dtype, opa, opb, m, n, k);
      break;
    default:
      break;
  }
  *config = (c >= 0 && c < kNumConfigs) ? c : heuristic_config(dtype, m, n, k);
  return EH_SUCCESS;
}

}  // namespace eh
