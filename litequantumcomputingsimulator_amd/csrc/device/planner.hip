// Cost-model guided pass planning (placeholder: greedy grouping).
#include "engine.hpp"

namespace lqcs_dev {

void apply_learned(amp_t* s, uint32_t nq, const std::vector<DecodedOp>& ops, uint32_t L,
                   uint32_t low, uint32_t batch) {
    apply_ops(s, nq, ops, LQCS_PLAN_GREEDY, L, low, batch);
}

}  // namespace lqcs_dev
