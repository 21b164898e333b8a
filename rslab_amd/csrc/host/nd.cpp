// Multilevel nested-dissection ordering (placeholder until the restatement of
// crates/rslab-metis lands).
#include <stdexcept>

#include "symbolic.hpp"

namespace rsl {
std::vector<idx> nd_order(const Pattern&, uint64_t) {
  throw std::runtime_error("MetisND ordering is not available in this build");
}
}  // namespace rsl
