// k_hash_filter instantiations for window lengths 1 and 2.
#define WS_FILTER_INSTANCES
#include "filter.hpp"

namespace ws {
WS_FILTER_INSTANCE(1)
WS_FILTER_INSTANCE(2)
}  // namespace ws
