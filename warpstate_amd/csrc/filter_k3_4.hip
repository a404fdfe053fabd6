// k_hash_filter instantiations for window lengths 3 and 4.
#define WS_FILTER_INSTANCES
#include "filter.hpp"

namespace ws {
WS_FILTER_INSTANCE(3)
WS_FILTER_INSTANCE(4)
}  // namespace ws
