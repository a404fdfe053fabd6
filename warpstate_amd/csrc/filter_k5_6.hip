// k_hash_filter instantiations for window lengths 5 and 6.
#define WS_FILTER_INSTANCES
#include "filter.hpp"

namespace ws {
WS_FILTER_INSTANCE(5)
WS_FILTER_INSTANCE(6)
}  // namespace ws
