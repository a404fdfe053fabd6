// k_hash_filter instantiations for window lengths 7 and 8.
#define WS_FILTER_INSTANCES
#include "filter.hpp"

namespace ws {
WS_FILTER_INSTANCE(7)
WS_FILTER_INSTANCE(8)
}  // namespace ws
