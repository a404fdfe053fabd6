// k_hash_filter instantiations for window lengths 11 and 12.
#define WS_FILTER_INSTANCES
#include "filter.hpp"

namespace ws {
WS_FILTER_INSTANCE(11)
WS_FILTER_INSTANCE(12)
}  // namespace ws
