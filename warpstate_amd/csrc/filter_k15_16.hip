// k_hash_filter instantiations for window lengths 15 and 16.
#define WS_FILTER_INSTANCES
#include "filter.hpp"

namespace ws {
WS_FILTER_INSTANCE(15)
WS_FILTER_INSTANCE(16)
}  // namespace ws
