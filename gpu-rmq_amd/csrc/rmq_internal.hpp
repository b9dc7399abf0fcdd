// rmq_internal.hpp — library-internal access to a built hierarchy (not part of the C ABI).
#pragma once

#include "../../include/rmq.h"
#include "rmq_device.hpp"

// The device-side level view of a built hierarchy (valid until rmq_free).
const rmq::HierView* rmq_internal_view(const rmq_hierarchy* h);
// Device bytes the hierarchy's level buffers occupy.
uint64_t rmq_internal_level_bytes(const rmq_hierarchy* h);
