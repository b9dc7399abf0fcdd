// rmq_internal.hpp — library-internal access to a built hierarchy (not part of the C ABI).
#pragma once

#include "../../include/rmq.h"
#include "rmq_device.hpp"

// The device-side level view of a built hierarchy (valid until rmq_free).
const rmq::HierView* rmq_internal_view(const rmq_hierarchy* h);
// Device bytes the hierarchy occupies (levels, offset entries, top-level table, small words).
uint64_t rmq_internal_level_bytes(const rmq_hierarchy* h);
// Byte offsets (roff[k]) of the offset entries of levels 1.. of a hierarchy with these level
// sizes and reduction factor 2^R_log; returns their total bytes.
uint64_t rmq_internal_rel_layout(uint32_t R_log, uint32_t num_levels, const uint64_t* size, uint32_t* roff);
// Levels 1.. of the hierarchy view H (reduction factor 2^R_log) into d_val / d_idx at its level
// offsets (rmq_read_levels); launches on `stream`.
rmq_status rmq_internal_read_levels(const rmq::HierView& H, uint32_t R_log, float* d_val, uint32_t* d_idx,
                                    void* stream);
