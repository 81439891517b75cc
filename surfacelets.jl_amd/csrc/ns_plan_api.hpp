// Internal interface of the NSSurfacelet plan (ns.hip) for code that drives
// several plans (the batch, ns_batch.cpp): the transforms without argument
// checks, for launch sequences recorded into a graph after the caller's
// buffers were checked once.
#pragma once

#include <cstddef>
#include <cstdint>

#include "../../include/surfhip.h"

namespace surfhip {

// Bytes of one volume and of one coefficient array of the plan.
size_t ns_volume_bytes(const surfhip_ns_plan *plan);
size_t ns_coefficient_bytes(const surfhip_ns_plan *plan);
// surfhip_ns_analyze / surfhip_ns_synthesize after their checks.
void ns_analyze_raw(surfhip_ns_plan *plan, const void *x, void *coeffs);
void ns_synthesize_raw(surfhip_ns_plan *plan, const void *coeffs, void *x);

}  // namespace surfhip
