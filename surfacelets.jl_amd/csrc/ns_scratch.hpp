// Internal interface between the NSSurfacelet plan (ns.hip) and the
// allocation-free analysis entry point (ns_noalloc.hip), which runs the
// plan's analysis with its working spectra and slabs carved from one region
// of device memory the caller side provides.
#pragma once

#include <cstddef>

#include "../../include/surfhip.h"

namespace surfhip {

// Bytes the plan's analysis needs when its working buffers come from one
// region (every piece 256-byte aligned): the factored path's forward-spectrum
// staging plane, one spectrum per channel and the slabs; the general path's
// spectra and at least one work spectrum of a batch.
size_t ns_analyze_working_bytes(const surfhip_ns_plan *plan);

// The device the plan's stream belongs to.
int ns_plan_device(const surfhip_ns_plan *plan);

// analyze! (the plan's kernels and launch order) with the working buffers
// carved from [ws, ws + bytes) instead of the plan's own; the coefficients
// equal surfhip_ns_analyze's bit for bit.  Checks the caller's buffers as
// surfhip_ns_analyze does.  Allocates nothing.
void ns_analyze_in(surfhip_ns_plan *plan, void *ws, size_t bytes, const void *x, void *coeffs);

}  // namespace surfhip
