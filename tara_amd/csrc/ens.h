// Library-internal access to an ensemble context (capi_ens.hip).
#pragma once
#include "common.h"

namespace tara {
// The stream that orders an ensemble's uploads, steps and downloads.
hipStream_t ens_stream(tara_ens* e);
// Grid size of one run (0 for an invalid run).
int ens_grid_of(tara_ens* e, int run);
}  // namespace tara
