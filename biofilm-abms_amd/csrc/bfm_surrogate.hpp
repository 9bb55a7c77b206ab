// bfm_surrogate.hpp -- JKR surrogate network hooks of the context.
#pragma once

#include <functional>

#include "bfm_ctx.hpp"

namespace bfm {

// Runs launch() bracketed by the context's per-kernel timing events.
void launch_timed(bfm_ctx* c, const char* name, const std::function<void()>& launch);

// Contact radius for every pair of the context's neighbor list (device
// array of max_pairs doubles); pairs outside the training envelope use the
// exact FP64 solve.
const double* surrogate_pair_radii(bfm_ctx* c, const double* d_cells, int ld, int n);
void surrogate_free(bfm_ctx* c);

}  // namespace bfm
