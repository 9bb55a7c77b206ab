// bfm_surrogate.hpp -- JKR surrogate network hooks of the context.
#pragma once

#include <hip/hip_runtime.h>

#include <functional>

#include "bfm_ctx.hpp"

namespace bfm {

struct DistRefresh;   // bfm_device.hpp

// Training envelope of the learned radius: overlap delta in [dlo R, dhi R]
// and gamma in [glo, ghi] Pa um. Rows inside it take the network's radius;
// the others the exact FP64 solve.
struct SurrogateEnvelope {
    double dlo, dhi;     // delta / R range
    double glo, ghi;     // gamma range (Pa um)
    double R;            // reference radius
};

__host__ __device__ inline bool sur_inside(double delta, double gamma, const SurrogateEnvelope& e)
{
    return delta >= e.dlo * e.R && delta <= e.dhi * e.R && gamma >= e.glo && gamma <= e.ghi;
}

// The envelope of the context's network; an empty one when none is set.
SurrogateEnvelope surrogate_envelope(const bfm_ctx* c);

// Runs launch() bracketed by the context's per-kernel timing events.
void launch_timed(bfm_ctx* c, const char* name, const std::function<void()>& launch);

// Network contact radius for the pairs of the context's neighbor list that
// overlap, adhere and lie inside the training envelope (device array of
// max_pairs doubles, indexed by row; other rows are not written: the pair
// force kernel solves them exactly, by the same sur_inside test). rf.geom
// set: the list's distances are refreshed from the packed segments first
// (updateNeighborDistances), in the staging kernel.
const double* surrogate_pair_radii(bfm_ctx* c, const double* d_cells, int ld, int n, const DistRefresh& rf);
void surrogate_free(bfm_ctx* c);

}  // namespace bfm
