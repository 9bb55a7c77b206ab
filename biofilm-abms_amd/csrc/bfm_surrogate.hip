// bfm_surrogate.hip -- JKR surrogate (placeholder until the MLP kernels land).
#include <hip/hip_runtime.h>

#include "bfm_surrogate.hpp"

namespace bfm {

struct SurrogateWeights {
    int unused;
};

const double* surrogate_pair_radii(bfm_ctx*, const double*, int, int) { return nullptr; }

void surrogate_free(bfm_ctx* c)
{
    delete c->sur;
    c->sur = nullptr;
}

}  // namespace bfm

extern "C" int bfm_set_surrogate(bfm_ctx* c, const uint16_t*, const float*, const uint16_t*, const float*,
                                 const uint16_t*, const float*, const uint16_t*, const float*, const double*)
{
    if (!c) return BFM_ERR_ARG;
    c->err = "surrogate not available in this build";
    return BFM_ERR_UNSUPPORTED;
}
