// The image-adaptive-P2 instantiation of the request's launch sequence (DenseStereoSgm.cs:267-274), compiled on
// its own so that it builds in parallel with kernels.hip.
#include "kernels_impl.h"

namespace lux {

template hipError_t launch_sgm_t<true>(const Geometry& g, const SgmBuffers& b, const SgmOutputs& o, int p1, int p2,
                                       uint16_t* agg_pass1, hipStream_t s, hipEvent_t* m);

}  // namespace lux
