// FP32 instantiations of the hop kernel (RB = 1, 2).
#include "wc_hop.h"

namespace wck {
template <> bool hop<float>(const HopLaunch &a, hipStream_t s) { return hop_launch<float>(a, s); }
}  // namespace wck
