// FP64 instantiations of the hop kernel (RB = 1).
#include "wc_hop.h"

namespace wck {
template <> bool hop<double>(const HopLaunch &a, hipStream_t s) { return hop_launch<double>(a, s); }
}  // namespace wck
