// FP32 instantiations of the hop kernel (RB = 1, 2).  FP16-storage hops of
// a multiple of 4 right-hand sides go to the matrix-core kernel
// (wc_hop_mma.h) unless the launch turns it off.  Both live in this one code
// object, which HIP loads on the first FP32 hop launch: an eager launch
// (tuning, admission, the Schur source terms) that precedes every captured
// iteration run, so no kernel of a captured run triggers a module load.
#include "wc_hop.h"
#include "wc_hop_mma.h"

namespace wck {
template <> bool hop<float>(const HopLaunch &a, hipStream_t s) {
  if (hop_mma_applies(a) && hop_mma(a, s)) return true;
  return hop_launch<float>(a, s);
}
}  // namespace wck
