// Instantiates the DCT/DST-2/3 single-pass kernels for float.
#include "nfft_launch_impl.h"

namespace nfft {
template hipError_t t23_impl<float>(int, int, const RtArgs &, hipStream_t);
}  // namespace nfft
