// Instantiates the realdft / invrealdft and DCT/DST-4 single-pass kernels.
#include "nfft_launch_impl.h"

namespace nfft {
template hipError_t r2c_impl<float>(int, bool, const RtArgs &, hipStream_t);
template hipError_t r2c_impl<double>(int, bool, const RtArgs &, hipStream_t);
template hipError_t t4_impl<float>(int, int, const RtArgs &, hipStream_t);
template hipError_t t4_impl<double>(int, int, const RtArgs &, hipStream_t);
}  // namespace nfft
