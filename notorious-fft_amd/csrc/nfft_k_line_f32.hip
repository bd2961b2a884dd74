// Instantiates the single-pass C2C lines and single-kernel Bluestein for float.
#include "nfft_launch_impl.h"

namespace nfft {
template hipError_t line_impl<float>(int, int, const LineArgs &, hipStream_t);
template hipError_t blue_line_impl<float>(int, const BlueArgs &, hipStream_t);
}  // namespace nfft
