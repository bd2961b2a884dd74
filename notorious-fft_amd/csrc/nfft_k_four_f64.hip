// Instantiates the column, transposed-row and Bluestein-row passes for double.
#include "nfft_launch_impl.h"

namespace nfft {
template hipError_t col_impl<double>(int, int, int, const ColArgs &, hipStream_t);
template hipError_t rowT_impl<double>(int, int, const RowTArgs &, hipStream_t);
template hipError_t rowT_r2c_impl<double>(int, const RowR2CArgs &, hipStream_t);
template hipError_t blue_row_impl<double>(int, const BlueRowArgs &, hipStream_t);
}  // namespace nfft
