// Binary32 FFT pass instantiations (every mesh size and launch variant).
#define HY_FFT_INSTANTIATE
#include "hy_fft_table.hpp"

namespace hy {
bool lookup_fft(int N, FftKernels<float>& out) { return lookup_fft_impl<float>(N, out); }
}  // namespace hy
