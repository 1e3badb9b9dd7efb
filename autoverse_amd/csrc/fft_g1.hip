// fft_g1.hip -- instantiates the LDS FFT passes for the lengths of group 1 (fft_passes.hpp).
#include "fft_passes.hpp"

namespace hp {
HP_DEFINE_FFT_LAUNCHERS(g1, HP_FFT_G1)
}  // namespace hp
