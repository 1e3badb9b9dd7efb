// fft_g2.hip -- instantiates the LDS FFT passes for the lengths of group 2 (fft_passes.hpp).
#include "fft_passes.hpp"

namespace hp {
HP_DEFINE_FFT_LAUNCHERS(g2, HP_FFT_G2)
}  // namespace hp
