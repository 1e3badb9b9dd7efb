// fft_g3.hip -- instantiates the LDS FFT passes for the lengths of group 3 (fft_passes.hpp).
#include "fft_passes.hpp"

namespace hp {
HP_DEFINE_FFT_LAUNCHERS(g3, HP_FFT_G3)
}  // namespace hp
