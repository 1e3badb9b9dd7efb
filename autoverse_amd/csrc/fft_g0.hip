// fft_g0.hip -- instantiates the LDS FFT passes for the lengths of group 0 (fft_passes.hpp).
#include "fft_passes.hpp"

namespace hp {
HP_DEFINE_FFT_LAUNCHERS(g0, HP_FFT_G0)
}  // namespace hp
