// fft_g4.hip -- instantiates the LDS FFT passes for the lengths of group 4 (fft_passes.hpp).
#include "fft_passes.hpp"

namespace hp {
HP_DEFINE_FFT_LAUNCHERS(g4, HP_FFT_G4)
}  // namespace hp
