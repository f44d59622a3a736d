// wave_media_spectral.hip — the frame wave of spectral programs with participating media.
#include "wave.h"

namespace ptmi {

void wave_media_spectral(pt_renderer* r, const Camera& cam, int frames, uint32_t sample0) {
    launch_wave<32, true, true, false>(r, cam, frames, sample0);
}

}  // namespace ptmi
