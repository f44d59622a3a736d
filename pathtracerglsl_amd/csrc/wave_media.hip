// wave_media.hip — the frame wave of RGB programs with participating media.
#include "wave.h"

namespace ptmi {

void wave_media(pt_renderer* r, const Camera& cam, int frames, uint32_t sample0) {
    launch_wave<32, true, false, false>(r, cam, frames, sample0);
}

}  // namespace ptmi
