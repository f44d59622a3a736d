// wave_adjoint_media.hip — the frame wave of the path-replay adjoint of RGB programs with media.
#include "wave.h"

namespace ptmi {

void wave_adjoint_media(pt_renderer* r, const Camera& cam, int frames, uint32_t sample0) {
    launch_wave<32, true, false, true>(r, cam, frames, sample0);
}

}  // namespace ptmi
