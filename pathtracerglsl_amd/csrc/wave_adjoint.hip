// wave_adjoint.hip — the frame wave of the path-replay adjoint of RGB programs without media.
#include "wave.h"

namespace ptmi {

void wave_adjoint(pt_renderer* r, const Camera& cam, int frames, uint32_t sample0) {
    launch_wave<32, false, false, true>(r, cam, frames, sample0);
}

}  // namespace ptmi
