// wave_rgb.hip — the frame wave of RGB programs without media (the 16- and 32-entry traversal stacks).
#include "wave.h"

namespace ptmi {

void wave_rgb(pt_renderer* r, const Camera& cam, int frames, uint32_t sample0, int stack) {
    if (stack == 16) launch_wave<16, false, false, false>(r, cam, frames, sample0);
    else launch_wave<32, false, false, false>(r, cam, frames, sample0);
}

}  // namespace ptmi
