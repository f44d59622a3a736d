#!/usr/bin/env python3
"""Build the spectral data files from the reference's shipped assets.

- spectral_curves.f32: the 128 x 2 RGBA32F curves texture (CIE 1931 X, Y, Z and D65 at 5 nm from
  360 nm; z nodes of the coefficient atlas), byte-identical to
  src/components/color/spectral/data/curves.bin (SHA-256 pinned in spectral.py).
- spectral_atlas.zlib: the Jakob-Hanika coefficient atlas src/components/color/spectral/data/
  coefficients.bin (3 x 64^3 RGBA32F texels whose w is always 0) stored as its three used
  float32 channels, byte planes shuffled (plane k holds byte k of every float) and deflated;
  spectral.atlas() inflates it and checks the reference's SHA-256 of the RGBA bytes
  (src/components/color/spectral/data.ts).

Usage: python3 pathtracerglsl_amd/data/make_spectral_data.py [/root/reference]
"""
import os
import shutil
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from pathtracerglsl_amd import spectral  # noqa: E402

def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    d = os.path.join(ref, "src", "components", "color", "spectral", "data")
    shutil.copyfile(os.path.join(d, "curves.bin"), spectral.CURVES_PATH)
    spectral._curves = None
    spectral.curves_texture()                       # verifies the SHA-256
    atlas = spectral.load_atlas(os.path.join(d, "coefficients.bin"))
    spectral.write_atlas(atlas, spectral.ATLAS_PATH)
    spectral._atlas = None
    assert np.array_equal(spectral.atlas(), atlas)
    print(f"wrote {spectral.CURVES_PATH} and {spectral.ATLAS_PATH}")

if __name__ == "__main__":
    main()
