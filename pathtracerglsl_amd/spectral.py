"""Spectral color model, host side: the data the four-wavelength transport reads and the lift of
authored RGB values into per-material spectra.

The device evaluates a spectrum at a path's wavelengths from a pt_spectrum record
(include/ptmi.h); this module builds those records the way the reference's GLSL would evaluate
its lifts for a constant authored value:

- reflectance role: spectral_bounded(rgb) (src/components/color/spectral/lookup.glsl:40-48);
- coefficient role: 2 max(rgb) * spectral_bounded(rgb / (2 max(rgb))) (lookup.glsl:52-56);
- emission role: the coefficient lift times the D65 curve (lookup.glsl:57-60);
- blackbody emission: scale * Planck relative to 550 nm with a host CIE-Y normalisation
  (src/components/color/spectra/blackbody/blackbody.ts:13-41, blackbody.glsl:4-10).

spectral_bounded's sigmoid-quadratic coefficients come from the Jakob-Hanika coefficient atlas
through the trilinear lookup of lookup.glsl:16-38 (spectral_coefficients), restated here in
float32 operations in the same order. The atlas (3 x 64^3 RGBA32F texels, 12.6 MB) is the
reference's shipped asset `coefficients.bin`; it ships as data/spectral_atlas.zlib (its three
used channels, byte planes shuffled, deflated: 7 MB) and is checked against the reference's
SHA-256 when it is inflated, so any authored colour lifts. The CIE 1931 / D65 curves (curves.bin,
4 KB) ship as data/spectral_curves.f32.
"""
import hashlib
import math
import os
import zlib

import numpy as np

from . import abi

DATA_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")
CURVES_PATH = os.path.join(DATA_DIR, "spectral_curves.f32")
ATLAS_PATH = os.path.join(DATA_DIR, "spectral_atlas.zlib")
CURVES_SHA256 = "86f1739e08f078a4665a7bc64a8b9fb0bf261c592f2696dd73af0a32ee52650e"
ATLAS_SHA256 = "a753491b81547d3f033a5fc91c523a21079ac43a7d2360c9258eb0912823748b"

# src/components/color/spectral/data.ts:15-45 (generated tables, jh-cie1931-d65-v1)
CIE_Y_INTEGRAL = 106.85702892807676
XYZ_TO_RGB = np.array([
    [3.2401742558155275, -1.5370057367330066, -0.4984883470485426],
    [-0.9692738386733962, 1.8760259581276126, 0.04155635229538314],
    [0.055653158206683866, -0.20406158009117398, 1.0574100059269926]])
RGB_TO_XYZ = np.array([
    [0.4124920696286926, 0.35757319712237773, 0.1804059457916708],
    [0.21269122340229463, 0.7151463942447555, 0.07216237831666832],
    [0.019335565763844945, 0.11919106570745922, 0.9501379811694662]])
PLANCK_C2 = 6.62607015e-34 * 299792458 / 1.380649e-23 * 1e9    # blackbody.ts:11, nm K

ROLE_REFLECTANCE, ROLE_COEFFICIENT, ROLE_EMISSION = "reflectance", "coefficient", "emission"
F32 = np.float32


class SpectralDataError(ValueError):
    pass


_curves = None


def curves_texture():
    """The 128 x 2 RGBA32F curves texture: row 0 = (X, Y, Z, D65) at 360 + 5 i nm for
    i < 95, row 1 .x = the 64 z nodes of the coefficient atlas."""
    global _curves
    if _curves is None:
        raw = open(CURVES_PATH, "rb").read()
        if hashlib.sha256(raw).hexdigest() != CURVES_SHA256:
            raise SpectralDataError(f"{CURVES_PATH}: SHA-256 does not match the shipped curves")
        _curves = np.frombuffer(raw, np.float32).reshape(2, 128, 4)
    return _curves


def curve_rows():
    """[95][4] float32: CIE X, Y, Z and D65 (what pt_scene_desc.spectral_curves carries)."""
    return np.ascontiguousarray(curves_texture()[0, :95, :])


def z_nodes():
    return curves_texture()[1, :64, 0]


ATLAS_TEXELS = 3 * 64 ** 3


def _check_atlas(rgba, where):
    h = hashlib.sha256(rgba.tobytes()).hexdigest()
    if h != ATLAS_SHA256:
        raise SpectralDataError(f"{where}: SHA-256 {h} does not match the reference's coefficient atlas")
    return rgba.reshape(-1, 4)


def load_atlas(path):
    """The reference's coefficients.bin (3 x 64^3 RGBA32F texels), SHA-256 checked."""
    raw = np.fromfile(path, np.float32)
    if raw.size != ATLAS_TEXELS * 4:
        raise SpectralDataError(f"{path}: expected 3 x 64^3 RGBA32F texels")
    return _check_atlas(raw, path)


def write_atlas(rgba, path):
    """The shipped form of the atlas: channels x, y, z as float32 (w is 0 in every texel), the
    four byte planes one after another, deflated."""
    rgba = np.asarray(rgba, np.float32).reshape(-1, 4)
    if rgba.shape[0] != ATLAS_TEXELS or rgba[:, 3].any():
        raise SpectralDataError("write_atlas: expected 3 x 64^3 texels with w = 0")
    planes = np.ascontiguousarray(rgba[:, :3]).view(np.uint8).reshape(-1, 4).T
    with open(path, "wb") as f:
        f.write(zlib.compress(np.ascontiguousarray(planes).tobytes(), 9))


_atlas = None


def atlas():
    """The coefficient atlas as [3 * 64^3][4] float32 (RGBA, w = 0), inflated from the shipped
    file and checked against the reference's SHA-256."""
    global _atlas
    if _atlas is None:
        raw = zlib.decompress(open(ATLAS_PATH, "rb").read())
        if len(raw) != ATLAS_TEXELS * 12:
            raise SpectralDataError(f"{ATLAS_PATH}: wrong inflated size")
        xyz = np.frombuffer(raw, np.uint8).reshape(4, -1).T.copy().view(np.float32).reshape(-1, 3)
        rgba = np.zeros((ATLAS_TEXELS, 4), np.float32)
        rgba[:, :3] = xyz
        _atlas = _check_atlas(rgba, ATLAS_PATH)
    return _atlas


def _mix(a, b, t):
    """GLSL mix(a, b, t) = a (1 - t) + b t, float32."""
    return F32(F32(a * F32(F32(1.0) - t)) + F32(b * t))


def spectral_coefficients(rgb, atlas):
    """spectral_coefficients (lookup.glsl:16-38) in float32 operations."""
    r = [F32(v) for v in rgb]
    ch = 0 if (r[0] >= r[1] and r[0] >= r[2]) else (1 if r[1] >= r[2] else 2)
    z = r[ch]
    x = F32(F32(r[(ch + 1) % 3] / z) * F32(63.0))
    y = F32(F32(r[(ch + 2) % 3] / z) * F32(63.0))
    ix, iy = min(int(x), 62), min(int(y), 62)
    fx, fy = F32(x - F32(ix)), F32(y - F32(iy))
    zn = z_nodes()
    lo, hi = 0, 63
    for _ in range(6):
        mid = (lo + hi) // 2
        if zn[mid] <= z:
            lo = mid
        else:
            hi = mid
    z0, z1 = zn[lo], zn[lo + 1]
    fz = F32(F32(z - z0) / F32(z1 - z0))

    def node(xx, yy, zz):
        return atlas[((ch * 64 + zz) * 64 + yy) * 64 + xx, :3]

    out = []
    for c in range(3):
        def plane(zz):
            return _mix(_mix(node(ix, iy, zz)[c], node(ix + 1, iy, zz)[c], fx),
                        _mix(node(ix, iy + 1, zz)[c], node(ix + 1, iy + 1, zz)[c], fx), fy)
        out.append(_mix(plane(lo), plane(lo + 1), fz))
    return out


def _lookup(bounded):
    """Coefficients of a non-grey bounded rgb from the atlas."""
    return spectral_coefficients(bounded, atlas())


def bounded_input(rgb, role):
    """(scale, bounded rgb) as the GLSL computes them for the role (float32)."""
    r = [F32(v) for v in rgb]
    if role == ROLE_REFLECTANCE:
        return None, r
    scale = F32(F32(2.0) * max(r[0], max(r[1], r[2])))
    if scale == 0.0:
        return scale, None
    return scale, [F32(v / scale) for v in r]


def is_gaussian(value):
    return isinstance(value, dict) and "gaussian" in value


def gaussian_params(value):
    """(center, width, peak, floor) of {'gaussian': {...}} (gaussian.ts fields), validated."""
    g = value["gaussian"]
    try:
        center, width, peak = float(g["center"]), float(g["width"]), float(g["peak"])
        floor = float(g.get("floor", 0.0))
    except (KeyError, TypeError, ValueError):
        raise SpectralDataError("gaussian values need numeric center, width and peak") from None
    if not math.isfinite(center):
        raise SpectralDataError("gaussian center must be finite")
    if not (width > 0.0 and math.isfinite(width)):
        raise SpectralDataError("gaussian width must be positive")
    if not (0.0 <= peak <= 1.0 and 0.0 <= floor <= 1.0):
        raise SpectralDataError("gaussian peak and floor must lie in [0, 1]")
    return center, width, peak, floor


def gaussian_reflectance(value, lam):
    """gaussianReflectance (gaussian.ts): the band's reflectance at lam nm, float64."""
    center, width, peak, floor = gaussian_params(value)
    d = (lam - center) / width
    return floor + (peak - floor) * math.exp(-0.5 * d * d)


def gaussian_stand_in(value):
    """gaussianStandIn (gaussian.ts): D65 reflected by the band in linear sRGB, each channel
    divided by the same projection of a perfect white reflector, Simpson's rule at 0.5 nm in
    each 5 nm interval of the tabulated observer and D65 (linear between knots), float64."""
    rows = curve_rows().astype(np.float64)
    num, den = [0.0, 0.0, 0.0], [0.0, 0.0, 0.0]
    for interval in range(rows.shape[0] - 1):
        for sample in range(11):
            fraction = sample / 10
            lam = 360 + 5 * (interval + fraction)
            weight = 0.5 / 3 * (1 if sample in (0, 10) else (4 if sample % 2 else 2))
            cmf = [rows[interval][j] * (1 - fraction) + rows[interval + 1][j] * fraction for j in range(3)]
            light = weight * (rows[interval][3] * (1 - fraction) + rows[interval + 1][3] * fraction)
            g = gaussian_reflectance(value, lam)
            for c in range(3):
                response = XYZ_TO_RGB[c][0] * cmf[0] + XYZ_TO_RGB[c][1] * cmf[1] + XYZ_TO_RGB[c][2] * cmf[2]
                num[c] += g * light * response
                den[c] += light * response
    return [num[c] / den[c] for c in range(3)]


def spectrum_record(value, role):
    """pt_spectrum for an authored value (rgb triple, scalar, {'blackbody': {...}} or
    {'gaussian': {...}})."""
    rec = np.zeros(1, abi.SPECTRUM_DTYPE)[0]
    if is_gaussian(value):
        if role != ROLE_REFLECTANCE:
            raise SpectralDataError("gaussian values are reflectances only")
        center, width, peak, floor = gaussian_params(value)
        # gaussian.ts emit: Spectrum(floor) + (peak - floor) * exp(-0.5 / width^2 * s * s), each
        # constant a GLSL literal of the host's float64 value
        rec["kind"] = abi.SPEC_GAUSSIAN
        rec["value"] = floor
        rec["c"][:3] = [center, -0.5 / (width * width), peak - floor]
        return rec
    if isinstance(value, dict) and "blackbody" in value:
        if role != ROLE_EMISSION:
            raise SpectralDataError("blackbody values are emission only")
        bb = value["blackbody"]
        kelvin = float(bb["kelvin"])
        if not (1000.0 <= kelvin <= 40000.0):
            raise SpectralDataError("blackbody kelvin must be in [1000, 40000]")
        rec["kind"] = abi.SPEC_BLACKBODY
        rec["scale"] = float(bb.get("scale", 1.0))
        rec["c"][:2] = [kelvin, blackbody_color(kelvin)[1]]
        return rec
    rgb = [float(value)] * 3 if not hasattr(value, "__len__") else [float(v) for v in value]
    scale, b = bounded_input(rgb, role)
    rec["scaled"] = int(role != ROLE_REFLECTANCE)
    rec["d65"] = int(role == ROLE_EMISSION)
    if role != ROLE_REFLECTANCE:
        rec["scale"] = scale
    if b is None:
        rec["kind"] = abi.SPEC_ZERO
    elif b[0] == b[1] and b[1] == b[2]:
        rec["kind"] = abi.SPEC_GREY
        rec["value"] = b[0]
    else:
        rec["kind"] = abi.SPEC_SIGMOID
        rec["c"][:3] = _lookup(b)
    return rec


def rgb_of(value):
    """RGB stand-in of an authored value in an RGB program (blackbody: blackbodyRGB; a Gaussian
    band: gaussianStandIn clipped to [0, 1])."""
    if isinstance(value, dict) and "blackbody" in value:
        bb = value["blackbody"]
        rgb, _ = blackbody_color(float(bb["kelvin"]))
        return [v * float(bb.get("scale", 1.0)) for v in rgb]
    if is_gaussian(value):          # a reflectance: its stand-in clipped to [0, 1] (rgbStandIn)
        return [min(1.0, max(0.0, c)) for c in gaussian_stand_in(value)]
    return [float(value)] * 3 if not hasattr(value, "__len__") else [float(v) for v in value]


def planck_shape(lam, kelvin):
    """Planck radiance relative to 550 nm (blackbody.ts:13-17), float64."""
    return (550.0 / lam) ** 5 * math.expm1(PLANCK_C2 / (550.0 * kelvin)) / math.expm1(PLANCK_C2 / (lam * kelvin))


def blackbody_color(kelvin):
    """blackbodyColor (blackbody.ts:21-41): (rgb at unit CIE Y, normalization), float64,
    Simpson over each 5 nm CMF interval with 0.5 nm steps on the float32 CMFs."""
    cmf = curve_rows()[:, :3].astype(np.float64)
    xyz = [0.0, 0.0, 0.0]
    for interval in range(cmf.shape[0] - 1):
        for sample in range(11):
            fraction = sample / 10
            lam = 360 + 5 * (interval + fraction)
            weight = 0.5 / 3 * (1 if sample in (0, 10) else (4 if sample % 2 else 2))
            value = weight * planck_shape(lam, kelvin) / CIE_Y_INTEGRAL
            for c in range(3):
                response = cmf[interval][c] * (1 - fraction) + cmf[interval + 1][c] * fraction
                xyz[c] += value * response
    rgb = [sum(XYZ_TO_RGB[r][i] * xyz[i] for i in range(3)) for r in range(3)]
    normalization = 1 / xyz[1]
    return [v * normalization for v in rgb], normalization


# ---- float64 references for tests: quadrature of lifted spectra against the observer ----

def curves_at(lam):
    """spectralCurves (lookup.ts:33-37): linear interpolation of the float32 samples."""
    c = curve_rows().astype(np.float64)
    x = (lam - 360.0) / 5.0
    i = min(int(math.floor(x)), 93)
    f = x - i
    return c[i] * (1 - f) + c[i + 1] * f


def sigmoid(c, lam):
    """evaluatePolynomial (lookup.ts:26-30), float64."""
    t = (lam - 360.0) / 470.0
    p = (c[0] * t + c[1]) * t + c[2]
    return 0.5 + 0.5 * p / math.sqrt(1 + p * p)


def project_rgb(spectrum, step=0.25):
    """Signed linear sRGB of a spectrum under the CIE 1931 observer (Simpson, 360-830 nm)."""
    n = int(round(470 / step))
    xyz = np.zeros(3)
    for i in range(n + 1):
        lam = 360 + step * i
        w = step / 3 * (1 if i in (0, n) else (4 if i % 2 else 2))
        xyz += w * spectrum(lam) * curves_at(lam)[:3]
    xyz /= CIE_Y_INTEGRAL
    return XYZ_TO_RGB @ xyz
