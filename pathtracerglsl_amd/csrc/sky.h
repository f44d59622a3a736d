// Procedural sky bodies (PT_SKY_*): the RGB of each formula at a unit direction in the table's
// azimuth frame (features/environment.ts:207-213 evaluates environment.glsl.source at
// env_rotate_cs(normalize(dir), u_envRotCS); the bake, EnvironmentBake.ts:61-67, at the chart's
// texel centres with rotation 0). One definition for the radiance kernels, the device bake and
// the host table builder, so the three agree bit for bit. f32, no contraction.
#ifndef PTMI_SKY_H
#define PTMI_SKY_H

#include "../../include/ptmi.h"
#include "../../include/ptmi_math.h"

// normalize(vec3(0.4, 0.35, 0.2)): v * (1 / sqrt(dot(v, v)))
PTM_FN void pt_sky_sun(float* sx, float* sy, float* sz) {
    const float x = 0.4f, y = 0.35f, z = 0.2f;
    const float inv = 1.0f / sqrtf(x * x + y * y + z * z);
    *sx = x * inv; *sy = y * inv; *sz = z * inv;
}

// rgb[0..2] of `body` at (dx, dy, dz); a body outside PT_SKY_* is black.
PTM_FN void pt_sky_body(int body, float dx, float dy, float dz, float* rgb) {
    if (body == PT_SKY_SUN || body == PT_SKY_BIG_SUN) {
        float sx, sy, sz;
        pt_sky_sun(&sx, &sy, &sz);
        const float c = dx * sx + dy * sy + dz * sz;
        if (body == PT_SKY_SUN) {
            // mix(vec3(0.06, 0.08, 0.12), vec3(0.35, 0.45, 0.65), s) + vec3(18, 16, 13) * smoothstep(0.9994, 0.9999, c)
            const float s = 0.5f * (dy + 1.0f);
            const float k = ptm_smoothstep(0.9994f, 0.9999f, c);
            rgb[0] = 0.06f * (1.0f - s) + 0.35f * s + 18.0f * k;
            rgb[1] = 0.08f * (1.0f - s) + 0.45f * s + 16.0f * k;
            rgb[2] = 0.12f * (1.0f - s) + 0.65f * s + 13.0f * k;
        } else {
            // vec3(0.02) + vec3(40.0) * smoothstep(0.985, 0.99, c)
            const float v = 0.02f + 40.0f * ptm_smoothstep(0.985f, 0.99f, c);
            rgb[0] = rgb[1] = rgb[2] = v;
        }
        return;
    }
    float v = 0.0f;
    if (body == PT_SKY_CAP) v = dy < 0.9f ? 0.0f : 1.0f;   // vec3(step(0.9, dir.y))
    else if (body == PT_SKY_WHITE) v = 1.0f;                 // vec3(1.0)
    rgb[0] = rgb[1] = rgb[2] = v;
}

// A body whose every value is an exact grey: its spectral lift is value * D65 with no atlas.
PTM_FN int pt_sky_body_grey(int body) { return body == PT_SKY_BIG_SUN || body == PT_SKY_CAP || body == PT_SKY_WHITE; }

#endif
