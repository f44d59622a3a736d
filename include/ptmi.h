/*
 * ptmi.h — C ABI of the MI355X transport-loop backend (libptmi.so).
 *
 * This is the drop-in for the reference's executor boundary: the Engine loads a
 * CompiledRenderer, registers scene data, renders frames (1 sample per pixel each) and
 * reads exports back (src/engine/Engine.ts:83-474, src/engine/RenderExecutor.ts:159-258).
 * The GLSL program the reference generator would emit is replaced by a pt_program record
 * (the Plan stage's decisions as plain data) that selects template instantiations of
 * the HIP kernels; scene data arrives as plain arrays instead of data textures.
 *
 * Conventions: every entry point returns pt_status; on failure pt_last_error() holds a
 * message (thread-local). Host arrays are caller-owned and copied during the call;
 * device memory is owned by the library. A pt_ctx owns one HIP stream and must be used
 * from one host thread at a time. No torch types cross this boundary.
 */
#ifndef PTMI_H
#define PTMI_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PT_ABI_VERSION 15u

typedef enum pt_status {
    PT_OK = 0,
    PT_E_INVALID = 1,        /* bad argument / inconsistent description */
    PT_E_MISSING_INPUT = 2,  /* a required scene channel or table was not uploaded */
    PT_E_DEVICE = 3,         /* HIP runtime error */
    PT_E_OOM = 4,            /* device allocation failed */
    PT_E_UNSUPPORTED = 5,    /* program asks for a feature this build does not provide */
    PT_E_CONTEXT_LOST = 6    /* the renderer or batch lost its device resources to pt_ctx_reset */
} pt_status;

/* ---- scene description (the data rail: src/components/data/channels.ts:11-20) ---- */

enum { PT_OBJ_MESH = 0, PT_OBJ_QUAD = 1, PT_OBJ_SPHERE = 2, PT_OBJ_INSTANCES = 3 };

/* One placement of an instance batch (src/glsl/core/placement.glsl:1-41): q is the INVERSE
 * rotation quaternion (world -> local, x y z w) and ts = (t_rigid, s) with t_rigid = -R^T t and
 * s the forward uniform scale. A world ray maps to the prototype's frame as
 * o' = (qrot(q, o) + t_rigid) / s, d' = qrot(q, d) / s (ray t preserved); normals return by
 * the forward rotation qrot(conj(q), n). An instance batch object (PT_OBJ_INSTANCES) is ONE
 * region for all of its placements (features/intersection.ts:981-1066 instanceLeafItem). */
typedef struct pt_instance {
    float q[4];
    float ts[4];
} pt_instance;

/* One scene object = one region (src/compiler/plan: objects are indexed by region id).
 * Meshes are triangle soups in WORLD space (identity placement) over a range of the
 * triangle arrays. closed != 0 claims an interior (watertight, consistent outward
 * winding; src/compiler/generate/features/intersection.ts:1196-1211); quads are thin. */
typedef struct pt_object {
    int32_t kind;
    int32_t material;
    int32_t closed;
    int32_t first_tri;       /* meshes: world triangles; instance batches: prototype triangles */
    int32_t tri_count;
    int32_t first_instance;  /* instance batches: placements [first_instance, + instance_count) */
    int32_t instance_count;
    int32_t smooth;          /* meshes / batches: barycentric per-vertex shading normals
                                (components/intersection/mesh/mesh.glsl:84-109; Planner.ts:142) */
    float   p[12];           /* quad: corner, edge1, edge2, normal; sphere: center, radius */
    float   box_min[4];      /* meshes: vertex AABB (containment quick reject); batches: local */
    float   box_max[4];
} pt_object;

/* Material models (src/components/materials/): PT_MAT_MIRROR is the perfect specular conductor
 * (mirror/mirror.glsl): a delta reflection weighted by Schlick's Fresnel of f0, no emission.
 * PT_MAT_ROUGH_DIELECTRIC is the two-lobe microfacet dielectric (rough_dielectric/
 * rough_dielectric.glsl, Walter et al. 2007): GGX reflection and transmission with the exact
 * Fresnel, alpha = max(1e-3, roughness^2), the region's index and the transmittance tint. */
enum { PT_MAT_LAMBERT = 0, PT_MAT_DIELECTRIC = 1, PT_MAT_GGX = 2, PT_MAT_NULL = 3, PT_MAT_MIRROR = 4,
       PT_MAT_ROUGH_DIELECTRIC = 5, PT_MAT_CHECKER = 6 };
/* PT_MAT_CHECKER (materials/checker/checker.glsl): Lambert transport whose albedo is
 * mix(albedo, f0, mod(floor(u s) + floor(v s), 2)) at the hit's surface coordinates (Hit.uv,
 * glsl/core/structs.glsl:40-43): quad_uv (quad.glsl:33-41) on a quad, sphere_uv (sphere.glsl:84-89)
 * on a sphere, the planar placeholder (x, z) * 0.1 on meshes and instances; s = pad1 (uv_scale).
 * The record reuses albedo / s_albedo for albedo_a, f0 / s_f0 for albedo_b. */

/* The arm a homogeneous medium is sampled by (src/compiler/generate/features/materials.ts:675-728,
 * routing predicates src/compiler/types.ts:351-417):
 *   PT_MED_ANALYTIC  scattering, non-emissive: channel-MIS distance sample (analytic.glsl:9-43)
 *   PT_MED_ABSORB    non-scattering: deterministic exp(-sigma_a t_max) and, when emissive,
 *                    radiance eps * integral_0^t_max exp(-sigma_a s) ds (analytic/emission.glsl)
 *   PT_MED_DELTA     scattering and emissive: delta tracking at the derived majorant
 *                    (delta_tracking.glsl:56-120), emission collected per tentative collision
 * The host decides the arm from the authored values, as the reference's Planner does. */
enum { PT_MED_ANALYTIC = 0, PT_MED_ABSORB = 1, PT_MED_DELTA = 2 };
enum { PT_IOR_CONSTANT = 0, PT_IOR_CAUCHY = 1, PT_IOR_SELLMEIER = 2 };
/* Volume scattering models (components/volume_scattering/): Henyey-Greenstein with hg_g
 * (hg/hg.glsl) or Rayleigh, (3 / 16 pi)(1 + cos^2), parameter-free (rayleigh/rayleigh.glsl). */
enum { PT_PHASE_HG = 0, PT_PHASE_RAYLEIGH = 1 };

/* A spectral value in one radiometric role (src/components/descriptors.ts:43), as the
 * spectral color model evaluates it at a path's four wavelengths
 * (src/components/color/spectral/lookup.glsl:40-60, color/spectra/blackbody/blackbody.glsl:4-10):
 *   PT_SPEC_ZERO       0
 *   PT_SPEC_GREY       value                                  (exact grey: no sigmoid)
 *   PT_SPEC_SIGMOID    clamp(0.5 + 0.5 p / sqrt(1 + p^2), 0, 1),  p = (c0 t + c1) t + c2,
 *                      t = (lambda - 360) / 470              (Jakob-Hanika coefficients c)
 *   PT_SPEC_BLACKBODY  scale * Planck(c0 = kelvin, c1 = normalization)
 *   PT_SPEC_GAUSSIAN   value + c2 exp((c1 (lambda - c0)) (lambda - c0)): a reflectance band
 *                      (color/spectra/gaussian/gaussian.ts emit): value = floor, c0 = center (nm),
 *                      c1 = -0.5 / width^2, c2 = peak - floor (the host's f64 values, in f32)
 * GREY and SIGMOID are multiplied by `scale` when scaled != 0 (role 'coefficient':
 * scale = 2 max(rgb)) and then by the D65 curve when d65 != 0 (role 'emission').
 * The host computes c from the authored RGB (spectral_coefficients, lookup.glsl:16-38). */
enum { PT_SPEC_ZERO = 0, PT_SPEC_GREY = 1, PT_SPEC_SIGMOID = 2, PT_SPEC_BLACKBODY = 3, PT_SPEC_GAUSSIAN = 4 };
typedef struct pt_spectrum {
    int32_t kind;
    int16_t scaled, d65;
    float   value;
    float   scale;
    float   c[3];
    float   pad;
} pt_spectrum;

/* Material record: the MaterialProperties the generator bakes per material id
 * (src/compiler/generate/features/materials.ts, components/materials/{lambert,dielectric,ggx}).
 * PT_MAT_NULL is the reference's model 'none': a null interface that only bounds a medium
 * (components/transport/integrators/pt/pt.ts:378-388). Interior medium (homogeneous, HG or
 * Rayleigh phase) when has_medium != 0 (features/materials.ts scene_medium_properties). */
typedef struct pt_material {
    int32_t model;
    int32_t emissive;        /* material_is_emissive */
    int32_t has_medium;
    int32_t medium_arm;      /* PT_MED_* */
    float   albedo[4];       /* lambert */
    float   emission[4];
    float   f0[4];           /* ggx */
    float   transmittance[4];/* dielectric */
    float   alpha;           /* ggx: max(1e-3, roughness^2), host-derived */
    float   ior;             /* dielectric */
    float   hg_g;            /* medium */
    float   pad1;
    float   sigma_s[4];      /* medium */
    float   sigma_a[4];      /* medium */
    /* spectral color model: the same properties in their roles */
    pt_spectrum s_albedo, s_emission, s_f0, s_transmittance, s_sigma_s, s_sigma_a;
    float   medium_emission[4];  /* medium: volume emission coefficient eps (dL/ds), RGB */
    pt_spectrum s_medium_emission;   /* spectral, role 'emission' */
    /* Delta-tracking majorant sigma_bar, derived by the library at upload and at every
     * material update (features/materials.ts:774-813 derivedMajorant + the color model's
     * mediumBound): [0] RGB = max(1e-6, max_c(sigma_a + sigma_s)); [1] spectral = the
     * float32 bound of components/color/spectral/majorant.ts. Values passed in are ignored. */
    float   majorant[2];
    /* Dispersive index of a dielectric (src/components/color/spectra/cauchy, sellmeier), read at
     * the hero wavelength by spectral programs (ior_of, features/intersection.ts:1609-1649);
     * RGB programs and PT_IOR_CONSTANT use `ior` (the index at 587.5618 nm, host-evaluated).
     *   PT_IOR_CAUCHY     ior_coef = A, B, C (micrometres): A + B / l^2 + C / l^4
     *   PT_IOR_SELLMEIER  ior_coef = B1 B2 B3 C1 C2 C3: n^2 = 1 + sum B_i l^2 / (l^2 - C_i) */
    int32_t ior_kind;
    int32_t phase;           /* PT_PHASE_*: the medium's volume scattering model (medium.model) */
    float   ior_coef[8];
} pt_material;

enum { PT_LIGHT_QUAD = 0, PT_LIGHT_POINT = 1 };

/* Finite light with its host-computed selection pdf and CDF bound
 * (src/compiler/generate/features/lighting.ts:537-620). */
typedef struct pt_light {
    int32_t kind;
    int32_t region;          /* object index of a hittable light (quad), -1 for delta */
    float   select_pdf;
    float   cdf;             /* upper CDF bound of this light in the selection chain */
    float   corner[4], edge1[4], edge2[4], normal[4];
    float   radiance[4];     /* quad: Le; point: intensity */
    float   area;
    float   pad[3];
    pt_spectrum s_radiance;  /* spectral color model, role 'emission' */
} pt_light;

enum { PT_ENV_NONE = 0, PT_ENV_SKY = 1, PT_ENV_CONSTANT = 2, PT_ENV_IMAGE = 3 };
/* PT_ENV_IMAGE (features/environment.ts:144-190): an equirect picture (pt_scene_desc.env_image,
 * the reference's env_map: RGB32F, NEAREST, clamped edges) read at equirect_uv(dir) (the additive
 * rotation) times intensity, lifted as an emission in a spectral program; its importance table is
 * built from the picture (components/env/importance.ts:16-127) on the equirect chart (the
 * picture's own W x H) or, on the octahedral chart, from the picture resampled to N x N, N =
 * table_h (256 default; resampleEquirectToOctahedral, env/octahedral/octahedral.ts:72-100, f64). */

/* Environment. PT_ENV_SKY is a procedural sky (features/environment.ts:191-231): one of the
 * formula bodies below evaluated at the direction turned into the table's azimuth, sampled
 * through an importance table the library bakes itself at texel centres
 * (src/compiler/EnvironmentBake.ts:29-80, src/components/env/importance.ts:16). */
/* Procedural sky bodies (environment.glsl.source), the formulas of the reference's scenes:
 *   PT_SKY_SUN      mix(vec3(0.06, 0.08, 0.12), vec3(0.35, 0.45, 0.65), 0.5 * (dir.y + 1.0))
 *                   + vec3(18, 16, 13) * smoothstep(0.9994, 0.9999, dot(dir, normalize(vec3(0.4, 0.35, 0.2))))
 *                   (proc-sky, generated-glsl.snapshot.test.ts.snap:63758)
 *   PT_SKY_BIG_SUN  vec3(0.02) + vec3(40.0) * smoothstep(0.985, 0.99, dot(dir, normalize(vec3(0.4, 0.35, 0.2))))
 *                   (proc-sky-rotated, tests/witnesses/scenes/estimatorAgreementWitness.ts:61-72)
 *   PT_SKY_CAP      vec3(step(0.9, dir.y))           (sky-missed-cap, witnesses/scenes/spectralSkies.ts:39-41)
 *   PT_SKY_WHITE    vec3(1.0)                        (spectral-comp-empty, spectralSkies.ts:48)
 * A spectral program lifts the body's RGB per direction as an emission
 * (color_lift_emission, components/color/spectral/lookup.glsl:50-60): an exact grey c is
 * c * D65, any other colour goes through the coefficient atlas of pt_scene_desc. */
enum { PT_SKY_SUN = 0, PT_SKY_BIG_SUN = 1, PT_SKY_CAP = 2, PT_SKY_WHITE = 3 };
typedef struct pt_env {
    int32_t type;
    int32_t table_w, table_h;   /* 512 x 256 default */
    int32_t compensated;        /* MIS compensation of the table (0 = plain) */
    float   intensity;
    float   rotation;           /* u_envRotation (radians) */
    float   select_prob;        /* two-stage env-vs-finite draw; ignored without finite lights */
    int32_t samplable;          /* PT_ENV_CONSTANT: sampled as a light (sampleAsLight); the sky always is */
    /* PT_ENV_CONSTANT (src/compiler/generate/features/environment.ts:65-133): radiance
     * color * intensity in every direction; sampled uniformly on the sphere, pdf 1/(4 pi) */
    float   color[4];
    pt_spectrum s_color;        /* spectral, role 'emission' (SETTING_ROLES.sky) */
    int32_t chart;              /* PT_ENV_SKY: the importance table's chart (estimator.envSampler.chart) */
    int32_t body;               /* PT_ENV_SKY: PT_SKY_* */
} pt_env;
/* Environment charts (src/components/env/): equirect W x H with sin(theta) row weights, or the
 * equal-area octahedral map (Clarberg 2008; octahedral.glsl) on a table_h x table_h grid, every
 * texel 4 pi / N^2 sr, rotated through (cos rotation, sin rotation). */
enum { PT_CHART_EQUIRECT = 0, PT_CHART_OCTAHEDRAL = 1 };

typedef struct pt_scene_desc {
    uint32_t abi_version;
    int32_t  n_tris;
    const float*   tri_vertices;   /* [n_tris][9]: a, b, c (world space) */
    const int32_t* tri_object;     /* [n_tris]: owning object (region) index */
    int32_t  n_objects;
    const pt_object* objects;
    int32_t  n_materials;
    const pt_material* materials;
    int32_t  n_lights;
    const pt_light* lights;
    pt_env   env;
    int32_t  ambient_material;     /* material filling region -1 (ambientMedium), -1 = none */
    int32_t  pad0;
    /* spectral color model data (the u_spectralCurves row 0 texture and the cie1931 sensor
     * constants, src/components/color/spectral/lookup.glsl:1-7, sensor/cie1931/cie1931.ts:6-10);
     * may be null for RGB programs */
    const float* spectral_curves;  /* [95][4]: CIE X, Y, Z, D65 at 360 + 5 i nm */
    float    xyz_to_rgb[9];        /* row-major XYZ -> linear sRGB */
    float    cie_y_integral;
    int32_t  pad1[2];
    /* instance batches: prototype triangles in their local frames [n_proto_tris][9] and the
     * placements [n_instances] the batch objects index */
    int32_t  n_proto_tris;
    int32_t  n_instances;
    const float*       proto_vertices;
    const pt_instance* instances;
    /* per-corner shading normals of smooth meshes / prototypes, the same layout as the
     * vertices ([n][9], object-local for prototypes); may be null when no object is smooth */
    const float* tri_normals;
    const float* proto_normals;
    /* the Jakob-Hanika coefficient atlas (u_spectralCoefficients, lookup.glsl:9-12: [3 * 64^3][4]
     * texels, channel-major, then z, y, x; w unused) and its 64 z nodes (u_spectralCurves row 1):
     * needed only by a spectral program whose procedural sky body is not grey; may be null */
    const float* spectral_atlas;
    const float* spectral_z_nodes;
    /* the light tree's inputs (sceneData.ts:99-143): per light its box (6 f64: min.xyz, max.xyz,
     * the kind's treeBounds: a quad's four corners, a point's position) and its selection power
     * (f64, > 0: pi A mean(Le) for a quad, 4 pi mean(I) for a point, floored at 1e-8); needed
     * only by programs with PT_SEL_BVH; may be null */
    const double* light_boxes;
    const double* light_powers;
    /* PT_ENV_IMAGE: the picture, env_image_w x env_image_h RGB f32 texels, row 0 at theta = 0
     * (finite and >= 0; width <= 4096) */
    const float* env_image;
    int32_t  env_image_w, env_image_h;
} pt_scene_desc;

/* ---- program (the Plan stage's decisions) ---- */

enum { PT_EST_PT = 0, PT_EST_NEE = 1, PT_EST_MIS = 2 };
/* measurement.color (src/compiler/types.ts): 'rgb' or 'spectral' with four hero wavelengths;
 * estimator.wavelengths: 'visible' (truncated sech^2) or 'uniform' on 360-830 nm
 * (src/components/color/spectral/visible.glsl, uniform.glsl). */
enum { PT_COLOR_RGB = 0, PT_COLOR_SPECTRAL = 1 };
enum { PT_WL_VISIBLE = 0, PT_WL_UNIFORM = 1 };
/* Medium NEE placement: at the scattering vertex (techniques/light/light_medium.glsl) or once
 * per medium segment by equiangular sampling toward a delta light (Kulla-Fajardo 2012,
 * techniques/equiangular/equiangular.glsl; pt-nee with point lights only). */
enum { PT_MLS_VERTEX = 0, PT_MLS_EQUIANGULAR = 1 };
enum { PT_VS_ANALYTIC = 0, PT_VS_DELTA_TRACKING = 1 };

typedef struct pt_program {
    uint32_t abi_version;
    int32_t  estimator;       /* directLighting: none / nee / mis */
    int32_t  max_bounces;     /* measurement.maxBounces */
    int32_t  rr_start;        /* russianRoulette.startDepth, < 0 = off */
    float    rr_max_survival; /* RR_MAX_SURVIVAL (0.95 default) */
    int32_t  color;
    int32_t  width, height;   /* film size */
    int32_t  frames_in_flight;/* samples per pixel traced per wave of paths (>= 1) */
    int32_t  max_null_crossings; /* measurement.maxNullCrossings (32 default) */
    int32_t  wavelengths;     /* PT_WL_* (spectral only) */
    int32_t  medium_light_sampling; /* PT_MLS_*: estimator.mediumLightSampling */
    int32_t  variance;        /* != 0: also accumulate the variance picture (export "variance",
                                 components/accumulator/variance/variance.glsl) */
    int32_t  volume_sampling; /* PT_VS_*: estimator.volumeSampling (validation: the arms are per material) */
    int32_t  aovs;            /* PT_AOV_* bits: measurement.aovs, the first-hit guide pictures (RGB only) */
    int32_t  env_chart;       /* estimator.envSampler.chart as PT_CHART_* + 1, 0 = not stated; when stated,
                               * a procedural sky's importance table (pt_env.chart) must be on that chart */
    int32_t  light_selection; /* PT_SEL_*: estimator.lightSelection (computeSelectPdf,
                               * compiler/generate/features/lighting.ts:537-542). The selection table itself
                               * is host-derived scene data (pt_light.select_pdf / cdf, pt_env.select_prob);
                               * under PT_SEL_UNIFORM the library checks that the resident scene holds the
                               * uniform table */
    int32_t  accumulation;    /* PT_ACC_*: estimator.accumulation.type (components/accumulator/) */
} pt_program;

/* Accumulators: the running average of every sample (average/average.glsl), or the latest sample
 * alone (oneshot/oneshot.glsl: the picture updates live, the sample count still advances, and it
 * keeps no statistic, so a oneshot program with variance != 0 is PT_E_INVALID). The guide
 * pictures follow the same rule. */
enum { PT_ACC_AVERAGE = 0, PT_ACC_ONESHOT = 1 };

/* Light selection (components/lights/index.ts LIGHT_SELECTIONS): the power-weighted CDF (the
 * reference's default), one of N lights with equal probability, the environment then taking
 * 1/(N+1) of the NEE draws (envSelectionProbability, lighting.ts:488-518), or the light tree
 * (lighting.ts:676-933): a stochastic descent of the tree pt_scene_upload builds over the lights
 * (pt_scene_desc.light_boxes / light_powers, pt_light_tree_build), each level choosing a child
 * in proportion to its importance at the shading point (accel/light_tree/light_tree.glsl:28-77:
 * power over clamped squared distance, the receiver's horizon cull and best-case cosine),
 * drawing a fresh random number per level and one more for the light's own sample; MIS
 * replays the chosen light's bit trail from the stored query, and the environment takes 1/2
 * (pt_env.select_prob). */
enum { PT_SEL_POWER = 0, PT_SEL_UNIFORM = 1, PT_SEL_BVH = 2 };

/* First-hit guides (src/components/aov/): at the first non-null surface a path reaches with no
 * scattering event before it, the material's albedo fact clamped to [0, 1] times the path's
 * weight on arrival ("albedo"), and the shading normal faced toward the arriving ray times the
 * weight's channel mean ("normal"); 0 where no surface is reached. Each is averaged like the
 * radiance; pt_read_export(r, "albedo" | "normal", ...) reads them. */
enum { PT_AOV_ALBEDO = 1, PT_AOV_NORMAL = 2 };

/* Live per-frame values (the engine builtins + camera uniforms,
 * src/engine/Engine.ts:231-254, src/components/camera/basis.ts:50-56). */
typedef struct pt_params {
    float    cam_position[4];
    float    cam_forward[4];
    float    cam_right[4];
    float    cam_up[4];
    float    tan_fov;         /* tan(fov/2), host f64 -> f32 */
    float    aspect;          /* width / height */
    uint32_t reset_salt;
    int32_t  camera;          /* PT_CAM_PINHOLE, PT_CAM_ORTHOGRAPHIC or PT_CAM_THINLENS */
    int32_t  pixel_offset[2];
    float    ortho_scale;     /* orthographic: world half-height of the view (u_orthoScale) */
    float    aperture;        /* thin lens: lens radius (u_aperture), >= 0; 0 images like the pinhole */
    float    focus_distance;  /* thin lens: distance of the sharp plane along forward (u_focusDistance), > 0 */
    int32_t  projection;      /* fisheye: PT_FISHEYE_* radial map (the FISHEYE_THETA selector) */
    float    fisheye_k;       /* fisheye: the map's radial constant u_fisheyeK (fisheye.ts fisheyeK) */
    float    cyl_focal;       /* cylindrical: cylinder radius in pixels u_cylFocal = width / hfov(rad) */
} pt_params;

/* Cameras (src/components/camera/pinhole/pinhole.glsl:4-18, orthographic/orthographic.glsl:6-21,
 * thinlens/thinlens.glsl:6-33): the pinhole spreads directions over tan(fov/2); the orthographic
 * camera shares the forward direction and slides the origin over the film plane; the thin lens
 * moves the pinhole ray's origin over a disk of radius `aperture` (the concentric map of the
 * lens sample) and re-aims it at the pinhole ray's point on the focus plane. */
enum { PT_CAM_PINHOLE = 0, PT_CAM_ORTHOGRAPHIC = 1, PT_CAM_THINLENS = 2, PT_CAM_FISHEYE = 3, PT_CAM_CYLINDRICAL = 4,
       PT_CAM_EQUIRECT = 5 };
/* Fisheye radial maps (camera/fisheye/fisheye.glsl): theta = rho K, 2 asin(rho K), 2 atan(rho K),
 * asin(rho K). The cylindrical camera sweeps azimuth = arc length over a cylinder of radius
 * cyl_focal pixels; the equirect camera maps the film to azimuth x elevation. */
enum { PT_FISHEYE_EQUIDISTANT = 0, PT_FISHEYE_EQUISOLID = 1, PT_FISHEYE_STEREOGRAPHIC = 2, PT_FISHEYE_ORTHOGRAPHIC = 3 };

/* ---- rays for the traversal entry points (SoA) ---- */
typedef struct pt_rays_soa {
    const float* ox; const float* oy; const float* oz;
    const float* dx; const float* dy; const float* dz;
    const float* tmax;        /* closest: initial hit.t (MAX_DIST); any: maxDist */
} pt_rays_soa;

typedef struct pt_ctx pt_ctx;
typedef struct pt_renderer pt_renderer;

/* Block until all work queued on the context's stream has finished (the gl.finish the
 * reference's GPUProfiler emulates with a 1-pixel readPixels, src/engine/GPUProfiler.ts:1-21). */
pt_status pt_finish(pt_ctx* ctx);

/* Thread-local message describing the last failure on this thread. */
const char* pt_last_error(void);

/* Engine constructor/destructor (src/engine/Engine.ts:83-152). */
pt_status pt_ctx_create(int device_index, pt_ctx** out);
void      pt_ctx_destroy(pt_ctx* ctx);

/* Context loss (Engine._handleContextLoss, src/engine/Engine.ts:476-491): every device resource
 * the context holds is released (the resident scene and its tables, each loaded renderer's film
 * and path state, each batch's scenes, films and graph). Renderer and batch handles stay valid
 * but lost: their calls fail with PT_E_CONTEXT_LOST (pt_sample_count reads 0) and only
 * pt_renderer_unload / pt_batch_destroy remain. The caller restores as the reference's
 * _handleContextRestored does: upload the scene again and load each renderer again from its
 * retained program (a fresh, zeroed accumulation). The context itself (device, stream) stays. */
pt_status pt_ctx_reset(pt_ctx* ctx);

/* Upload and build the scene: triangle wide BVH, containment structures, env table.
 * Replaces registerDataTexture[U32]/registerEnvironmentTable (Engine.ts:438-460) plus
 * the App-side packing (app/App.ts:301-345). */
pt_status pt_scene_upload(pt_ctx* ctx, const pt_scene_desc* scene);

/* Engine.loadRenderer (Engine.ts:153-164): select the kernel instantiation for program. */
pt_status pt_renderer_load(pt_ctx* ctx, const pt_program* program, pt_renderer** out);

/* The same from the reference's own plan record: `program_json` is a ProgramDescription
 * (src/compiler/plan/types.ts:39-249) as JSON.stringify writes it; width, height and
 * frames_in_flight are the Engine state it does not carry (Engine.setImageSize). Fields this
 * backend implements are lowered to a pt_program; a field selecting a component it does not
 * implement fails with PT_E_UNSUPPORTED and pt_last_error() names the field and value, as the
 * reference's Validator names a rejected option. pt_program_from_description needs no device. */
/* The importance table's luminance mass of a tabulated environment: sum of the blurred
 * (and, when compensated, compensated) luminance times sin(theta) over the W x H cells, the
 * `totalWeight` of src/components/env/importance.ts:100-127 the reference ships as the
 * env.totalWeight parameter (the derived env selection probability reads it). 0 for
 * untabulated environments. Needs no device. */
double pt_env_total_weight(const pt_env* env);
/* The same for an image environment: the mass of the table `env` (its chart, table_h, compensation)
 * builds from the w x h picture rgb; 0 on a bad argument (see pt_last_error). */
double pt_env_image_total_weight(const pt_env* env, const float* rgb, int32_t w, int32_t h);

pt_status pt_program_from_description(const char* program_json, size_t len, int32_t width, int32_t height,
                                      int32_t frames_in_flight, pt_program* out);
pt_status pt_renderer_load_description(pt_ctx* ctx, const char* program_json, size_t len, int32_t width, int32_t height,
                                       int32_t frames_in_flight, pt_renderer** out);
void      pt_renderer_unload(pt_renderer* r);

/* Engine.renderFrame x nframes (Engine.ts:190-204): frame k traces 1 spp per pixel with
 * sampleCount = current count + k and averages into the film. Asynchronous. */
pt_status pt_render_frames(pt_renderer* r, const pt_params* params, uint32_t nframes);

/* pt_render_frames without any device allocation. It has no counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc). The path state lives in one static device
 * array of the library per GPU (about 70 MiB: PT_NOALLOC_MAX_PATHS slots, held until the
 * process ends) instead of the renderer's own buffers; the film is the renderer's, allocated
 * at load. The film may have at most PT_NOALLOC_MAX_PATHS pixels (PT_E_INVALID otherwise),
 * and frames in flight shrink so that width * height * frames fits; the film equals
 * pt_render_frames' bit for bit. Every call on a GPU uses that GPU's one array, which keeps
 * the last call's contents: calls that overlap on one GPU (two contexts' streams) race.
 * pt_last_ray_counts reports pt_render_frames calls only. Asynchronous. */
#define PT_NOALLOC_MAX_PATHS 262144
pt_status pt_render_frames_noalloc(pt_renderer* r, const pt_params* params, uint32_t nframes);

/* Engine.readExport('radiance') (Engine.ts:351-398): RGBA32F film, row 0 = bottom
 * (gl_FragCoord convention), width*height*4 floats. Synchronises the stream. "variance"
 * reads the Welford statistic picture (per-channel population variance of the samples) of a
 * program compiled with variance != 0. */
pt_status pt_read_export(pt_renderer* r, const char* name, float* out, size_t out_floats);

/* Engine.clearAccumulation / pinResetSalt (Engine.ts:400-436). */
pt_status pt_clear_accumulation(pt_renderer* r);
uint32_t  pt_sample_count(const pt_renderer* r);

/* Parity entry points: closest-hit (t, global primitive id; -1 = miss) and any-hit over
 * the uploaded scene, for n rays in host memory. Primitive ids: triangles 0..n_tris-1,
 * analytic object j at n_tris + j. Equal-t ties resolve to the smaller id. */
pt_status pt_trace_closest(pt_ctx* ctx, const pt_rays_soa* rays, uint32_t n, float* t_out, int32_t* prim_out);
pt_status pt_trace_any(pt_ctx* ctx, const pt_rays_soa* rays, uint32_t n, uint8_t* hit_out);

/* Device-resident variants for timing: pointers are device memory already holding the
 * ray SoA (7 arrays of n floats, packed back to back) and receiving the results. */
pt_status pt_trace_closest_device(pt_ctx* ctx, const float* d_rays, uint32_t n, float* d_t, int32_t* d_prim);
pt_status pt_trace_any_device(pt_ctx* ctx, const float* d_rays, uint32_t n, uint8_t* d_hit);

/* Kernel timing of the last pt_render_frames / trace call, measured with HIP events on
 * the context stream: fills up to n_slots (name, total ms, launches) entries. */
typedef struct pt_kernel_timing { char name[32]; float total_ms; uint32_t launches; uint64_t items; } pt_kernel_timing;
pt_status pt_set_timing(pt_ctx* ctx, int enabled);
int       pt_get_timings(pt_ctx* ctx, pt_kernel_timing* out, int n_slots);

/* Statistics of the last render: rays traced by kind (closest, any). */
pt_status pt_last_ray_counts(pt_renderer* r, uint64_t* closest, uint64_t* any);

/* Live material values (the reference's setParameter for material uniforms,
 * src/engine/Engine.ts:206-236): replace the n material records of the uploaded scene
 * without rebuilding geometry. model, has_medium and emissive must not change. */
pt_status pt_scene_update_materials(pt_ctx* ctx, const pt_material* materials, int32_t n);

/* Live light values (setParameter on driven emission, src/engine/Engine.ts:258, with the
 * CPU-recomputed u_light_selpdf / u_light_cdf arrays of src/compiler/generate/features/
 * lighting.ts:166-183): replace the n light records of the uploaded scene. Only the emission
 * (radiance, s_radiance) and the selection table (select_pdf, cdf) may change: kind, region and
 * the geometry rows are constant, as the reference's contract holds them (resolveLightValues,
 * lighting.ts:337-348). Emission must be finite and >= 0, select_pdf in (0, 1], and cdf must
 * rise to 1 (within 1e-5) across the records. A hittable light's emission also lives in its
 * quad's material record (pt_scene_update_materials). Renderers read the new records from their
 * next call. */
pt_status pt_scene_update_lights(pt_ctx* ctx, const pt_light* lights, int32_t n);

/* Live environment values (setParameter on env.intensity, env.rotation, env.selectProb and a
 * constant sky's colour; src/compiler/generate/features/environment.ts:119-160, lighting.ts
 * envSelectionDeps :520-531): type, samplable, chart, compensated and the table size must
 * equal the resident environment's (the table is rebuilt by pt_env_table_upload or a scene
 * upload); intensity, rotation, select_prob, color and s_color take effect from the next call.
 * The procedural sky's table is baked at rotation 0, so rotation never invalidates it. */
pt_status pt_scene_update_env(pt_ctx* ctx, const pt_env* env);

/* Engine.registerEnvironmentTable (src/engine/Engine.ts:450-460, build-environment-sampler.ts
 * :20-43): the importance table of the resident tabulated environment (PT_ENV_SKY) rebuilt from
 * the caller's W x H linear RGB picture on the chart (rgb: W*H*3 floats, texel (i, j) at chart
 * coordinates ((i + 0.5) / W, (j + 0.5) / H); equirect row 0 is the +Y pole) by
 * buildEnvImportanceTable (src/components/env/importance.ts:16-127): luminance, 3 x 3 tent blur,
 * optional MIS compensation, sin(theta)-weighted (equirect) or equal-area (octahedral, W == H)
 * conditional and marginal CDFs, on the device. It replaces the table the scene upload baked;
 * chart and compensated become the resident table's. Every sample must be finite and >= 0
 * (PT_E_INVALID otherwise, as the reference throws); 1 <= W <= 4096, H >= 1. total_weight
 * (may be NULL) receives the table's mass, the reference's returned totalWeight. Synchronous. */
pt_status pt_env_table_upload(pt_ctx* ctx, const float* rgb, int32_t width, int32_t height, int32_t chart,
                              int32_t compensated, double* total_weight);

/* ---- differentiation: path-replay adjoint (RGB programs) ----
 * Parameters: the homogeneous medium of material `medium_material` (sigma_s RGB, sigma_a RGB,
 * HG g) and the grey Lambert albedos of up to 16 materials. Gradient layout:
 * [0..2] d/d sigma_s, [3..5] d/d sigma_a, [6] d/d g, [7 + k] d/d albedo of albedo_material[k].
 * The backward pass replays the forward pass's paths (same sampleCount / resetSalt streams)
 * and accumulates, per path, the derivative of the log of the measurement contribution with
 * the sampling decisions held fixed (detached sampling): every contribution c scatters
 * sum_ch dL/dI_ch * c_ch * dlog f_ch/dtheta. Sampling densities, MIS weights and Russian
 * roulette are estimator choices and carry no derivative. A channel with sigma_s = 0 scatters
 * nothing, so its d/dsigma_s holds the extinction term only. */
enum { PT_REDUCE_ATOMIC = 0, PT_REDUCE_DETERMINISTIC = 1 };
enum { PT_ADJ_MAX_ALBEDO = 16, PT_ADJ_N_PARAMS = 7 + PT_ADJ_MAX_ALBEDO };
typedef struct pt_adjoint_params {
    int32_t medium_material;                       /* -1: no medium parameters */
    int32_t n_albedo;                              /* 0..16 */
    int32_t albedo_material[PT_ADJ_MAX_ALBEDO];    /* Lambert materials with grey albedo */
    int32_t reduce_mode;                           /* PT_REDUCE_* */
    int32_t pad;
} pt_adjoint_params;

/* ---- batched persistent wavefront (many scenes in one launch) ----
 * No counterpart in the reference (it renders one compiled scene per Engine): the batch
 * holds n scenes of uneven size, each with its own compiled program (same estimator, bounce
 * budget, RR and RGB color model), camera and film, and renders all of them as one sequence
 * of waves whose path slots mix scenes, captured into one hipGraph. Per-pixel output equals
 * pt_render_frames on each scene alone (same sample streams and arithmetic). */
typedef struct pt_batch pt_batch;
typedef struct pt_batch_config {
    int32_t wave_paths;          /* path slots per wave (queue capacity), >= the largest film */
    int32_t trace_block;         /* traversal block size: 64, 128 or 256 */
    int32_t trace_blocks_per_cu; /* persistent traversal / shadow grid */
    int32_t shade_blocks_per_cu; /* bin / shade grid */
    int32_t stack;               /* LDS stack entries per lane: 16 or 32, >= every scene's need */
    int32_t use_graph;           /* capture the whole batch into one hipGraph */
    int32_t pad[2];
} pt_batch_config;
pt_status pt_batch_create(pt_ctx* ctx, int32_t n_scenes, const pt_scene_desc* scenes, const pt_program* programs,
                          const pt_params* params, pt_batch** out);
void      pt_batch_destroy(pt_batch* b);
pt_status pt_batch_set_config(pt_batch* b, const pt_batch_config* cfg);
pt_status pt_batch_get_config(const pt_batch* b, pt_batch_config* cfg);
/* Clear every film and render spp[s] frames of scene s from sampleCount 0. Asynchronous. */
pt_status pt_batch_render(pt_batch* b, const uint32_t* spp);
pt_status pt_batch_read_export(pt_batch* b, int32_t scene, float* out, size_t out_floats);
/* Autotuner: times the batch at spp[s] / probe_div (at least 1), first over wave sizes of 2^23
 * to 2^28 path slots at the current launch shape, then over traversal block size, blocks per CU
 * and stack depth at the best wave size; sets and returns the fastest. The sweep (configs, ms)
 * is written to the first max_results entries of the optional arrays. */
pt_status pt_batch_autotune(pt_batch* b, const uint32_t* spp, uint32_t probe_div, pt_batch_config* best,
                            pt_batch_config* results, float* ms, int32_t max_results, int32_t* n_results);

/* Forward pass: clear the film and render nframes from sampleCount 0 (the primal image of
 * the adjoint pair; read it with pt_read_export). */
pt_status pt_forward(pt_renderer* r, const pt_params* params, uint32_t nframes);
/* Backward pass: replay the same nframes and write dL/dparams (PT_ADJ_N_PARAMS floats) for
 * the adjoint image dL_dimage (width*height*4 floats, host memory, row 0 = bottom; the film
 * average's 1/nframes is applied inside). PT_REDUCE_DETERMINISTIC sums per-pixel partials in
 * a fixed order (bitwise reproducible); PT_REDUCE_ATOMIC scatters block partials with fp32
 * atomics. Synchronises. */
pt_status pt_backward(pt_renderer* r, const pt_params* params, uint32_t nframes, const pt_adjoint_params* adj,
                      const float* dL_dimage, float* dparams);

/* ---------------------------------------------------------------- on-disk inputs (host only)
 * Readers for the two file formats the reference's authoring layer loads, so a scene can be
 * assembled without the TypeScript App. Neither needs a device or a context. */

/* Wavefront OBJ text -> indexed mesh (authoring/loadOBJ.ts:62-163, parseOBJ): v / vt / vn / f
 * records; polygons fan-triangulated; 1-based and negative (relative) indices resolved
 * against the records read so far; every distinct resolved (v, vt, vn) corner is one output
 * vertex. Normals and uvs are present only when the file has them. PT_OBJ_SMOOTH_NORMALS is
 * MeshAuthoring.smoothNormals: area-weighted vertex normals for a file without normals
 * (authoring/loadOBJ.ts:35-60); corners the file leaves without a normal in a file that has
 * normals always get them. A file without faces fails with PT_E_INVALID ("parseOBJ: no faces
 * found ..."). The result is a library-owned pt_mesh_data: query its sizes with
 * pt_mesh_data_info, copy it out with pt_mesh_data_copy (any output may be NULL), free it with
 * pt_mesh_data_free. */
enum { PT_OBJ_SMOOTH_NORMALS = 1 };
typedef struct pt_mesh_data pt_mesh_data;
pt_status pt_obj_parse(const char* text, size_t len, uint32_t flags, pt_mesh_data** out);
pt_status pt_mesh_data_info(const pt_mesh_data* m, uint32_t* n_vertices, uint32_t* n_indices, int32_t* has_normals,
                            int32_t* has_uvs);
pt_status pt_mesh_data_copy(const pt_mesh_data* m, float* positions, uint32_t* indices, float* normals, float* uvs);
void      pt_mesh_data_free(pt_mesh_data* m);

/* A parsed `.inst` instance cloud (authoring/loadInstances.ts:9-22, InstanceTable; layout
 * tools/inst-format.mjs:8-19, version 1, little-endian). Offsets are byte offsets into the
 * caller's buffer of 4-byte-aligned f32 blocks: positions 3N, sizes N, colors 3N (linear RGB),
 * orientations 4N (unit quaternions x, y, z, w), then n_scalars columns of N each; 0 = block
 * absent (flags bit0 sizes, bit1 colors, bit2 orientations). Column c's name is the
 * zero-padded 32 bytes at names_offset + 32 c; the provenance is provenance_len UTF-8 bytes.
 * aabb is the header's positions-only box (min xyz, max xyz). */
typedef struct pt_inst_table {
    uint32_t count, flags, n_scalars, pad;
    float    aabb[6];
    uint64_t names_offset, provenance_offset, provenance_len;
    uint64_t positions_offset, sizes_offset, colors_offset, orientations_offset, scalars_offset;
} pt_inst_table;
/* parseInstances (authoring/loadInstances.ts:30-98): validates the header (magic, version,
 * nonzero count, no unknown flag bits, finite non-empty box, column names present and
 * distinct, byte length equal to the header arithmetic) and fills `out`. Failures are
 * PT_E_INVALID with the reference's "inst: ..." message. Nothing is copied. */
pt_status pt_inst_parse(const void* buffer, size_t len, pt_inst_table* out);

/* ---------------------------------------------------------------- scene-data rail
 * The App's data packing run natively (app/sceneData.ts over components/accel/bvh/bvh.ts and
 * components/data/ledger.ts): the reference's BVH builder byte for byte, its data ledger and
 * the mesh channels written at the ledger's bases. Host-only: no device and no context. */

/* buildBVHNodesFlat (components/accel/bvh/bvh.ts:323-360 over buildBVHCore :116-320): the binned
 * SAH binary BVH over n boxes (6 f64 each: min.xyz, max.xyz), leaves of <= leaf_size items,
 * byte for byte (nodes as 8 f32: min.xyz, A, max.xyz, B; A >= 0 leaf of A items at offset B of
 * order, A < 0 internal with axis -A-1, right child B, left child next). PT_BVH_ALWAYS_SPLIT
 * splits above leaf_size even where SAH would stop; PT_BVH_SORTED_MEDIAN takes buildBVHCore's
 * sorted-median fallback instead of the median. nodes_out holds 8 (2n - 1) floats, order_out n. */
enum { PT_BVH_ALWAYS_SPLIT = 1, PT_BVH_SORTED_MEDIAN = 2 };
pt_status pt_bvh_build(const double* boxes, int32_t n, int32_t leaf_size, uint32_t flags, float* nodes_out,
                       uint32_t* order_out, int32_t* node_count, int32_t* max_depth);
/* buildLightTree (src/components/accel/light_tree/light_tree.ts:46-77): the light tree over n
 * lights (boxes: 6 f64 each, min.xyz, max.xyz, degenerate allowed; powers: n f64, finite and
 * > 0), the binned-SAH core weighted by power, one light per leaf, always split, depth <= 48
 * (the balanced sorted-median split where the cap bites). nodes_out: 8 (2n - 1) floats, node k
 * = (min.xyz, phi) (max.xyz, link), phi the subtree power, link >= 0 the right child (the left
 * is k + 1), link < 0 a leaf of light -link - 1; trails_out: 4 n floats, per light (lo, hi, 0, 0),
 * its root-to-leaf left/right bits LSB first, levels 0-23 in lo and 24-47 in hi. */
pt_status pt_light_tree_build(const double* boxes, const double* powers, int32_t n, float* nodes_out, float* trails_out,
                              int32_t* node_count, int32_t* max_depth);
/* buildBVH (bvh.ts:444-470): a mesh BLAS over n_tris triangles (3 indices each into n_vertices
 * positions of 3 f32), leaf size 2; nodes as pt_bvh_build, reindexed_out the 3 n_tris indices
 * in leaf order. */
pt_status pt_mesh_bvh_build(const float* positions, int32_t n_vertices, const uint32_t* indices, int32_t n_tris,
                            float* nodes_out, uint32_t* reindexed_out, int32_t* node_count, int32_t* max_depth);

/* DataTenants and DataLayout (components/data/ledger.ts:47-114). A region base is a texel index
 * in its channel; -1 where the structure is absent. */
typedef struct pt_mesh_tenant { int32_t vertex_count, tri_count; } pt_mesh_tenant;
typedef struct pt_batch_tenant {
    int32_t instance_count, placement_texels;   /* placement_texels: 2 rigid frame, 1 folded params */
    int32_t attr_texels, cwbvh_node_texels, cwbvh_record_texels;
} pt_batch_tenant;
typedef struct pt_mesh_light_tenant { int32_t mesh_ordinal, vertex_count, tri_count; } pt_mesh_light_tenant;
typedef struct pt_data_tenants {
    int32_t n_meshes; const pt_mesh_tenant* meshes;
    int32_t n_batches; const pt_batch_tenant* batches;
    int32_t n_mesh_lights; const pt_mesh_light_tenant* mesh_lights;
    int32_t has_scene_table, table_leaf_count, table_analytic_texels;
    int32_t has_light_tree, light_count, light_table_texels;
    int32_t has_region_materials, region_count;
} pt_data_tenants;
typedef struct pt_mesh_slot { int32_t vbase, tbase, nbase; } pt_mesh_slot;
typedef struct pt_batch_slot { int32_t placements_base, tlas_base, attrs_base, cwbvh_nodes_base, cwbvh_records_base; } pt_batch_slot;
typedef struct pt_mesh_light_slot { int32_t mesh_ordinal, wpos_base, cdf_base; } pt_mesh_light_slot;
typedef struct pt_data_layout {
    int32_t total_vertices, total_normals, total_uvs, total_indices, total_nodes, total_records, total_nodesq;
    pt_mesh_slot* meshes;             /* caller arrays the planner fills: n_meshes, */
    pt_batch_slot* batches;           /* n_batches and */
    pt_mesh_light_slot* mesh_lights;  /* n_mesh_lights entries (tenant order) */
    int32_t table_leaf_list_base, table_analytic_base, table_tlas_base;
    int32_t light_table_base, light_trails_base, light_tree_base;
    int32_t region_materials_base;
} pt_data_layout;
/* planDataLayout (ledger.ts:122-184): every region from counts alone, always-built regions
 * before optional ones. */
pt_status pt_plan_data_layout(const pt_data_tenants* tenants, pt_data_layout* layout);

/* One mesh of the packing (MeshObject): positions (3 f32 per vertex), triangle indices, and
 * per-vertex normals (3 f32) / uvs (2 f32) or NULL. */
typedef struct pt_mesh_view {
    const float* positions; int32_t n_vertices;
    const uint32_t* indices; int32_t n_tris;
    const float* normals;
    const float* uvs;
} pt_mesh_view;
/* packSceneData's geometry pass (app/sceneData.ts:36-47, components/data/pack.ts): for mesh slot
 * s, its positions at layout.meshes[s].vbase of the vertices channel (xyz, w 0), normals and uvs
 * likewise when given, its BLAS's reindexed triangles at tbase of the indices channel (exact
 * floats, w 0) and its nodes at nbase of the nodes channel. Channels are caller RGBA32F arrays
 * of 4 x the layout's totals (zero-filled by the caller); a mesh shared by several slots builds
 * once; a BLAS that outgrows its declared region fails with PT_E_INVALID (assertFits). root_boxes
 * (6 f32 per mesh, or NULL) receive each BLAS's root box (the prototype box of an instance batch). */
pt_status pt_pack_mesh_channels(const pt_data_layout* layout, int32_t n_meshes, const pt_mesh_view* meshes,
                                float* vertices, float* normals, float* uvs, float* indices, float* nodes,
                                float* root_boxes);

#ifdef __cplusplus
}
#endif

#endif /* PTMI_H */
