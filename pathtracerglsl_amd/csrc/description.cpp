// description.cpp — lowering of the reference's ProgramDescription (the Planner's link map,
// src/compiler/plan/types.ts:39-249, serialized with JSON.stringify) to a pt_program.
//
// The reference's Generate stage turns a ProgramDescription into GLSL; here the same record
// selects the kernel instantiations a pt_program names. Fields this backend implements are
// lowered; a field whose value selects a component outside it is rejected with
// PT_E_UNSUPPORTED, naming the field and value, as the reference's Validator names a rejected
// option (src/compiler/analyze/Validator.ts). Fields that only choose between equivalent
// implementations (objectDispatch, meshTraversal, instanceAccel, regionLookup) or that belong
// to the display (view) are accepted as they are.
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "description.h"

namespace {

// ---------------------------------------------------------------- a small JSON reader
struct JVal {
    enum Kind { NUL, BOOL, NUM, STR, ARR, OBJ } kind = NUL;
    bool b = false;
    double num = 0.0;
    std::string str;
    std::vector<JVal> arr;
    std::map<std::string, JVal> obj;
    const JVal* get(const char* k) const {
        if (kind != OBJ) return nullptr;
        auto it = obj.find(k);
        return it == obj.end() ? nullptr : &it->second;
    }
};

struct JParser {
    const char* p;
    const char* end;
    std::string err;
    void ws() { while (p < end && (*p == ' ' || *p == '\n' || *p == '\r' || *p == '\t')) p++; }
    bool fail(const char* m) { if (err.empty()) err = m; return false; }
    bool lit(const char* s) {
        size_t n = std::strlen(s);
        if ((size_t)(end - p) < n || std::strncmp(p, s, n) != 0) return false;
        p += n;
        return true;
    }
    bool string(std::string& out) {
        if (p >= end || *p != '"') return fail("expected a string");
        p++;
        while (p < end && *p != '"') {
            if (*p == '\\') {
                if (++p >= end) return fail("bad escape");
                switch (*p) {
                    case 'n': out += '\n'; break;
                    case 't': out += '\t'; break;
                    case 'r': out += '\r'; break;
                    case 'b': out += '\b'; break;
                    case 'f': out += '\f'; break;
                    case 'u': {                       // keep ASCII, replace others
                        if (end - p < 5) return fail("bad \\u escape");
                        unsigned v = (unsigned)std::strtoul(std::string(p + 1, p + 5).c_str(), nullptr, 16);
                        out += v < 128 ? (char)v : '?';
                        p += 4;
                        break;
                    }
                    default: out += *p;
                }
                p++;
            } else {
                out += *p++;
            }
        }
        if (p >= end) return fail("unterminated string");
        p++;
        return true;
    }
    bool value(JVal& v, int depth) {
        if (depth > 64) return fail("nesting too deep");
        ws();
        if (p >= end) return fail("unexpected end");
        if (*p == '{') {
            v.kind = JVal::OBJ; p++; ws();
            if (p < end && *p == '}') { p++; return true; }
            for (;;) {
                ws();
                std::string k;
                if (!string(k)) return false;
                ws();
                if (p >= end || *p != ':') return fail("expected ':'");
                p++;
                if (!value(v.obj[k], depth + 1)) return false;
                ws();
                if (p < end && *p == ',') { p++; continue; }
                if (p < end && *p == '}') { p++; return true; }
                return fail("expected ',' or '}'");
            }
        }
        if (*p == '[') {
            v.kind = JVal::ARR; p++; ws();
            if (p < end && *p == ']') { p++; return true; }
            for (;;) {
                v.arr.emplace_back();
                if (!value(v.arr.back(), depth + 1)) return false;
                ws();
                if (p < end && *p == ',') { p++; continue; }
                if (p < end && *p == ']') { p++; return true; }
                return fail("expected ',' or ']'");
            }
        }
        if (*p == '"') { v.kind = JVal::STR; return string(v.str); }
        if (lit("true")) { v.kind = JVal::BOOL; v.b = true; return true; }
        if (lit("false")) { v.kind = JVal::BOOL; v.b = false; return true; }
        if (lit("null")) { v.kind = JVal::NUL; return true; }
        char* e = nullptr;
        std::string tok(p, (size_t)std::min<ptrdiff_t>(end - p, 64));
        double d = std::strtod(tok.c_str(), &e);
        if (e == tok.c_str()) return fail("unexpected character");
        v.kind = JVal::NUM; v.num = d;
        p += (e - tok.c_str());
        return true;
    }
};

struct Lower {
    std::string msg;
    bool unsupported(const std::string& field, const std::string& value) {
        msg = "ProgramDescription: " + field + " " + value + " is not supported by this backend";
        return false;
    }
    bool invalid(const std::string& m) { msg = "ProgramDescription: " + m; return false; }
};

const JVal* path(const JVal& root, std::initializer_list<const char*> keys) {
    const JVal* v = &root;
    for (const char* k : keys) { v = v->get(k); if (!v) return nullptr; }
    return v;
}
std::string quoted(const JVal* v) {
    if (!v) return "(missing)";
    if (v->kind == JVal::STR) return "'" + v->str + "'";
    if (v->kind == JVal::BOOL) return v->b ? "true" : "false";
    if (v->kind == JVal::NUM) { char b[64]; std::snprintf(b, sizeof b, "%g", v->num); return b; }
    return "(value)";
}
bool is_str(const JVal* v, const char* s) { return v && v->kind == JVal::STR && v->str == s; }
bool truthy(const JVal* v) { return v && v->kind == JVal::BOOL && v->b; }

// The procedural sky bodies the library evaluates for PT_ENV_SKY (PT_SKY_* in ptmi.h): proc-sky,
// proc-sky-rotated, sky-missed-cap and spectral-comp-empty.
const char* const kSkyBodies[] = {
    "mix(vec3(0.06, 0.08, 0.12), vec3(0.35, 0.45, 0.65), 0.5 * (dir.y + 1.0)) + vec3(18.0, 16.0, 13.0) * "
    "smoothstep(0.9994, 0.9999, dot(dir, normalize(vec3(0.4, 0.35, 0.2))))",
    "vec3(0.02) + vec3(40.0) * smoothstep(0.985, 0.99, dot(dir, normalize(vec3(0.4, 0.35, 0.2))))",
    "vec3(step(0.9, dir.y))",
    "vec3(1.0)",
};
// a GLSL source without its whitespace
std::string squash(const std::string& src) {
    std::string out;
    for (char c : src)
        if (!std::isspace((unsigned char)c)) out += c;
    return out;
}
bool known_sky_body(const std::string& src) {
    for (const char* b : kSkyBodies)
        if (squash(src) == squash(b)) return true;
    return false;
}

bool lower(const JVal& d, int32_t width, int32_t height, int32_t fif, pt_program& pg, Lower& L) {
    if (d.kind != JVal::OBJ) return L.invalid("not a JSON object");
    std::memset(&pg, 0, sizeof pg);
    pg.abi_version = PT_ABI_VERSION;
    if (width <= 0 || height <= 0) return L.invalid("film size must be positive");
    pg.width = width; pg.height = height;
    pg.frames_in_flight = fif > 0 ? fif : 1;

    // measurement (compiler/types.ts:561-640)
    const JVal* m = d.get("measurement");
    if (!m || m->kind != JVal::OBJ) return L.invalid("missing 'measurement'");
    const JVal* color = m->get("color");
    if (is_str(color, "rgb")) pg.color = PT_COLOR_RGB;
    else if (is_str(color, "spectral")) pg.color = PT_COLOR_SPECTRAL;
    else return L.unsupported("measurement.color", quoted(color));
    const JVal* sensor = m->get("sensor");
    if (sensor && !is_str(sensor, pg.color == PT_COLOR_RGB ? "ideal" : "cie1931"))
        return L.unsupported("measurement.sensor", quoted(sensor));
    const JVal* mb = m->get("maxBounces");
    if (!mb || mb->kind != JVal::NUM || mb->num < 0 || mb->num != std::floor(mb->num)) return L.invalid("measurement.maxBounces must be an integer >= 0");
    pg.max_bounces = (int32_t)mb->num;
    const JVal* mn = m->get("maxNullCrossings");
    pg.max_null_crossings = mn && mn->kind == JVal::NUM ? (int32_t)mn->num : 32;
    if (const JVal* r = m->get("response"); r && !is_str(r, "radiance")) return L.unsupported("measurement.response", quoted(r));
    if (const JVal* s = m->get("shadows"); s && !is_str(s, "opaque-dielectrics")) return L.unsupported("measurement.shadows", quoted(s));
    if (const JVal* a = m->get("ambient"); a && !is_str(a, "euclidean")) return L.unsupported("measurement.ambient", quoted(a));
    if (const JVal* sc = m->get("scattering"); sc && !is_str(sc, "full") && !is_str(sc, "ignored"))
        return L.unsupported("measurement.scattering", quoted(sc));
    // measurement.aovs: the first-hit guides (components/aov/index.ts AOVS), RGB programs only
    if (const JVal* aov = m->get("aovs"); aov && aov->kind == JVal::ARR && !aov->arr.empty()) {
        if (pg.color != PT_COLOR_RGB) return L.unsupported("measurement.aovs", "(first-hit guides of a spectral program)");
        for (const JVal& a : aov->arr) {
            if (is_str(&a, "albedo")) pg.aovs |= PT_AOV_ALBEDO;
            else if (is_str(&a, "normal")) pg.aovs |= PT_AOV_NORMAL;
            else return L.unsupported("measurement.aovs entry", quoted(&a));
        }
    }
    const JVal* cam = m->get("camera");
    // pinhole, orthographic and thin lens (the pt_params camera); the lens draws are consumed either way
    if (!cam || !(is_str(cam->get("type"), "pinhole") || is_str(cam->get("type"), "orthographic") ||
                  is_str(cam->get("type"), "thinlens") || is_str(cam->get("type"), "fisheye") ||
                  is_str(cam->get("type"), "cylindrical") || is_str(cam->get("type"), "equirect")))
        return L.unsupported("measurement.camera.type", quoted(cam ? cam->get("type") : nullptr));

    // estimator (compiler/types.ts:642-721)
    const JVal* e = d.get("estimator");
    if (!e || e->kind != JVal::OBJ) return L.invalid("missing 'estimator'");
    const JVal* lighting = e->get("lighting");
    if (!lighting || lighting->kind == JVal::NUL) pg.estimator = PT_EST_PT;
    else {
        const JVal* method = lighting->get("method");
        if (is_str(method, "nee")) pg.estimator = PT_EST_NEE;
        else if (is_str(method, "mis")) pg.estimator = PT_EST_MIS;
        else return L.unsupported("estimator.lighting.method", quoted(method));
        const JVal* sel = lighting->get("selection");
        if (sel && is_str(sel, "uniform")) pg.light_selection = PT_SEL_UNIFORM;
        else if (sel && is_str(sel, "bvh")) pg.light_selection = PT_SEL_BVH;
        else if (sel && !is_str(sel, "power")) return L.unsupported("estimator.lighting.selection", quoted(sel));
    }
    const JVal* rr = e->get("russianRoulette");
    pg.rr_max_survival = 0.95f;
    if (!rr || rr->kind == JVal::NUL) pg.rr_start = -1;
    else {
        const JVal* sd = rr->get("startDepth");
        if (!sd || sd->kind != JVal::NUM || sd->num < 0) return L.invalid("estimator.russianRoulette.startDepth must be >= 0");
        pg.rr_start = (int32_t)sd->num;
        if (const JVal* ms = rr->get("maxSurvival"); ms && ms->kind == JVal::NUM) pg.rr_max_survival = (float)ms->num;
    }
    if (const JVal* acc = e->get("accumulation")) {
        if (is_str(acc->get("type"), "average")) pg.accumulation = PT_ACC_AVERAGE;
        else if (is_str(acc->get("type"), "oneshot")) pg.accumulation = PT_ACC_ONESHOT;
        else return L.unsupported("estimator.accumulation.type", quoted(acc->get("type")));
        pg.variance = truthy(acc->get("variance")) ? 1 : 0;
        if (pg.variance && pg.accumulation == PT_ACC_ONESHOT)
            return L.unsupported("estimator.accumulation", "'oneshot' with variance (oneshot keeps no statistic)");
    }
    if (const JVal* es = e->get("envSampler")) {
        if (const JVal* ch = es->get("chart"); ch && ch->kind != JVal::NUL) {
            // the table itself is built at scene upload on pt_env.chart; the program records the chart
            // so that loading it over a scene whose table is on the other chart fails
            if (is_str(ch, "equirect")) pg.env_chart = PT_CHART_EQUIRECT + 1;
            else if (is_str(ch, "octahedral")) pg.env_chart = PT_CHART_OCTAHEDRAL + 1;
            else return L.unsupported("estimator.envSampler.chart", quoted(ch));
        }
    }
    const JVal* vs = e->get("volumeSampling");
    if (!vs || vs->kind == JVal::NUL || is_str(vs, "none") || is_str(vs, "analytic")) pg.volume_sampling = PT_VS_ANALYTIC;
    else if (is_str(vs, "delta-tracking")) pg.volume_sampling = PT_VS_DELTA_TRACKING;
    else return L.unsupported("estimator.volumeSampling", quoted(vs));
    const JVal* wl = e->get("wavelengths");
    if (!wl || wl->kind == JVal::NUL || is_str(wl, "visible")) pg.wavelengths = PT_WL_VISIBLE;
    else if (is_str(wl, "uniform")) pg.wavelengths = PT_WL_UNIFORM;
    else return L.unsupported("estimator.wavelengths", quoted(wl));
    const JVal* mls = e->get("mediumLightSampling");
    if (!mls || is_str(mls, "vertex")) pg.medium_light_sampling = PT_MLS_VERTEX;
    else if (is_str(mls, "equiangular")) pg.medium_light_sampling = PT_MLS_EQUIANGULAR;
    else return L.unsupported("estimator.mediumLightSampling", quoted(mls));

    // the components the plan links (plan/types.ts: materials, media, intersection, spectra)
    if (const JVal* mats = d.get("materials")) {
        if (const JVal* models = mats->get("models"); models && models->kind == JVal::ARR)
            for (const JVal& md : models->arr)
                if (!(is_str(&md, "lambert") || is_str(&md, "dielectric") || is_str(&md, "ggx") || is_str(&md, "none") ||
                      is_str(&md, "mirror") || is_str(&md, "rough_dielectric") || is_str(&md, "checker")))
                    return L.unsupported("material model", quoted(&md));
        // materials.twoSidedShading is read only by the light tree's horizon cull
        // (accel/light_tree/light_tree.glsl, through LightQuery.two_sided); the kernels answer
        // material_two_sided per material (the rough dielectric: sphere support, non-delta), which
        // is what the flag folds away for scenes without such a material.
        // Hit.uv readers: the checker material reads the shape charts; any other reader is an image
        // texture or an expression material
        bool checker = false;
        if (const JVal* models = mats->get("models"); models && models->kind == JVal::ARR)
            for (const JVal& md : models->arr) checker = checker || is_str(&md, "checker");
        if (truthy(mats->get("materialsReadUv")) && !checker) return L.unsupported("materials.materialsReadUv", "true (textures)");
        // materials.albedo marks the material_albedo dispatch an albedo AOV needs (plan/types.ts:246-248);
        // the albedo guide itself is lowered from measurement.aovs, so the flag adds nothing here.
        if (const JVal* cones = mats->get("emissionCones"); cones && cones->kind == JVal::ARR && !cones->arr.empty())
            return L.unsupported("materials.emissionCones", "(emitter profiles)");
    }
    if (const JVal* media = d.get("media")) {
        if (truthy(media->get("deflecting"))) return L.unsupported("media.deflecting", "true (GRIN media)");
        if (truthy(media->get("heterogeneousArms"))) return L.unsupported("media.heterogeneousArms", "true (expression media)");
        if (const JVal* models = media->get("models"); models && models->kind == JVal::ARR)
            for (const JVal& md : models->arr)
                if (!(is_str(&md, "hg") || is_str(&md, "rayleigh"))) return L.unsupported("phase model", quoted(&md));
    }
    if (const JVal* isec = d.get("intersection")) {
        if (truthy(isec->get("drivenPlacement"))) return L.unsupported("intersection.drivenPlacement", "true");
    }
    if (const JVal* kinds = d.get("spectrumKinds"); kinds && kinds->kind == JVal::ARR)
        for (const JVal& k : kinds->arr)
            if (!(is_str(&k, "blackbody") || is_str(&k, "cauchy") || is_str(&k, "sellmeier") || is_str(&k, "gaussian")))
                return L.unsupported("spectrum kind", quoted(&k));
    const JVal* env = d.get("environment");
    const JVal* et = env ? env->get("type") : nullptr;
    if (!et || is_str(et, "none")) {
    } else if (is_str(et, "procedural")) {
        const JVal* src = path(*env, {"glsl", "source"});
        if (!src || src->kind != JVal::STR || !known_sky_body(src->str)) return L.unsupported("environment", "'procedural' with this GLSL body");
    } else if (is_str(et, "constant")) {
        // color * intensity in every direction; the colour itself travels with the scene (pt_env)
    } else if (is_str(et, "image")) {
        // the equirect picture travels with the scene (pt_scene_desc.env_image, PT_ENV_IMAGE)
    } else {
        return L.unsupported("environment.type", quoted(et));
    }
    return true;
}

}  // namespace

// Lower a ProgramDescription; on failure err names the field (PT_E_UNSUPPORTED) or the defect.
pt_status ptmi_lower_description(const char* json, size_t len, int32_t width, int32_t height, int32_t frames_in_flight,
                                 pt_program* out, std::string& err) {
    JParser P{json, json + len, {}};
    JVal root;
    if (!P.value(root, 0)) { err = "ProgramDescription: malformed JSON (" + P.err + ")"; return PT_E_INVALID; }
    Lower L;
    pt_program pg;
    if (!lower(root, width, height, frames_in_flight, pg, L)) {
        err = L.msg;
        return L.msg.find("not supported") != std::string::npos ? PT_E_UNSUPPORTED : PT_E_INVALID;
    }
    *out = pg;
    return PT_OK;
}
