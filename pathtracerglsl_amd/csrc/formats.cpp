// formats.cpp — on-disk scene inputs, host C++.
//
// 1. Wavefront OBJ → indexed mesh (authoring/loadOBJ.ts:62-163): v / vt / vn / f records,
//    fan-triangulated polygons, 1-based and negative (relative) indices resolved against the
//    records read so far, each distinct resolved (v, vt, vn) corner one output vertex, and
//    area-weighted vertex normals where the file has none (authoring/loadOBJ.ts:35-60).
// 2. `.inst` instance clouds (authoring/loadInstances.ts:30-98; the byte layout of
//    tools/inst-format.mjs:8-19): header validation, then byte offsets of each f32 block in
//    the caller's buffer (the reference returns Float32Array views over it; nothing is copied).
//
// The reference runs in JavaScript, so its numbers are IEEE doubles rounded into Float32Array
// storage at the points it stores them; this file does the same (double arithmetic, a float
// store wherever the reference writes a Float32Array element). Error texts are the reference's.
#include "formats.h"

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>

namespace {

bool is_space(char c) { return c == ' ' || c == '\t' || c == '\r' || c == '\v' || c == '\f' || c == '\n'; }

// JavaScript parseFloat: the longest prefix that is a StrDecimalLiteral (optional sign, then
// "Infinity", or digits with an optional fraction and exponent), NaN when there is none. The
// prefix is converted by strtod, which rounds correctly like the ECMAScript conversion.
double js_parse_float(const std::string& s) {
    size_t i = 0, n = s.size();
    while (i < n && is_space(s[i])) i++;
    size_t start = i;
    if (i < n && (s[i] == '+' || s[i] == '-')) i++;
    if (s.compare(i, 8, "Infinity") == 0) return s[start] == '-' ? -INFINITY : INFINITY;
    size_t digits = 0;
    while (i < n && s[i] >= '0' && s[i] <= '9') { i++; digits++; }
    if (i < n && s[i] == '.') {
        size_t j = i + 1, frac = 0;
        while (j < n && s[j] >= '0' && s[j] <= '9') { j++; frac++; }
        if (digits + frac > 0) { i = j; digits += frac; }
    }
    if (digits == 0) return NAN;
    if (i < n && (s[i] == 'e' || s[i] == 'E')) {
        size_t j = i + 1;
        if (j < n && (s[j] == '+' || s[j] == '-')) j++;
        size_t ed = 0;
        while (j < n && s[j] >= '0' && s[j] <= '9') { j++; ed++; }
        if (ed > 0) i = j;
    }
    std::string lit = s.substr(start, i - start);
    return std::strtod(lit.c_str(), nullptr);
}

// JavaScript parseInt(s, 10): optional sign and the leading decimal digits, NaN when there
// are none. Digit strings beyond 2^53 keep their double value, as in JavaScript.
double js_parse_int(const std::string& s) {
    size_t i = 0, n = s.size();
    while (i < n && is_space(s[i])) i++;
    bool neg = false;
    if (i < n && (s[i] == '+' || s[i] == '-')) { neg = s[i] == '-'; i++; }
    size_t d0 = i;
    while (i < n && s[i] >= '0' && s[i] <= '9') i++;
    if (i == d0) return NAN;
    double v = std::strtod(s.substr(d0, i - d0).c_str(), nullptr);
    return neg ? -v : v;
}

// Math.hypot as V8 computes it (src/builtins/math.tq, MathHypot): scale by the largest
// magnitude, Kahan-sum the squares, sqrt, rescale.
double js_hypot(double a, double b, double c) {
    double v[3] = {std::fabs(a), std::fabs(b), std::fabs(c)};
    bool nan = false;
    double mx = 0.0;
    for (double x : v) {
        if (std::isnan(x)) nan = true;
        else if (x > mx) mx = x;
    }
    if (mx == INFINITY) return INFINITY;
    if (nan) return NAN;
    if (mx == 0.0) return 0.0;
    double sum = 0.0, comp = 0.0;
    for (double x : v) {
        double q = x / mx;
        double summand = q * q - comp;
        double pre = sum + summand;
        comp = (pre - sum) - summand;
        sum = pre;
    }
    return std::sqrt(sum) * mx;
}

// A resolved OBJ index as JavaScript holds it: a double, NaN for an unparseable token. Keys
// compare NaN equal to NaN (the reference keys on the decimal string "NaN").
struct Key {
    double v, t, n;
    static double canon(double x) { return std::isnan(x) ? NAN : x; }
    static bool less(double a, double b) {   // NaN after every number, equal to itself
        if (std::isnan(a) || std::isnan(b)) return !std::isnan(a) && std::isnan(b);
        return a < b;
    }
    bool operator<(const Key& o) const {
        if (less(v, o.v) || less(o.v, v)) return less(v, o.v);
        if (less(t, o.t) || less(o.t, t)) return less(t, o.t);
        return less(n, o.n);
    }
};

// stream[i * per + k] as JavaScript reads a plain array: undefined (NaN once stored in a
// Float32Array) unless i is an integer index inside the stream.
double at(const std::vector<double>& stream, double i, int per, int k) {
    if (!(i >= 0.0) || i != std::floor(i) || i * per + k >= (double)stream.size()) return NAN;
    return stream[(size_t)i * per + k];
}

// authoring/loadOBJ.ts:35-60, over the Float32Array positions of the output mesh.
std::vector<float> vertex_normals(const std::vector<float>& P, const std::vector<uint32_t>& I) {
    const size_t nv = P.size() / 3;
    std::vector<float> n(P.size(), 0.0f), largest(P.size(), 0.0f), largest_len(nv, 0.0f);
    for (size_t t = 0; t + 2 < I.size(); t += 3) {
        size_t ia = 3 * (size_t)I[t], ib = 3 * (size_t)I[t + 1], ic = 3 * (size_t)I[t + 2];
        double e1x = (double)P[ib] - P[ia], e1y = (double)P[ib + 1] - P[ia + 1], e1z = (double)P[ib + 2] - P[ia + 2];
        double e2x = (double)P[ic] - P[ia], e2y = (double)P[ic + 1] - P[ia + 1], e2z = (double)P[ic + 2] - P[ia + 2];
        double cx = e1y * e2z - e1z * e2y, cy = e1z * e2x - e1x * e2z, cz = e1x * e2y - e1y * e2x;
        double cl = js_hypot(cx, cy, cz);
        for (size_t i : {ia, ib, ic}) {
            n[i] = (float)((double)n[i] + cx);
            n[i + 1] = (float)((double)n[i + 1] + cy);
            n[i + 2] = (float)((double)n[i + 2] + cz);
            if (cl > (double)largest_len[i / 3]) {
                largest_len[i / 3] = (float)cl;
                largest[i] = (float)cx; largest[i + 1] = (float)cy; largest[i + 2] = (float)cz;
            }
        }
    }
    for (size_t i = 0; i < n.size(); i += 3) {
        double l = js_hypot(n[i], n[i + 1], n[i + 2]);
        if (l == 0.0) {
            n[i] = largest[i]; n[i + 1] = largest[i + 1]; n[i + 2] = largest[i + 2];
            l = largest_len[i / 3];
        }
        if (l == 0.0) { n[i] = 0.0f; n[i + 1] = 0.0f; n[i + 2] = 1.0f; continue; }
        n[i] = (float)((double)n[i] / l);
        n[i + 1] = (float)((double)n[i + 1] / l);
        n[i + 2] = (float)((double)n[i + 2] / l);
    }
    return n;
}

void split_ws(const char* s, const char* e, std::vector<std::string>& parts) {
    parts.clear();
    while (s < e) {
        while (s < e && is_space(*s)) s++;
        const char* b = s;
        while (s < e && !is_space(*s)) s++;
        if (s > b) parts.emplace_back(b, s);
    }
}

uint32_t rd32(const uint8_t* p) { uint32_t v; std::memcpy(&v, p, 4); return v; }

}  // namespace

pt_status ptmi_parse_obj(const char* text, size_t len, uint32_t flags, pt_mesh_data& out, std::string& err) {
    std::vector<double> v, vt, vn;
    std::vector<double> pos, nrm, uv;     // output streams before the Float32Array store
    std::vector<uint32_t> idx;
    std::vector<bool> corner_has_normal;
    std::map<Key, uint32_t> corners;
    bool used_normals = false, used_uvs = false;
    auto resolve = [](double raw, double count) { return raw < 0 ? count + raw : raw - 1; };
    auto corner = [&](const std::string& tok) -> uint32_t {
        // token.split('/'): absent or empty vt / vn fields select nothing
        size_t s1 = tok.find('/');
        std::string vs = tok.substr(0, s1), vts, vns;
        bool has_t = false, has_n = false;
        if (s1 != std::string::npos) {
            size_t s2 = tok.find('/', s1 + 1);
            vts = tok.substr(s1 + 1, s2 == std::string::npos ? std::string::npos : s2 - s1 - 1);
            has_t = !vts.empty();
            if (s2 != std::string::npos) {
                size_t s3 = tok.find('/', s2 + 1);
                vns = tok.substr(s2 + 1, s3 == std::string::npos ? std::string::npos : s3 - s2 - 1);
                has_n = !vns.empty();
            }
        }
        double vi = resolve(js_parse_int(vs), (double)(v.size() / 3));
        double ti = has_t ? resolve(js_parse_int(vts), (double)(vt.size() / 2)) : -1.0;
        double ni = has_n ? resolve(js_parse_int(vns), (double)(vn.size() / 3)) : -1.0;
        Key key{Key::canon(vi), Key::canon(ti), Key::canon(ni)};
        auto it = corners.find(key);
        if (it != corners.end()) return it->second;
        uint32_t o = (uint32_t)(pos.size() / 3);
        for (int k = 0; k < 3; k++) pos.push_back(at(v, vi, 3, k));
        if (ti >= 0) { uv.push_back(at(vt, ti, 2, 0)); uv.push_back(at(vt, ti, 2, 1)); used_uvs = true; }
        else { uv.push_back(0.0); uv.push_back(0.0); }
        if (ni >= 0) { for (int k = 0; k < 3; k++) nrm.push_back(at(vn, ni, 3, k)); used_normals = true; }
        else { for (int k = 0; k < 3; k++) nrm.push_back(0.0); }
        corner_has_normal.push_back(ni >= 0);
        corners.emplace(key, o);
        return o;
    };
    std::vector<std::string> parts;
    size_t p = 0;
    if (len >= 3 && (uint8_t)text[0] == 0xEF && (uint8_t)text[1] == 0xBB && (uint8_t)text[2] == 0xBF) p = 3;   // U+FEFF
    for (bool last = false; !last;) {                 // text.split('\n')
        const char* nl = (const char*)std::memchr(text + p, '\n', len - p);
        size_t le = nl ? (size_t)(nl - text) : len;
        last = nl == nullptr;
        split_ws(text + p, text + le, parts);
        p = le + 1;
        if (parts.empty() || parts[0][0] == '#') continue;
        const std::string& tag = parts[0];
        auto num = [&](size_t k) { return k < parts.size() ? js_parse_float(parts[k]) : NAN; };
        if (tag == "v") {
            v.push_back(num(1)); v.push_back(num(2)); v.push_back(num(3));
        } else if (tag == "vt") {
            vt.push_back(num(1)); vt.push_back(parts.size() > 2 ? js_parse_float(parts[2]) : 0.0);
        } else if (tag == "vn") {
            vn.push_back(num(1)); vn.push_back(num(2)); vn.push_back(num(3));
        } else if (tag == "f") {
            std::vector<uint32_t> c;
            for (size_t k = 1; k < parts.size(); k++) c.push_back(corner(parts[k]));
            for (size_t i = 1; i + 1 < c.size(); i++) { idx.push_back(c[0]); idx.push_back(c[i]); idx.push_back(c[i + 1]); }
        }
    }
    if (pos.empty() || idx.empty()) { err = "parseOBJ: no faces found (empty or unsupported OBJ)"; return PT_E_INVALID; }
    out = pt_mesh_data();
    out.positions.assign(pos.begin(), pos.end());
    out.indices = std::move(idx);
    if (used_normals) {
        out.normals.assign(nrm.begin(), nrm.end());
        bool missing = false;
        for (bool h : corner_has_normal) missing = missing || !h;
        if (missing) {
            std::vector<float> syn = vertex_normals(out.positions, out.indices);
            for (size_t i = 0; i < corner_has_normal.size(); i++)
                if (!corner_has_normal[i]) for (int k = 0; k < 3; k++) out.normals[3 * i + k] = syn[3 * i + k];
        }
        out.has_normals = true;
    } else if (flags & PT_OBJ_SMOOTH_NORMALS) {
        out.normals = vertex_normals(out.positions, out.indices);
        out.has_normals = true;
    }
    if (used_uvs) { out.uvs.assign(uv.begin(), uv.end()); out.has_uvs = true; }
    return PT_OK;
}

pt_status ptmi_parse_inst(const uint8_t* buf, size_t len, pt_inst_table& out, std::string& err) {
    auto bad = [&](const std::string& m) { err = "inst: " + m; return PT_E_INVALID; };
    if (len < 48) return bad("buffer too small for a header (" + std::to_string(len) + " bytes)");
    if (rd32(buf) != 0x54534e49u) return bad("bad magic — not a '.inst' file");
    uint32_t version = rd32(buf + 4);
    if (version != 1) return bad("unsupported version " + std::to_string(version) + " (loader speaks 1)");
    uint32_t count = rd32(buf + 8);
    if (count == 0) return bad("count is 0 — nothing to place");
    uint32_t flags = rd32(buf + 12);
    if (flags & ~7u) {
        char hex[16];
        std::snprintf(hex, sizeof hex, "%x", flags & ~7u);
        return bad(std::string("unknown flag bits 0x") + hex + " — file written by a newer format version");
    }
    float aabb[6];
    std::memcpy(aabb, buf + 16, sizeof aabb);
    bool ok = true;
    for (int a = 0; a < 6; a++) ok = ok && std::isfinite(aabb[a]);
    for (int a = 0; a < 3; a++) ok = ok && !(aabb[a] > aabb[3 + a]);
    if (!ok) return bad("header AABB is non-finite or empty");
    uint32_t k = rd32(buf + 40);
    uint64_t names_end = 44 + (uint64_t)k * 32;
    if ((uint64_t)len < names_end + 4) return bad("buffer truncated in the column-name table");
    for (uint32_t c = 0; c < k; c++) {
        const uint8_t* raw = buf + 44 + 32 * (size_t)c;
        if (raw[0] == 0) return bad("scalar column " + std::to_string(c) + " has an empty name");
        size_t lc = strnlen((const char*)raw, 32);
        for (uint32_t d = 0; d < c; d++) {
            const uint8_t* other = buf + 44 + 32 * (size_t)d;
            if (strnlen((const char*)other, 32) == lc && std::memcmp(raw, other, lc) == 0)
                return bad("duplicate scalar column names");
        }
    }
    uint64_t prov_len = rd32(buf + names_end);
    uint64_t header = names_end + 4 + prov_len + ((4 - prov_len % 4) % 4);
    if ((uint64_t)len < header) return bad("buffer truncated in the provenance string");
    uint64_t n = count;
    uint64_t floats = 3 * n + ((flags & 1) ? n : 0) + ((flags & 2) ? 3 * n : 0) + ((flags & 4) ? 4 * n : 0) + (uint64_t)k * n;
    if ((uint64_t)len != header + floats * 4)
        return bad("byte length " + std::to_string(len) + " does not match header arithmetic (" + std::to_string(header) +
                   " header + " + std::to_string(floats) + " floats) — truncated or corrupt");
    std::memset(&out, 0, sizeof out);
    out.count = count;
    out.flags = flags;
    out.n_scalars = k;
    std::memcpy(out.aabb, aabb, sizeof aabb);
    out.names_offset = 44;
    out.provenance_offset = names_end + 4;
    out.provenance_len = prov_len;
    uint64_t off = header;
    out.positions_offset = off; off += 12 * n;
    if (flags & 1) { out.sizes_offset = off; off += 4 * n; }
    if (flags & 2) { out.colors_offset = off; off += 12 * n; }
    if (flags & 4) { out.orientations_offset = off; off += 16 * n; }
    out.scalars_offset = k ? off : 0;
    return PT_OK;
}
